"""Units and constants: lengths in m, times in ns, wavelengths in nm, photon energies in eV (the reference's units,
``units.py``), so every number crosses the drop-in surface unchanged."""

C_M_S = 2.99792458e8
H_PLANCK_J_S = 6.62607015e-34
E_CHARGE_C = 1.602176634e-19
FINE_STRUCTURE = 1.0 / 137.035999084

C_VACUUM = 0.299792458  # m / ns
HC_EV_M = H_PLANCK_J_S * C_M_S / E_CHARGE_C  # h c in eV m
HC = HC_EV_M * 1e9  # h c in eV nm: wavelength <-> energy
MM = 1e-3
NM_PER_M = 1e9

# Geant4's Rayleigh rule for a material named Water (G4OpRayleigh, as the reference restates it at
# ``rules/optics.py:100-104``): isothermal compressibility (m3 / MeV), temperature (K), Boltzmann (MeV / K)
WATER_COMPRESSIBILITY = 7.658e-23
WATER_TEMPERATURE = 283.15
K_BOLTZMANN = 8.617333e-11


def energy_of(lam_nm):
    """Return the photon energy (eV) of a wavelength (nm)."""
    return HC / lam_nm


def wavelength_of(e_ev):
    """Return the wavelength (nm) of a photon energy (eV)."""
    return HC / e_ev
