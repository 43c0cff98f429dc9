"""Host logic on the CPU: the navigator-array format, the BVH, the detectors, the packing, the C ABI's exports."""

import ctypes
import os
import re

import numpy as np
import pytest

import lumen_amd as L
from lumen_amd import _lib, pack
from lumen_amd.detector import NavArrays, bvh, example_tank, navarrays, solids, water_tank

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_abi_exports_every_declared_symbol():
    """The library loads (no GPU needed to load it) and exports every function include/lumen_hip.h declares."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("the library is not built (python -m lumen_amd.build)")
    header = open(os.path.join(ROOT, "include", "lumen_hip.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*\**(lh_\w+)\(", header, re.M))
    assert {"lh_model_create", "lh_trace", "lh_adjoint", "lh_query", "lh_last_error"} <= declared
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert set(_lib.SYMBOLS) <= declared


def test_abi_structs_match_the_header():
    """ctypes' structs have the header's sizes and order (lh_source: 4 ints, 14 floats, 2 int64; lh_controls: 10
    words; lh_geometry's pointers after `reach`: skin, touch, box_lo, box_hi)."""  # noqa: D205, D209
    assert ctypes.sizeof(_lib.LhSource) == 4 * 4 + 14 * 4 + 2 * 8
    assert ctypes.sizeof(_lib.LhControls) == 10 * 4
    assert _lib.LhControls.reach_bound.offset == 9 * 4
    assert _lib.LhGeometry.skin.offset == _lib.LhGeometry.reach.offset + 8  # the float, padding, the pointer
    g = _lib.LhGeometry
    assert (g.touch.offset, g.box_lo.offset, g.box_hi.offset) == tuple(g.skin.offset + 8 * k for k in (1, 2, 3))
    assert ctypes.sizeof(g) == g.box_hi.offset + 8
    header = open(os.path.join(ROOT, "include", "lumen_hip.h")).read()
    assert "const float *box_lo, *box_hi;\n} lh_geometry;" in header
    assert "int32_t reach_bound;" in header.split("} lh_controls;")[0].split("typedef struct {")[-1]


def _leaf(T, li, c=None, h_shift=0.0):
    """Leaf ``li`` of the CSG table T as a solids.Leaf (optionally with another c, and its planes moved out)."""
    pl = slice(int(T["p0"][li]), int(T["p0"][li] + T["np_"][li]))
    return solids.Leaf(int(T["typ"][li]), np.array([T[f"o{a}"][li] for a in range(3)], float),
                       np.array([T[f"q{a}"][li] - T[f"o{a}"][li] for a in range(3)], float),
                       np.array([[T[f"A{a}{b}"][li] for b in range(3)] for a in range(3)], float),
                       np.array([T[f"b{a}"][li] for a in range(3)], float), float(T["c"][li] if c is None else c),
                       np.stack([T["n0"], T["n1"], T["n2"]], 1)[pl].astype(float),
                       np.asarray(T["h"][pl], float) + h_shift)


def test_daughter_free_cores_hold_no_daughter():
    """Each volume's core (its skin leaf) holds no point of any daughter: points drawn in every daughter's solid
    are outside the core (the mother's frame); a core a quarter of the skin thicker does hold some (the check
    can fail), and the core holds the mother's centre."""
    from lumen_amd.detector import lar_box

    rng = np.random.default_rng(0)
    for det in (water_tank(), example_tank(), lar_box()):
        nav, h = det.nav, det.nav.host
        T = nav.csg["T"]
        assert nav.stats["skin_cores"] == 1
        v = int(np.flatnonzero(h["skin"] >= 0)[0])
        li = int(h["skin"][v])
        core = _leaf(T, li)
        # the outer solid's leaf: the core's depth D from the plane offsets of the two
        outer_li = int(nav.csg["K"]["l0"][int(h["kind"][v])])
        D = float(T["h"][T["p0"][outer_li]] - T["h"][T["p0"][li]])
        c_big = -(np.sqrt(-T["c"][li]) + 0.25 * D) ** 2 if T["typ"][li] == 1 else None
        bigger = _leaf(T, li, c=c_big, h_shift=0.25 * D)
        Rm, tm = h["R"][v].astype(float), h["t"][v].astype(float)
        assert core.inside(np.zeros((1, 3)) @ Rm - tm @ Rm)[0]
        n_pts = n_big = 0
        for k in np.flatnonzero(h["parent"][:nav.n] == v):
            x = rng.uniform(nav.lo_w[k], nav.hi_w[k], size=(2000, 3))
            Rk, tk = h["R"][k].astype(float), h["t"][k].astype(float)
            xin = x[_inside_kind(nav, int(h["kind"][k]), (x - tk) @ Rk)]
            n_pts += len(xin)
            assert not np.any(core.inside((xin - tm) @ Rm)), (det.name, int(k))
            n_big += int(np.sum(bigger.inside((xin - tm) @ Rm)))
        assert n_pts > 1000 and n_big > 0, (det.name, n_pts, n_big)


def _inside_kind(nav, kind, x):
    """Membership of points x (kind frame) in kind ``kind`` from the CSG tables."""
    T, K, G = nav.csg["T"], nav.csg["K"], nav.csg["G"]
    l0, nl, p0, pl = (int(K[k][kind]) for k in ("l0", "nl", "p0", "pl"))
    vals = []
    for li in range(l0, l0 + nl):
        lf = solids.Leaf(int(T["typ"][li]), np.array([T[f"o{a}"][li] for a in range(3)], float),
                         np.array([T[f"q{a}"][li] - T[f"o{a}"][li] for a in range(3)], float),
                         np.array([[T[f"A{a}{b}"][li] for b in range(3)] for a in range(3)], float),
                         np.array([T[f"b{a}"][li] for a in range(3)], float), float(T["c"][li]),
                         np.stack([T["n0"], T["n1"], T["n2"]], 1)[T["p0"][li]:T["p0"][li] + T["np_"][li]].astype(float),
                         np.asarray(T["h"][T["p0"][li]:T["p0"][li] + T["np_"][li]], float))
        vals.append(lf.inside(x))
    st = []
    for j in range(p0, p0 + pl):
        op, arg = int(G["op"][j]), int(G["arg"][j])
        if op == 0:
            st.append(vals[arg])
        else:
            b, a = st.pop(), st.pop()
            st.append(a | b if op == 1 else a & b if op == 2 else a & ~b)
    return st[-1]


def _check_bvh(t, lo, hi):
    n = len(t["skip"])
    for i in range(n):
        assert i < t["skip"][i] <= n
        if t["count"][i] > 0:
            assert t["skip"][i] == i + 1
            for it in t["items"][t["start"][i]: t["start"][i] + t["count"][i]]:
                assert np.all(lo[it] >= t["lo"][i] - 1e-12) and np.all(hi[it] <= t["hi"][i] + 1e-12)
        else:  # an inner node's children lie inside it
            assert np.all(t["lo"][i + 1] >= t["lo"][i] - 1e-12) and np.all(t["hi"][i + 1] <= t["hi"][i] + 1e-12)
    assert sorted(t["items"]) == list(range(len(lo)))


def test_bvh_skip_links_cover_every_box():
    rng = np.random.default_rng(1)
    c = rng.uniform(-1, 1, (300, 3))
    h = rng.uniform(0.01, 0.1, (300, 3))
    t = bvh.build(c - h, c + h)
    _check_bvh(t, c - h, c + h)


def test_navarrays_round_trip(tmp_path):
    det = water_tank()
    det.nav.save(str(tmp_path / "nav"))
    back = NavArrays.load(str(tmp_path / "nav"))
    for k, v in det.nav.host.items():
        np.testing.assert_array_equal(v, back.host[k])
    for tab in ("T", "K", "G"):
        for k, v in det.nav.csg[tab].items():
            np.testing.assert_array_equal(v, back.csg[tab][k])
    assert back.n == det.nav.n and back.names == det.nav.names


def test_example_tank_structure():
    det = example_tank()
    assert det.n_sensors == 8 and det.nav.n == 9
    names, pos = det.sensor_table()
    np.testing.assert_allclose(pos[1], [0, 0, 0.48], atol=1e-7)
    np.testing.assert_allclose(pos[6], [-0.25, 0, -0.48], atol=1e-7)
    assert [r.type for r in det.rows] == ["SPECULAR_METAL", "ABSORBER"]
    assert set(det.params().keys()) >= {"materials/Water/ABSLENGTH", "materials/Water/RINDEX",
                                        "surfaces/CathodeSurface/REFLECTIVITY", "sensors/pmt/qe"}
    p = det.params().scale("materials/Water/ABSLENGTH", 2.0)
    np.testing.assert_allclose(p["materials/Water/ABSLENGTH"].value, 2 * det.params()["materials/Water/ABSLENGTH"].value)
    assert p.explain("materials/Water/ABSLENGTH") == "set"
    assert det.params().info("materials/Water/ABSLENGTH").method == "reweight"


def test_water_tank_pmts_inside_and_apart():
    """1,000 caps, each inside the cylinder (every corner of its box) and no two boxes overlapping."""
    det = water_tank()
    assert det.n_sensors == 1000
    lo, hi = det.nav.lo_w[1:], det.nav.hi_w[1:]
    cap = solids.sphere_cap(0.10, 0.04)
    R, T = det.nav.host["R"].astype(float), det.nav.host["t"].astype(float)
    for k in range(1, det.nav.n):
        # sample the cap's surface: every point inside the tank
        u = np.random.default_rng(k).normal(size=(64, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        pts = 0.1 * u
        pts = pts[pts[:, 2] >= 0.06]
        w = pts @ R[k].T + T[k]
        assert np.all(np.hypot(w[:, 0], w[:, 1]) < 2.0) and np.all(np.abs(w[:, 2]) < 1.5)
    order = np.argsort(lo[:, 0])
    for a in range(len(order)):  # sweep for overlapping boxes
        i = order[a]
        for b in range(a + 1, len(order)):
            j = order[b]
            if lo[j, 0] > hi[i, 0]:
                break
            assert not (np.all(lo[i] < hi[j]) and np.all(lo[j] < hi[i])), (i, j)
    assert cap.inside(np.array([0.0, 0.0, 0.08])) and not cap.inside(np.array([0.0, 0.0, 0.05]))


def test_pack_layout_and_slots():
    det = water_tank()
    pk = pack.pack(det, grad_keys=["materials/Water/ABSLENGTH", "surfaces/Liner/REFLECTIVITY"])
    assert pk.layout[0] == pack.MAGIC and pk.n_grad == 17
    assert pk.grad_keys[0] == ("materials/Water/ABSLENGTH", 0, 16)
    # the slotted values are exactly the table's values
    absl = det.params()["materials/Water/ABSLENGTH"].value
    np.testing.assert_allclose(pk.values[pk.slots[:] >= 0][:16], absl, rtol=1e-6)
    # a new value changes the values, never the layout
    pk2 = pack.pack(det, det.params().scale("materials/Water/ABSLENGTH", 1.1))
    np.testing.assert_array_equal(pk.layout, pk2.layout)
    assert not np.array_equal(pk.values, pk2.values)


def test_edge_rows_follow_the_structure():
    """Rows reading a keyed cut or reflection constant are edge rows whatever its value (api.py:797-814)."""
    assert [r.edge() for r in example_tank().rows] == [True, False]
    assert [r.edge() for r in water_tank().rows] == [True, True, False]


def test_simulate_refuses_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="no GPU"):
        L.simulate(example_tank(), L.sources.Point((0, 0, 0), lam=400.0), samples=16, seed=0)


def test_csg_program_membership():
    """A tube with a bore is the outer cylinder minus the inner one."""
    t = solids.tube(1.0, 0.5, rmin=0.4)
    assert t.inside(np.array([0.7, 0, 0])) and not t.inside(np.array([0.2, 0, 0]))
    assert not t.inside(np.array([0.7, 0, 0.6]))
    tabs = navarrays.csg_tables([t])
    assert list(tabs["G"]["op"]) == [0, 0, 3]


def test_native_bvh_builder_builds_the_rules_trees():
    """The native builder (lh_bvh_build) builds exactly the trees the SAH rules give (tests/bvh_py.py, numpy):
    random boxes, coincident centroids, leaves of 1 and 2; and it is fast on 100,000 boxes."""
    import time

    from bvh_py import build_py

    rng = np.random.default_rng(7)
    cases = []
    for n in (1, 2, 3, 17, 500):
        c = rng.uniform(-3, 3, (n, 3))
        h = rng.uniform(0.01, 0.3, (n, 3))
        cases.append((c - h, c + h))
    same = np.zeros((40, 3)) + rng.uniform(0, 1, 3)
    cases.append((same - 0.1, same + 0.2 + rng.uniform(0, 0.1, (40, 3))))  # every centroid nearly one point
    cases.append((np.zeros((9, 3)), np.ones((9, 3))))  # every centroid one point: the halves
    for lo, hi in cases:
        for leaf in (1, 2):
            a, b = bvh.build(lo, hi, leaf), build_py(lo, hi, leaf)
            for k in ("lo", "hi", "skip", "start", "count", "items"):
                np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)
            assert a["depth"] == b["depth"]
            _check_bvh(a, lo, hi)
    c = rng.uniform(-50, 50, (100_000, 3))
    t0 = time.perf_counter()
    big = bvh.build(c - 0.05, c + 0.05)
    assert time.perf_counter() - t0 < 5.0 and len(big["skip"]) == 2 * 100_000 - 1


def test_more_solids_membership():
    """G4Cons (with an inner cone and an azimuth cut), G4Polycone, G4Ellipsoid (z cuts) and G4Trd as CSG leaves:
    membership of random points against their analytic definitions, and every member inside the solid's box."""
    rng = np.random.default_rng(3)

    def check(s, inside_fn, lo, hi, n=60000):
        x = rng.uniform(lo, hi, (n, 3))
        want, got = inside_fn(x), s.inside(x)
        assert np.array_equal(want, got), (s.name, int((want != got).sum()))
        assert want.sum() > 100
        assert np.all(x[got] >= s.lo - 1e-12) and np.all(x[got] <= s.hi + 1e-12)

    def frustum(x, r1, r2, hz, zc=0.0):
        z = x[:, 2] - zc
        return (np.abs(z) <= hz) & (np.hypot(x[:, 0], x[:, 1]) <= r1 + (r2 - r1) * (z + hz) / (2 * hz))

    s = solids.cons(0.1, 0.5, 0.15, 0.3, 0.4, phi0=0.3, dphi=4.0)
    check(s, lambda x: frustum(x, 0.5, 0.3, 0.4) & ~frustum(x, 0.1, 0.15, 0.4)
          & (np.mod(np.arctan2(x[:, 1], x[:, 0]) - 0.3, 2 * np.pi) <= 4.0), -0.6, 0.6)
    s = solids.cons(0.0, 0.5, 0.0, 0.5, 0.4, phi0=1.0, dphi=1.2)
    check(s, lambda x: frustum(x, 0.5, 0.5, 0.4) & (np.mod(np.arctan2(x[:, 1], x[:, 0]) - 1.0, 2 * np.pi) <= 1.2),
          -0.6, 0.6)
    zp, rmin, rmax = [-0.5, 0.0, 0.2, 0.6], [0.0, 0.1, 0.1, 0.0], [0.3, 0.5, 0.5, 0.2]

    def poly(x):
        out = np.zeros(len(x), bool)
        for k in range(3):
            hz, zc = 0.5 * (zp[k + 1] - zp[k]), 0.5 * (zp[k] + zp[k + 1])
            out |= frustum(x, rmax[k], rmax[k + 1], hz, zc) & ~frustum(x, rmin[k], rmin[k + 1], hz, zc)
        return out

    check(solids.polycone(zp, rmin, rmax), poly, -0.7, 0.7)
    check(solids.ellipsoid(0.3, 0.5, 0.7, -0.4, 0.6),
          lambda x: ((x[:, 0] / 0.3) ** 2 + (x[:, 1] / 0.5) ** 2 + (x[:, 2] / 0.7) ** 2 <= 1) & (x[:, 2] >= -0.4)
          & (x[:, 2] <= 0.6), -0.8, 0.8)

    def trd_in(x, x1=0.2, x2=0.5, y1=0.4, y2=0.1, hz=0.3):
        f = (x[:, 2] + hz) / (2 * hz)
        return (np.abs(x[:, 2]) <= hz) & (np.abs(x[:, 0]) <= x1 + (x2 - x1) * f) & (np.abs(x[:, 1]) <= y1 + (y2 - y1) * f)

    check(solids.trd(0.2, 0.5, 0.4, 0.1, 0.3), trd_in, -0.6, 0.6)




def test_gdml_reads_the_added_solids_and_mie(tmp_path):
    """The GDML reader's cone, tube with a phi cut, polycone, ellipsoid and trd (full lengths, mm) equal the solids
    built directly (membership of random points in each placement's frame), and a material's MIEHG constants
    become the Mie medium's keys."""
    from lumen_amd.detector.gdml import Recipe, load
    from toys import GDML_SOLIDS

    f = tmp_path / "solids.gdml"
    f.write_text(GDML_SOLIDS)
    det = load(str(f), Recipe("solids", "World"))
    names = [n.rsplit("/", 1)[-1] for n in det.nav.names]
    m = det.media[0]
    assert m.mie_length == "materials/Liq/MIEHG" and det.params()[m.mie_forward] == pytest.approx(0.6)
    assert det.params()[m.mie_ratio] == pytest.approx(0.7) and det.params()["materials/Liq/MIEHG"].value[0] == 20.0
    want = {"pc": solids.cons(0.1, 0.5, 0.15, 0.3, 0.4), "pt": solids.cons(0, 0.3, 0, 0.3, 0.2, np.radians(30),
                                                                          np.radians(120)),
            "pp": solids.polycone([-0.3, 0.0, 0.2], [0, 0.05, 0], [0.2, 0.3, 0.1]),
            "pe": solids.ellipsoid(0.2, 0.3, 0.4, -0.3, 0.35), "pd": solids.trd(0.2, 0.1, 0.15, 0.25, 0.3)}
    rng = np.random.default_rng(5)
    h = det.nav.host
    for pv, s in want.items():
        v = names.index(pv)
        x = rng.uniform(-0.6, 0.6, (20000, 3))
        got = _inside_kind(det.nav, int(h["kind"][v]), x)
        assert np.array_equal(got, s.inside(x)), pv
        assert np.allclose(h["t"][v], {"pc": (-1, -1, 0), "pt": (1, -1, 0), "pp": (-1, 1, 0), "pe": (1, 1, 0),
                                       "pd": (0, 0, 1.2)}[pv])


def test_navigator_build_cache(tmp_path):
    """``navarrays.cached`` stores the arrays under build/nav-<sha256> and reads them back unchanged; the key follows
    the inputs (a moved volume, another precision) and the same inputs give the same key."""
    vols = [navarrays.Volume("box", 0, np.eye(3), np.zeros(3), -1, 0),
            navarrays.Volume("ball", 1, np.eye(3), np.array([0.1, 0.0, 0.0]), 0, 1)]
    kinds = [solids.box(1.0, 1.0, 1.0), solids.orb(0.2)]
    nav, k = navarrays.cached(vols, kinds, 2, root=str(tmp_path))
    assert re.fullmatch(r"[0-9a-f]{64}", k) and (tmp_path / "build" / f"nav-{k}" / "meta.json").is_file()
    assert sorted(p.name for p in (tmp_path / "build").iterdir()) == [f"nav-{k}"]  # no temporary left behind
    again, k2 = navarrays.cached(vols, kinds, 2, root=str(tmp_path))
    assert k2 == k and again is not nav
    for name, a in nav.host.items():
        assert a.dtype == again.host[name].dtype and np.array_equal(a, again.host[name]), name
    for tab in ("T", "K", "G"):
        for name, a in nav.csg[tab].items():
            assert np.array_equal(a, again.csg[tab][name]), (tab, name)
    assert again.names == nav.names and again.stats == nav.stats and again.n == nav.n
    moved = [vols[0], navarrays.Volume("ball", 1, np.eye(3), np.array([0.1, 0.0, 1e-7]), 0, 1)]
    assert navarrays.key(moved, kinds, 2, precision="float32", leaf_size=bvh.LEAF_SIZE) != k
    assert navarrays.cached(vols, kinds, 2, precision="float64", root=str(tmp_path))[1] != k


# ------------------------------------------------------------------------------------ the cache entry layout
def test_nav_entry_in_the_reference_layout(tmp_path):
    """The reference's cache entry layout (geometry/cache.py:187-230): the fixture tests/golden/nav_entry_example_tank
    (written by tests/golden/make_nav_entry.py, an encoder of its own) reads back as the example tank's arrays, and
    ``NavArrays.save`` writes the same tree and the same arrays."""
    import json

    from lumen_amd.detector import example_tank
    from lumen_amd.detector.navarrays import NavArrays

    golden = os.path.join(ROOT, "tests", "golden", "nav_entry_example_tank")
    nav = example_tank().nav
    got = NavArrays.load(golden)
    assert got.n == nav.n and got.names == nav.names and got.n_mesh == 0
    for k, v in nav.host.items():
        assert np.array_equal(got.host[k], v), k
    for t in "TKG":
        for k, v in nav.csg[t].items():
            assert np.array_equal(got.csg[t][k], v), (t, k)
    np.testing.assert_array_equal(got.lo_w, nav.lo_w)
    nav.save(str(tmp_path))
    mine, ref = json.load(open(tmp_path / "meta.json")), json.load(open(os.path.join(golden, "meta.json")))
    assert mine == ref
    assert mine["dc"] == "lumen.geometry.navigator_build:NavArrays" and list(mine["f"])[:3] == ["host", "csg", "n"]
    with np.load(tmp_path / "arrays.npz") as a, np.load(os.path.join(golden, "arrays.npz")) as b:
        assert sorted(a.files) == sorted(b.files) and all(a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k])
                                                          for k in a.files)


def test_nav_entry_refuses_foreign_records_and_meshes(tmp_path):
    import json
    import shutil

    from lumen_amd.detector.navarrays import NavArrays

    golden = os.path.join(ROOT, "tests", "golden", "nav_entry_example_tank")
    for change, what in ((lambda m: m.__setitem__("dc", "lumen.geometry.other:Thing"), "not a NavArrays"),
                         (lambda m: m["f"].__setitem__("n_mesh", 2), "meshes")):
        d = tmp_path / what.replace(" ", "_")
        shutil.copytree(golden, d)
        meta = json.load(open(d / "meta.json"))
        change(meta)
        json.dump(meta, open(d / "meta.json", "w"))
        with pytest.raises(ValueError, match=what):
            NavArrays.load(str(d))


def test_nav_cache_rebuilds_a_damaged_entry(tmp_path):
    from lumen_amd.detector import navarrays
    from toys import sphere_toy

    det = sphere_toy()
    vols = [navarrays.Volume("shell", 0, np.eye(3), np.zeros(3), -1, 1),
            navarrays.Volume("shell/water", 1, np.eye(3), np.zeros(3), 0, 0)]
    from lumen_amd.detector import solids

    kinds = [solids.orb(1.05, name="shell"), solids.orb(1.0, name="water")]
    nav, k = navarrays.cached(vols, kinds, 2, root=str(tmp_path))
    entry = tmp_path / "build" / f"nav-{k}"
    (entry / "meta.json").write_text("{\"format\": \"something else\"}")
    again, k2 = navarrays.cached(vols, kinds, 2, root=str(tmp_path))
    assert k2 == k and again.n == nav.n and np.array_equal(again.host["lo"], nav.host["lo"])
    assert navarrays.NavArrays.load(str(entry)).n == nav.n  # the entry was written again
    assert det.nav.n == nav.n


def test_gate_grid():
    from lumen_amd.api import gated

    det = example_tank()
    g = gated(det, (2.0, 12.0), 0.5)
    assert g.time_grid.t0 == 2.0 and g.time_grid.dt == 0.5 and g.time_grid.nt == 21
    assert gated(det, (2.0, 12.0), 0.5) is g  # one model per gate
    assert gated(det, (0.0, 10.2)).time_grid.nt == 12  # 1 ns nodes by default, the last at or past the gate's end
    with pytest.raises(ValueError):
        gated(det, (5.0, 5.0))


def test_gdml_glisur_surfaces_map_to_geant4s_paths():
    """G4OpBoundaryProcess's readings of a glisur surface (the reference's rules/optics.py:28-50, 453-522): ground
    with polish < 1 is the facet path (metal, or an interface that refracts), any other finish or polish 1 the
    polished reading; polish is a parameter at the surface's value (so the facets are drawn under the cover), and a
    complex index on a ground finish is refused."""
    from lumen_amd.detector.gdml import _surface_row
    from lumen_amd.detector.model import FLAG_COVER, FLAG_GLISUR, FLAG_REFRACTS, SurfaceRow

    def row(typ, fin, model, value, props=()):
        return _surface_row("S", dict(props={p: "m" for p in props}, type=typ, finish=fin, model=model, value=value))

    r, dflt = row("dielectric_metal", "ground", "glisur", 0.3)
    assert r.type == "GLISUR_METAL" and dflt == {"surfaces/S/REFLECTIVITY": 1.0, "surfaces/S/polish": 0.3}
    assert r.flags() & 0xFF == FLAG_GLISUR | FLAG_COVER
    assert row("dielectric_metal", "polished", "glisur", 0.3)[0].type == "SPECULAR_METAL"
    assert row("dielectric_metal", "ground", "glisur", 1.0)[0].type == "SPECULAR_METAL"
    r, dflt = row("dielectric_dielectric", "ground", "glisur", 0.5, ("TRANSMITTANCE",))
    assert r.type == "GLISUR_DIELECTRIC" and r.columns["T"] == "surfaces/S/TRANSMITTANCE"
    assert r.flags() & 0xFF == FLAG_GLISUR | FLAG_REFRACTS | FLAG_COVER
    assert row("dielectric_dielectric", "ground", "glisur", 1.0)[0].type == "DIELECTRIC"
    assert row("dielectric_metal", "ground", "unified", 0.1)[0].type == "UNIFIED_METAL"
    with pytest.raises(ValueError, match="complex index on a ground"):
        row("dielectric_metal", "ground", "unified", 0.1, ("REALRINDEX", "IMAGINARYRINDEX"))
    with pytest.raises(ValueError, match="polish"):
        SurfaceRow("S", "GLISUR_METAL", {"R": 0.9, "polish": 1.0, "eff": 0.0})
    with pytest.raises(ValueError, match="decides its own side"):
        SurfaceRow("S", "GLISUR_DIELECTRIC", {"R": 1.0, "polish": 0.5}, far_opaque=True)
    # a fixed polish draws from p_m itself; STRAIGHT is in the support only with a T that is not fixed at 0
    r = SurfaceRow("S", "GLISUR_DIELECTRIC", {"R": 1.0, "polish": 0.5})
    assert r.flags() & 0xFF == FLAG_GLISUR | FLAG_REFRACTS and (r.flags() >> 8) == 1 << 4
