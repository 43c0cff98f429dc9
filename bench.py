"""Benchmark: photons traced per second on one MI355X, the 1,000-PMT water tank with a 1 m Cherenkov track.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--photons 2**24] [--no-cpu] [--work-model]

A step is one forward pass of the photon loop over one batch (2^24 photons, FP32 tracing, FP64 tallies) through
the C ABI (``lh_trace``), the inputs resident on the device.  Prints one JSON line: the metric, whole-job photons/s,
the roofline of the photon-loop kernel (the algorithmic FP32 work per photon from the oracle's op count, over the
kernel's HIP-event time, against the MI355X FP32 vector peak) and the CPU baseline: the oracle (the FP64 C++/OpenMP
restatement of the reference's loop) timed on this host's cores on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photons traced/s (fwd; fwd+adjoint) on 1 MI355X at fixed MC std-error"
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector/matrix peak (TFLOP/s)
PEAK_HBM_GBS = 8000.0
WORK_MODEL = os.path.join(ROOT, "profiles", "work_model.json")
PMC_SUMMARY = os.path.join(ROOT, "profiles", "trace_kernel_water_tank_2p24_pmc.txt")


def work_model(det, src, n=1 << 14):
    """Return the algorithmic work per photon of this workload from the oracle's op-counting run (frozen in
    profiles/work_model.json; recomputed with --work-model)."""
    if os.path.exists(WORK_MODEL):
        with open(WORK_MODEL) as f:
            wm = json.load(f)
        if "water_tank_cherenkov" in wm:
            return wm["water_tank_cherenkov"]
    from oracle import oracle as O

    fp, ints, vert = O.OracleRun(det).count_ops(src, n, n, seed=0)
    return dict(photons=n, fp_ops_per_photon=fp / n, int_ops_per_photon=ints / n, vertices_per_photon=vert / n,
                fp_ops_per_vertex=fp / max(vert, 1))


def pmc_traffic(key="hbm_bytes_per_launch"):
    """Return a figure of the committed rocprofv3 counter summary of the photon-loop kernel (HBM bytes per launch by
    default; ``valu_issue_frac``, ``valu_busy``, ``valu_active_lanes``), or None."""
    if not os.path.exists(PMC_SUMMARY):
        return None
    vals = {}
    with open(PMC_SUMMARY) as f:
        for line in f:
            if "=" in line and not line.startswith("#"):
                k, v = line.split("=", 1)
                try:
                    vals[k.strip()] = float(v.split()[0])
                except ValueError:
                    pass
    return vals.get(key)


def cpu_model() -> str:
    """The host CPU's model name (lscpu's, from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(det, src, seconds=10.0):
    """Time the oracle (OpenMP on the host cores this process may use) on a bounded sample of the workload, in FP64
    and with the photon's arithmetic in FP32 (the reference's CPU path traces in FP32): per precision a warm-up, a
    pilot that sizes the sample to about ``seconds``, then the timed sample.  ``value`` is the FP32 figure."""  # noqa: D205, D209
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1)
    run = O.OracleRun(det)
    out = {}
    for fp32 in (True, False):
        run.trace(src, 1 << 10, seed=98, threads=threads, bvh=True, fp32=fp32)
        n = 1 << 14
        while True:
            t0 = time.perf_counter()
            run.trace(src, n, seed=99, threads=threads, bvh=True, fp32=fp32)
            dt = time.perf_counter() - t0
            if dt > 1.0 or n >= 1 << 22:
                break
            n *= 4
        n = int(min(1 << 24, max(n, n * seconds / max(dt, 1e-3))))
        log(f"cpu baseline {'fp32' if fp32 else 'fp64'}: {n} photons on {threads} threads")
        t0 = time.perf_counter()
        run.trace(src, n, seed=100, threads=threads, bvh=True, fp32=fp32)
        dt = time.perf_counter() - t0
        out["fp32" if fp32 else "fp64"] = (n, dt)
    (n, dt), (n64, dt64) = out["fp32"], out["fp64"]
    return dict(value=n / dt, unit="photons/s", cores=threads, kind="port", cpu=cpu_model(),
                value_fp64=n64 / dt64,
                sample=f"{n} photons of the same workload (1,000-PMT tank, 1 m Cherenkov track) in {dt:.1f} s on "
                       f"{threads} threads: the oracle's loop with the photon's arithmetic in FP32 and the reference's "
                       f"BVH walk (FP64 tallies); FP64 beside it: {n64} photons in {dt64:.1f} s")


def copy_calibration(nbytes=1 << 30, repeats=3):
    """A plain device-to-device copy of known size, the scale for the HBM byte counters (FETCH_SIZE, WRITE_SIZE)
    of the photon-loop kernel: run under the same counters, it reads and writes ``nbytes`` per dispatch."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(repeats):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "dispatches": repeats}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--photons", type=int, default=1 << 24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-lar", action="store_true", help="skip the liquid-argon voxel batch")
    ap.add_argument("--lar-side", type=int, default=16, help="voxels per side (16: the 4,096 sources)")
    ap.add_argument("--lar-photons", type=int, default=1 << 16, help="photons per voxel source")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--steps-adjoint", type=int, default=3)
    ap.add_argument("--batch-repeats", type=int, default=3)
    ap.add_argument("--work-model", action="store_true", help="recompute and store the op-count work model")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (the scale of the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not run (the photon loop shards by photon "
              "id ranges with one all-reduce of the FP64 tallies)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import lumen_amd as L
    from lumen_amd.api import engine
    from lumen_amd.detector import muon_track, water_tank

    det = water_tank()
    x0, d = muon_track()
    src = L.sources.CherenkovSegment("Water", start=tuple(x0), direction=tuple(d), length=1.0)
    if args.work_model:
        if os.path.exists(WORK_MODEL):
            os.remove(WORK_MODEL)
        wm = work_model(det, src)
        os.makedirs(os.path.dirname(WORK_MODEL), exist_ok=True)
        with open(WORK_MODEL, "w") as f:
            json.dump({"water_tank_cherenkov": wm, "source": "oracle op count (bench.py --work-model)"}, f, indent=1)
    wm = work_model(det, src)
    n = int(args.photons)
    eng = engine(det)
    dev = eng.device
    for w in range(args.warmup):
        L.simulate(det, src, samples=n, seed=w)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    lights = []
    for k in range(args.steps):  # each call only enqueues: the next step's host work overlaps this kernel
        lights.append(L.simulate(det, src, samples=n, seed=1000 + k))
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    kernel_ms = [x.kernel_ms for x in lights[-256:]]  # each launch's own HIP events, read after the timed region
    light = lights[-1]
    a = light.arrivals.double()
    rel_err = float(torch.sqrt(light.arrival2.sum()) / a.sum())
    ms_step = 1e3 * wall / args.steps
    value = n * args.steps / wall
    k_ms = float(np.mean(kernel_ms))
    fp_per_photon = wm["fp_ops_per_photon"]
    achieved = n * fp_per_photon / (k_ms * 1e-3) / 1e12
    traffic = pmc_traffic()
    out = {
        "metric": METRIC, "value": value, "unit": "photons/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "water_tank_1000pmt_cherenkov_fwd", "photons_per_step": n, "sensors": 1000,
                   "source": "1.0 m beta=1 track, seed 2", "tracing": "f32", "tallies": "f64",
                   "rel_std_error_total_arrivals": rel_err},
        "kernel_ms": k_ms,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "fp_ops_per_photon": fp_per_photon, "kernel": "lh::trace_kernel",
                     "valu_issue_frac": pmc_traffic("valu_issue_frac"), "valu_busy": pmc_traffic("valu_busy"),
                     "valu_active_lanes": pmc_traffic("valu_active_lanes")},
    }
    log(f"forward: {k_ms:.2f} ms per launch")
    if not args.no_adjoint:
        out["fwd_adjoint"] = fwd_adjoint_leg(det, src, n, args.steps_adjoint)
        log("forward + adjoint done")
        out["value_fwd_adjoint"] = out["fwd_adjoint"]["value"]
    if not args.no_batch:
        out["batched"] = batched_leg(det, args.batch_repeats)
        log("batched done")
    if not args.no_lar:
        out["lar_voxels"] = lar_leg(args.lar_side, args.lar_photons)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(det, src)
        log("cpu baseline done")
    print(json.dumps(out), flush=True)


C3_KEYS = ("materials/Water/ABSLENGTH", "materials/Water/RAYLEIGH", "surfaces/Liner/REFLECTIVITY",
           "surfaces/Liner/sigma_alpha")


def fwd_adjoint_leg(det, src, n, steps, fd_seeds=4):
    """Forward plus adjoint on the same tank: the Poisson NLL of pe against a target simulated at parameters
    perturbed by N(0, 5 %) (lumen_amd.gradcheck.TankLikelihood); gradient over the 16 absorption knots, the 16
    Rayleigh knots, the liner's reflectivity and roughness, and the 1,000 PMT QEs (after tracking).  Then the
    finite-difference comparison on the seeded 32-parameter subset: common-random-number central differences at
    h = 1e-3 and the adjoint over ``fd_seeds`` seeds each, timed; the gradients' spread over the seeds gives the
    photons finite differences need for the adjoint's error, so the speed-up is measured at matched MC error."""  # noqa: D205, D209
    import numpy as np
    import torch

    from lumen_amd.gradcheck import TankLikelihood, after_tracking, crn_central, values_of

    prob = TankLikelihood.build(det, src, target_samples=n)
    prob.adjoint(samples=n, seed=500)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lights = []
    for k in range(steps):  # enqueued back to back: the kernel times are read after the timed region
        grads, light = prob.adjoint(samples=n, seed=600 + k, to_host=False)
        lights.append(light)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    runs = [x.provenance["run"] for x in lights]
    fwd_ms = [r.kernel_ms for r in runs]
    adj_ms = [r.adjoint_ms for r in runs]
    n_params = sum(len(values_of(prob.params, k)) for k in prob.keys())
    n_track = sum(len(values_of(prob.params, k)) for k in C3_KEYS)
    t_step = wall / steps
    out = {"value": n / t_step, "unit": "photons/s", "ms_per_step": 1e3 * t_step, "photons_per_step": n,
           "forward_kernel_ms": float(np.mean(fwd_ms)), "adjoint_kernel_ms": float(np.mean(adj_ms)),
           "parameters": n_params, "loss": "Poisson NLL of pe vs a target at N(0, 5%)-perturbed parameters"}
    log("forward + adjoint timed")
    picks = prob.picks(32, seed=4)
    track = np.array([not after_tracking(k) for k, _ in picks])
    ga, gf, t_adj, t_fd = [], [], 0.0, 0.0
    for s in range(fd_seeds):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        grads, _ = prob.adjoint(samples=n, seed=900 + s)
        torch.cuda.synchronize()
        t_adj += time.perf_counter() - t0
        ga.append([grads[k][i] for k, i in picks])
        t0 = time.perf_counter()
        gf.append(crn_central(prob.loss, det, src, prob.params, [p for p, t in zip(picks, track) if t],
                              samples=n, seed=900 + s, h=1e-3))
        torch.cuda.synchronize()
        t_fd += time.perf_counter() - t0
    ga = np.asarray(ga)[:, track]
    gf = np.asarray(gf)
    mean_a, mean_f = ga.mean(0), gf.mean(0)
    sd_a, sd_f = ga.std(0, ddof=1), gf.std(0, ddof=1)
    rel = np.abs(mean_a - mean_f) / np.maximum(np.abs(mean_f), 1e-30)
    ratio = float(np.median((sd_f / np.maximum(sd_a, 1e-300)) ** 2))  # photons FD needs per adjoint photon
    t_forward = t_fd / (fd_seeds * 2 * int(track.sum()))  # one CRN forward of n photons, measured
    t_grad = t_adj / fd_seeds  # forward + adjoint + readout of n photons, measured
    out["finite_differences"] = {
        "subset": len(picks), "tracking_parameters_in_subset": int(track.sum()), "h": 1e-3, "seeds": fd_seeds,
        "crn_forward_ms": 1e3 * t_forward, "fwd_adjoint_ms": 1e3 * t_grad,
        "max_rel_diff_of_seed_means": float(rel.max()), "median_rel_diff_of_seed_means": float(np.median(rel)),
        "variance_ratio_fd_over_adjoint": ratio,
        "speedup_all_parameters": (2 * n_params + 1) * t_forward * ratio / t_grad,
        "speedup_tracking_parameters": (2 * n_track + 1) * t_forward * ratio / t_grad}
    out["speedup_vs_central_fd"] = out["finite_differences"]["speedup_all_parameters"]
    return out


def log(msg):
    """A progress line on stderr (the JSON result is stdout's one line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _poisson_cotangent(S, qe, target):
    """Return loss(tally) -> dLoss/dtally of the per-source Poisson NLL sum_s pe - target log pe, pe = qe charge."""
    import torch

    def loss(tally):
        pe = tally[:, 3 * S:4 * S] * qe
        cot = torch.zeros_like(tally)
        cot[:, 3 * S:4 * S] = qe * (1.0 - target / torch.clamp(pe, min=1e-30))
        return cot

    return loss


def batched_leg(det, repeats):
    """256 point sources with uneven photon counts (2^12..2^20): one graph-captured persistent launch with the
    autotuned work queue, forward and forward plus adjoint (the per-source Poisson NLL of pe against a target 5 %
    above the forward's, gradient over the absorption and Rayleigh knots, the liner's reflectivity and roughness),
    against one plain launch per source."""  # noqa: D205, D209
    import torch

    from lumen_amd.batch import BatchRun, c4_batch

    batch = c4_batch(det)
    run = BatchRun(det, batch, seed=11, grad_keys=C3_KEYS)
    S = run.eng.S
    target = 1.05 * run.forward()[:, 3 * S:4 * S].clone() * 0.25
    loss = _poisson_cotangent(S, 0.25, target)
    run.forward_adjoint(loss)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        run.forward()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / repeats
    t0 = time.perf_counter()
    for _ in range(repeats):
        run.forward_adjoint(loss)
    torch.cuda.synchronize()
    t_both = (time.perf_counter() - t0) / repeats
    run.sequential()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        run.sequential()
    torch.cuda.synchronize()
    t_seq = (time.perf_counter() - t0) / repeats
    run.sequential_adjoint(loss)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run.sequential_adjoint(loss)
    torch.cuda.synchronize()
    t_seq_both = time.perf_counter() - t0
    return {"sources": len(batch.sources), "photons": int(run.total), "batched_ms": 1e3 * t_batch,
            "batched_fwd_adjoint_ms": 1e3 * t_both, "sequential_ms": 1e3 * t_seq,
            "sequential_fwd_adjoint_ms": 1e3 * t_seq_both, "speedup": t_seq / t_batch,
            "speedup_fwd_adjoint": t_seq_both / t_both,
            "photons_per_s_batched": run.total / t_batch, "photons_per_s_batched_fwd_adjoint": run.total / t_both,
            "autotuned": {"chunk": run.controls.chunk, "blocks_per_cu": run.controls.blocks_per_cu,
                          "adjoint_chunk": run.adj_controls.chunk,
                          "adjoint_blocks_per_cu": run.adj_controls.blocks_per_cu}}


LAR_KEYS = ("materials/LAr/RAYLEIGH", "surfaces/PTFE/REFLECTIVITY")


def lar_leg(n_side, photons):
    """The liquid-argon box: n_side^3 voxel point sources of ``photons`` photons each at 128 nm in one graph-captured
    batched launch, forward and forward plus adjoint over the Rayleigh length and the foils' reflectivity (the 120
    QEs act after tracking) for the per-source Poisson NLL of pe against a target 5 % above the forward's."""  # noqa: D205, D209
    import torch

    from lumen_amd.batch import BatchRun
    from lumen_amd.detector import lar_box, voxel_points
    from lumen_amd.sources import Batch

    from lumen_amd.api import Controls

    det = lar_box()
    pts = voxel_points(n=n_side)
    batch = Batch(tuple(pts), tuple([int(photons)] * len(pts)))
    # one untimed forward (the target's); no graph (its capture would run the adjoint once more untimed)
    run = BatchRun(det, batch, seed=13, grad_keys=LAR_KEYS, graph=False, tune=False, controls=Controls())
    S = run.eng.S
    target = 1.05 * run.forward()[:, 3 * S:4 * S].clone() * 0.25
    loss = _poisson_cotangent(S, 0.25, target)
    torch.cuda.synchronize()
    log(f"lar: {len(pts)} sources, {run.total} photons: warm forward done")
    t0 = time.perf_counter()
    run.forward()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    log(f"lar: forward {t_fwd:.2f} s")
    t0 = time.perf_counter()
    run.forward_adjoint(loss)
    torch.cuda.synchronize()
    t_both = time.perf_counter() - t0
    log(f"lar: forward + adjoint {t_both:.2f} s")
    h = run.health.cpu().numpy()
    return {"workload": f"lar_box_120pmt_{n_side}^3_voxels_2^{int(photons).bit_length() - 1}_photons",
            "sources": len(pts), "photons": int(run.total), "forward_ms": 1e3 * t_fwd,
            "fwd_adjoint_ms": 1e3 * t_both, "photons_per_s": run.total / t_fwd,
            "photons_per_s_fwd_adjoint": run.total / t_both, "vertices_per_photon": float(h[12] / max(h[13], 1.0)),
            "parameters": int(run.n_grad) + S}


if __name__ == "__main__":
    main()
