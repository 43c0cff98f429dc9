"""The launch autotuner: picks the work-queue chunk (photons a wavefront claims at once: the photons per wavefront of
the persistent loop, whose 64 lanes each refill from that claim) and the persistent grid's resident blocks per CU for
a staged batch, for its forward and its adjoint launch separately, by timing launches of it, and remembers the
choices per (detector, launch kind, photons, sources, GPU) in ``$LUMEN_AMD_CACHE`` (default ``~/.cache/lumen_amd/autotune.json``).

Small chunks balance uneven sources at the end of a launch; large ones cut queue traffic.  Blocks per CU stay a
search dimension (``tune(blocks=...)``), but the kernels' one 768-thread workgroup per CU leaves a single value.  The search is exhaustive over a small grid; each point is timed on the
real launch (best of two), which costs a few launches once per configuration.
"""

from __future__ import annotations

import json
import os
import time

CHUNKS = (64, 128, 256, 512, 1024, 4096)
# the persistent grid's resident workgroups per CU: each photon-loop kernel's 768-thread workgroup of 12 wavefronts
# fills a CU at its register budget, so "as many as fit" (0) is one and the search has nothing else to try
BLOCKS = (0,)


def _cache_path():
    root = os.environ.get("LUMEN_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "lumen_amd"))
    return os.path.join(root, "autotune.json")


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d):
    try:
        os.makedirs(os.path.dirname(_cache_path()), exist_ok=True)
        with open(_cache_path(), "w") as f:
            json.dump(d, f, indent=1)
    except OSError:
        pass


def _key(run, what):
    import torch

    name = torch.cuda.get_device_name(run.eng.device)
    return f"{run.det.name}|{what}|photons={run.total}|sources={len(run.rows)}|{name}"


ADJ_CHUNKS = (128, 512, 2048)  # the adjoint's launches cost ~4 forward ones: a coarser grid, timed once each


def tune(run, what="forward", chunks=None, blocks=BLOCKS, repeats=None):
    """Time the staged batch ``run`` (a ``lumen_amd.batch.BatchRun``) at every (chunk, blocks_per_cu) of the
    forward launch (``what`` "forward") or of the adjoint launch ("adjoint": its cotangent as ``run.cot`` holds
    it); return (the best Controls, {point: ms}).  Ties go to the first point timed (as many blocks as fit)."""  # noqa: D205, D209
    import torch

    from .api import Controls

    adj = what == "adjoint"
    chunks = chunks or (ADJ_CHUNKS if adj else CHUNKS)
    repeats = repeats or (1 if adj else 2)
    base = Controls()
    times = {}
    for ch in chunks:
        for bl in blocks:
            ctl = base._replace(chunk=ch, blocks_per_cu=bl)
            best = float("inf")
            for _ in range(repeats):
                torch.cuda.synchronize(run.eng.device)
                t0 = time.perf_counter()
                if adj:
                    run.grad.zero_()
                    run._launch_adjoint(ctl)
                else:
                    run._fwd_body(ctl)
                torch.cuda.synchronize(run.eng.device)
                best = min(best, time.perf_counter() - t0)
            times[(ch, bl)] = 1e3 * best
    (ch, bl) = min(times, key=times.get)
    return base._replace(chunk=ch, blocks_per_cu=bl), times


def tuned_controls(run, what="forward"):
    """Return the remembered best Controls of this configuration and launch kind, tuning it first when it is
    new."""  # noqa: D205, D209
    from .api import Controls

    d = _load()
    k = _key(run, what)
    if k in d:
        return Controls()._replace(chunk=int(d[k]["chunk"]), blocks_per_cu=int(d[k]["blocks_per_cu"]))
    ctl, times = tune(run, what)
    d[k] = dict(chunk=ctl.chunk, blocks_per_cu=ctl.blocks_per_cu,
                ms={f"{c},{b}": round(t, 4) for (c, b), t in times.items()})
    _store(d)
    return ctl
