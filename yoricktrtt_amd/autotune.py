"""Autotuner of the axial kernel launch shapes.

For a :class:`~yoricktrtt_amd.projectors.Projector` on the closed-form
axial path, time every candidate launch shape (forward tile, backward tile,
forward LDS window, views per forward workgroup) on the GPU with the plan's
own inputs, check that each candidate reproduces the default launch, and
keep the fastest forward + back projection.  Results are cached in a JSON
table keyed by the geometry signature (kind, sizes, view count, footprint
width bound, precision) and the GPU name, so a geometry is tuned once per
machine.

Results differ between launch shapes only in FP32 summation order
(checked here to 1e-5 relative before a candidate is timed); a candidate
whose first run takes over twice the best time so far is not timed further.
"""
from __future__ import annotations

import itertools
import json
import os
import time

from . import _lib

FWD_TILES = {0: "8x4", 1: "4x4", 2: "8x8"}
# backward tiles 3 to 5 span 128 slices (two voxels per lane); plans whose
# column profiles over 128 slices exceed 256 pixels reject them
BWD_TILES = {0: "8x8", 1: "8x4", 2: "4x4", 3: "8x4x128", 4: "4x4x128", 5: "8x8x128"}
# LDS window sizes tried per forward tile (a window too small for a tile's
# detector footprint sends its voxels down the slow global-atomic path)
WINDOWS = {0: (12288, 16384, 24576), 1: (7168, 7424, 8192, 12288), 2: (24576, 32768)}
VPBS = (0, 1, 2, 3, 4)
CHUNKS = (0, 2, 4)  # views per forward scratch round (0: as many as fit the budget)
BWD_WINDOWS = {0: (16384, 24576, 32768), 1: (12288, 16384, 24576), 2: (8192, 12288, 16384),
               3: (18432, 20480, 22528, 24576), 4: (14336, 16384, 24576), 5: (28672, 32768)}
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "yoricktrtt_amd", "tuning.json")


def signature(P) -> str:
    """Cache key of a plan: geometry kind, sizes, views, width bound, type."""
    g = P.geom
    return (f"{g.kind}:{g.nx}x{g.ny}x{g.nz}x{P.nt}:{g.nu}x{g.nv}:{g.nviews}:W{P.path}:"
            f"{'f64' if P.precision == _lib.F64 else 'f32'}")


def get_tuning(P) -> dict:
    t = _lib.Tuning()
    if P.lib.trtt_hip_plan_get_tuning(P.handle, t) != 0:
        raise RuntimeError(_lib.last_error())
    return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}


def set_tuning(P, **kw):
    cur = get_tuning(P)
    cur.update(kw)
    t = _lib.Tuning(**cur)
    if P.lib.trtt_hip_plan_set_tuning(P.handle, t) != 0:
        raise RuntimeError(_lib.last_error())


def candidates(windows=WINDOWS, vpbs=VPBS, chunks=None):
    """Forward and backward candidates, tuned independently (the two
    kernels share no launch parameter).  ``windows`` maps a forward tile to
    its window sizes, or is one tuple for every tile.  Forward windows are
    added to the projections with atomics (one round over all views), or
    go through scratch and the reduction kernel in rounds of ``chunks``
    views."""
    per_tile = windows if isinstance(windows, dict) else {t: windows for t in FWD_TILES}
    chunks = CHUNKS if chunks is None else chunks
    modes = [(1, 0)] + [(0, c) for c in chunks]
    fwd = [dict(fwd_tile=t, fwd_window_bytes=wb, fwd_views_per_block=v, fwd_flush=f,
                fwd_chunk_views=c)
           for t in FWD_TILES for wb, v, (f, c) in itertools.product(per_tile[t], vpbs, modes)]
    bwd = [dict(bwd_tile=t, bwd_window_bytes=wb) for t in BWD_TILES for wb in BWD_WINDOWS[t]]
    return fwd, bwd


def _load_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(path, table):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def device_name(torch) -> str:
    """Architecture and CU count of the current GPU (the marketing name is
    not reliable across driver installs)."""
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = getattr(p, "gcnArchName", "") or p.name
    return f"{arch.split(':')[0]}x{p.multi_processor_count}"



def autotune(P, x=None, y=None, cache=DEFAULT_CACHE, reps=3, windows=WINDOWS, vpbs=VPBS,
             chunks=None, force=False, log=None):
    """Tune plan P in place and return the chosen tuning with its timings.

    x (volume) and y (sinogram) are device tensors of the plan's shapes; by
    default uniform random inputs.  With a cache hit (same signature and GPU)
    the stored tuning is applied without timing."""
    import torch

    if P.path <= 0:
        return dict(tuning=None, reason="general kernels have no launch shape to tune")
    key = f"{device_name(torch)}|{signature(P)}"
    table = _load_cache(cache) if cache else {}
    if not force and key in table:
        set_tuning(P, **table[key]["tuning"])
        return dict(table[key], cached=True)
    g = torch.Generator(device="cuda").manual_seed(0)
    if x is None:
        x = torch.rand(P.vol_shape, generator=g, device="cuda", dtype=P.dtype)
    if y is None:
        y = torch.rand(P.sino_shape, generator=g, device="cuda", dtype=P.dtype)
    sino = torch.zeros(P.sino_shape, dtype=P.dtype, device="cuda")
    vol = torch.zeros(P.vol_shape, dtype=P.dtype, device="cuda")
    default = get_tuning(P)
    ref_f = P.forward(x, sino.zero_()).clone()
    ref_b = P.backward(y, vol.zero_()).clone()
    scale_f = float(ref_f.abs().max())
    scale_b = float(ref_b.abs().max())

    def once(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def timed(fn):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2]

    fwd_c, bwd_c = candidates(windows, vpbs, chunks)
    best_f, best_b = None, None
    rows = []
    for c in fwd_c:
        try:
            set_tuning(P, **c)
        except RuntimeError:  # a tile the plan's geometry does not allow
            rows.append(dict(kind="fwd", **c, ms=None, rejected=True))
            continue
        t1 = once(lambda: P.forward(x, sino.zero_()))
        err = float((sino - ref_f).abs().max()) / scale_f
        if err > 1e-5:
            raise RuntimeError(f"tuning {c} changes the forward projection ({err:.2e})")
        if best_f is not None and t1 > 2.0 * best_f[1]:
            rows.append(dict(kind="fwd", **c, ms=1e3 * t1, dropped=True))
            continue
        t = timed(lambda: P.forward(x, sino.zero_()))
        rows.append(dict(kind="fwd", **c, ms=1e3 * t))
        if log:
            log(f"fwd {c}: {1e3 * t:.2f} ms")
        if best_f is None or t < best_f[1]:
            best_f = (c, t)
    set_tuning(P, **best_f[0])
    for c in bwd_c:
        try:
            set_tuning(P, **c)
        except RuntimeError:  # a tile the plan's geometry does not allow
            rows.append(dict(kind="bwd", **c, ms=None, rejected=True))
            continue
        t1 = once(lambda: P.backward(y, vol.zero_()))
        err = float((vol - ref_b).abs().max()) / scale_b
        if err > 1e-5:
            raise RuntimeError(f"tuning {c} changes the back projection ({err:.2e})")
        if best_b is not None and t1 > 2.0 * best_b[1]:
            rows.append(dict(kind="bwd", **c, ms=1e3 * t1, dropped=True))
            continue
        t = timed(lambda: P.backward(y, vol.zero_()))
        rows.append(dict(kind="bwd", **c, ms=1e3 * t))
        if log:
            log(f"bwd {c}: {1e3 * t:.2f} ms")
        if best_b is None or t < best_b[1]:
            best_b = (c, t)
    tuning = dict(default)
    tuning.update(best_f[0])
    tuning.update(best_b[0])
    set_tuning(P, **tuning)
    res = dict(tuning=tuning, fwd_ms=1e3 * best_f[1], bwd_ms=1e3 * best_b[1],
               default=default, signature=signature(P), device=device_name(torch), rows=rows)
    if cache:
        table = _load_cache(cache)
        table[key] = {k: v for k, v in res.items() if k != "rows"}
        _store_cache(cache, table)
    return dict(res, cached=False)
