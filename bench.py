"""Cone-beam 256^3 forward+back projection throughput on one MI355X.

Workload (the 3-D cone-beam configuration): 256^3 voxels of 1 mm on a cubic
B-spline basis, 360 views over 360 deg, flat detector of 512 (transaxial, v)
x 384 (axial, u) pixels at 1 mm, source-to-isocentre 1000 mm,
source-to-detector 1500 mm, FP32 data, FP64 geometry.  One step = one
forward projection of a volume over all 360 views plus one back projection
of a sinogram over all 360 views (closed-form cubic footprint,
trtt_CB3D_cubic semantics), inputs resident in HBM.

Prints one JSON line with the throughput, the per-kernel roofline (HIP
events on the launch stream) and the reference C plugin timed on the host
cores over a bounded sample of views.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

NX = NY = NZ = 256
NU, NV = 384, 512
NVIEWS = 360
RSC, RSD = 1000.0, 1500.0
HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E
FP32_PEAK = 157.3e12       # FLOP/s, MI355X FP32 vector
# Footprint weights of one pass over the 360 views (sum over voxels and views
# of npix_u * npix_v, the reference's window sizes), counted with the oracle's
# FP64 geometry pass by tests/golden/make_footprint_count.py and frozen in
# tests/golden/footprint_count_cone256.json; each weight is one multiply-add
# in the forward and one in the back projection: 4 flops per weight a pair.
FOOTPRINT_WEIGHTS = 290543347804


def geometry():
    from yoricktrtt_amd.geometry import cone_geometry

    theta = np.arange(NVIEWS) * (2 * math.pi / NVIEWS)
    return cone_geometry(NX, NY, NZ, 1.0, NU, NV, 1.0, 1.0, theta, RSC, RSD)


def phantom(seed=3, n_ell=30):
    """Seeded random ellipsoids (values in [0,1]) on the 256^3 grid."""
    rng = np.random.default_rng(seed)
    vol = np.zeros((NZ, NY, NX), dtype=np.float32)
    half = np.array([(NX - 1) / 2, (NY - 1) / 2, (NZ - 1) / 2])
    for _ in range(n_ell):
        c = rng.uniform(-60, 60, size=3)
        a = rng.uniform(8, 50, size=3)
        mu = np.float32(rng.uniform(0, 1))
        lo = np.maximum(np.floor(c - a + half).astype(int), 0)
        hi = np.minimum(np.ceil(c + a + half).astype(int) + 1, [NX, NY, NZ])
        zz, yy, xx = np.meshgrid(*(np.arange(lo[d], hi[d]) - half[d] for d in (2, 1, 0)),
                                 indexing="ij")
        q = ((xx - c[0]) / a[0]) ** 2 + ((yy - c[1]) / a[1]) ** 2 + ((zz - c[2]) / a[2]) ** 2
        vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += np.where(q <= 1.0, mu, np.float32(0))
    return vol


TRAFFIC_FILE = os.path.join(HERE, "profiles", "axial_fwd_cone256_traffic.txt")
ISSUE_FILE = os.path.join(HERE, "profiles", "axial_cone256_issue_counters.txt")


def profile_values(path, tuning):
    """`name = value` lines of a committed rocprofv3 summary, or None when
    the summary is absent or was measured with another launch shape than
    `tuning` (its `# tuning:` line)."""
    try:
        text = open(path).read()
    except OSError:
        return None
    vals, shape = {}, None
    for line in text.splitlines():
        if line.startswith("# tuning:"):
            shape = dict(kv.split("=") for kv in line[len("# tuning:"):].replace(",", " ").split())
        elif "=" in line and not line.startswith("#"):
            k, v = line.split("=")
            vals[k.strip()] = float(v)
    if not tuning or not shape or any(int(shape[k]) != int(tuning[k]) for k in shape):
        return None
    return vals


def measured_traffic(tuning, kernel):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated on a 1 GiB copy),
    when it was measured with the launch shape in use; else None."""
    vals = profile_values(TRAFFIC_FILE, tuning)
    key = "bwd_traffic_bytes_per_launch" if kernel == "axial::bwd" else "fwd_traffic_bytes_per_launch"
    return int(vals[key]) if vals and key in vals else None


def measured_issue(tuning, kernel, launch_ms):
    """VALU issue rate of the dominant kernel: its wave64 VALU instructions
    per launch (committed SQ_INSTS_VALU summary, same launch shape) over the
    launch time measured now, against 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles
    (non-packed instructions); None without a matching summary."""
    vals = profile_values(ISSUE_FILE, tuning)
    key = "bwd_valu_insts_per_launch" if kernel == "axial::bwd" else "fwd_valu_insts_per_launch"
    if not vals or key not in vals or launch_ms <= 0:
        return None
    rate = vals[key] / (launch_ms * 1e-3)
    peak = vals["valu_issue_peak_per_s"]
    return {"unit": "wave64 VALU instructions/s", "insts_per_launch": int(vals[key]),
            "achieved": float(f"{rate:.4g}"), "peak": peak, "frac": round(rate / peak, 4)}


def algorithmic_bytes(esz=4):
    """Per-view-operator bytes: each view reads the volume and writes its
    projection (forward), or reads its projection and read-modify-writes the
    volume (transpose)."""
    vol = NX * NY * NZ * esz
    det = NU * NV * esz
    fwd = NVIEWS * (vol + det)
    bwd = NVIEWS * (det + 2 * vol)
    compulsory = 2 * vol + 2 * NVIEWS * det
    return fwd, bwd, compulsory


def _reference_rate(path, geo, threads):
    """View pairs per second of the reference C plugin built at `path` on
    `threads` host threads, each projecting one view forward and one view
    back at the full size (the ctypes calls release the GIL)."""
    from oracle.oracle import Reference

    ref = Reference(path)
    rng = np.random.default_rng(1)
    vol = rng.uniform(size=(NZ, NY, NX))
    sino = rng.uniform(size=(NV, NU))
    views = np.linspace(0, NVIEWS - 1, threads).astype(int)
    outs_f = [np.zeros((NV, NU)) for _ in range(threads)]
    outs_b = [np.zeros((NZ, NY, NX)) for _ in range(threads)]

    def work(t):
        v = views[t]
        ref.CB3D_cubic(outs_f[t], vol, NX, NY, NZ, 1.0, NU, NV, 1.0, 1.0, geo.xos[v], geo.xsd[v], 0)
        ref.CB3D_cubic(outs_b[t], sino, NX, NY, NZ, 1.0, NU, NV, 1.0, 1.0, geo.xos[v], geo.xsd[v], 1)

    t0 = time.perf_counter()
    ths = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    dt = time.perf_counter() - t0
    return threads / dt, dt


def cpu_reference_baseline(geo, threads):
    """The reference C plugin (oracle/_ref, compiled from its own sources)
    on every host core this process may use, at -O2 (the parity build) and
    at -O3 -march=x86-64-v3; the faster build is the baseline value."""
    from oracle.oracle import REF_O3_SO, REF_SO, Reference

    if not Reference.available(REF_SO):
        return None
    builds = {"O2": REF_SO}
    if Reference.available(REF_O3_SO):
        builds["O3_x86-64-v3"] = REF_O3_SO
    rates = {}
    for name, path in builds.items():
        vps, dt = _reference_rate(path, geo, threads)
        rates[name] = {"pairs_per_s": vps / NVIEWS, "wall_s": round(dt, 2)}
    best = max(rates, key=lambda k: rates[k]["pairs_per_s"])
    return {"value": rates[best]["pairs_per_s"], "unit": "pairs/s", "cores": threads,
            "kind": "reference", "nproc": os.cpu_count(), "build": best, "builds": rates,
            "sample": f"{threads} threads x (1 forward + 1 back view of trtt_CB3D_cubic, FP64, "
                      f"256^3 -> 512x384) per build, scaled to 360 views"}


def error_vs_fp64(geo, x_host, y_host, sino, vol, views=(0, 90, 180, 270), slab=(124, 132)):
    """Max relative error of the FP32 GPU results of the last timed step
    against the FP64 oracle: the forward on a few views, the back
    projection of all 360 views on a slab of slices."""
    from oracle.oracle import Oracle

    o = Oracle()
    views = np.asarray(views)
    rf = o.views_fwd3d(x_host.astype(np.float64), NX, NY, NZ, 1.0, NU, NV, 1.0, 1.0,
                       geo.xos[views], geo.xsd[views], True)
    got_f = sino[views].double().cpu().numpy()
    rb = o.views_bwd3d(y_host.astype(np.float64), NX, NY, NZ, 1.0, NU, NV, 1.0, 1.0, geo.xos,
                       geo.xsd, True, k0=slab[0], k1=slab[1])
    got_b = vol[slab[0]:slab[1]].double().cpu().numpy()
    ef = float(max(np.abs(got_f[q] - rf[q]).max() / np.abs(rf[q]).max() for q in range(views.size)))
    eb = float(np.abs(got_b - rb).max() / np.abs(rb).max())
    return {"forward_max_rel": float(f"{ef:.3g}"), "backward_max_rel": float(f"{eb:.3g}"),
            "measure": "max|gpu - fp64| / max|fp64|",
            "sample": f"forward views {list(map(int, views))}; back projection of all "
                      f"{NVIEWS} views on slices {slab[0]}..{slab[1] - 1}"}


def _pool_time(fn, n, threads):
    """Wall seconds of fn(i) for i < n over `threads` host threads (the
    ctypes calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(fn, range(n)))
    return time.perf_counter() - t0


def _rel(a, b):
    return float(f"{float(np.abs(a - b).max() / np.abs(b).max()):.3g}")


def _gpu_pair(P, x, reps=5):
    """Median forward and back projection ms of plan P on x (HBM-resident)."""
    import torch

    sino = torch.zeros(P.sino_shape, dtype=P.dtype, device=x.device)
    vol = torch.zeros(P.vol_shape, dtype=P.dtype, device=x.device)
    ts = []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        P.forward(x, sino.zero_())
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        P.backward(sino, vol.zero_())
        torch.cuda.synchronize()
        ts.append((t1 - t0, time.perf_counter() - t1))
    f, b = np.median(np.array(ts[1:]), axis=0)
    return 1e3 * f, 1e3 * b, sino, vol


def _report_2d(cfg, geo, fan, prec, workload, ref, threads):
    """GPU plan vs the reference C plugin on every view of a 2-D
    configuration: both operators over all views on the host threads (the
    full workload, timed), and the GPU results against those FP64 outputs."""
    import torch

    from yoricktrtt_amd.projectors import Projector

    rng = np.random.default_rng(cfg)
    x = rng.uniform(size=(geo.ny, geo.nx))
    P = Projector(geo, model="cubic", precision=prec)
    dt = torch.float64 if prec == "f64" else torch.float32
    fms, bms, sino, vol = _gpu_pair(P, torch.tensor(x, dtype=dt, device="cuda"))
    ys = sino.double().cpu().numpy()   # the back projection's input, as on the GPU
    back_gpu = vol.double().cpu().numpy()
    P.close()
    fn = ref.FB2D_cubic if fan else ref.PB2D_cubic
    V = geo.nviews
    fwd = np.zeros((V, geo.nv))
    chunks = np.array_split(np.arange(V), threads)
    parts = [np.zeros((geo.ny, geo.nx)) for _ in chunks]

    def f(i):
        fn(fwd[i], x, geo.nx, geo.ny, geo.s_scl, geo.nv, geo.v_scl, geo.xos[i], geo.xsd[i], 0)

    def b(c):
        for i in chunks[c]:
            fn(parts[c], ys[i], geo.nx, geo.ny, geo.s_scl, geo.nv, geo.v_scl, geo.xos[i],
               geo.xsd[i], 1)

    tf = _pool_time(f, V, threads)
    tb = _pool_time(b, len(chunks), threads)
    back = np.sum(parts, axis=0)
    # short operators: repeat to about 2 s of host time, keep the median
    reps = int(min(20, max(0, 2.0 / (tf + tb) - 1)))
    if reps:
        fwd_keep = fwd.copy()
        ts = [(tf, tb)]
        for _ in range(reps):
            for q in parts:
                q[...] = 0.0
            ts.append((_pool_time(f, V, threads), _pool_time(b, len(chunks), threads)))
        tf, tb = np.median(np.array(ts), axis=0)
        fwd = fwd_keep
    return {"config": cfg, "workload": workload,
            "gpu": {"fwd_ms": round(fms, 3), "bwd_ms": round(bms, 3),
                    "pairs_per_s": round(1e3 / (fms + bms), 2), "dtype": prec},
            "cpu_reference": {"pairs_per_s": round(1.0 / (tf + tb), 3), "wall_s": round(tf + tb, 3),
                              "cores": threads, "kind": "reference",
                              "sample": f"every view, forward and back ({'trtt_FB2D_cubic' if fan else 'trtt_PB2D_cubic'}, FP64), median of {reps + 1} runs"},
            "gpu_over_cpu": round((tf + tb) * 1e3 / (fms + bms), 1),
            "error_vs_fp64": {"forward_max_rel": _rel(ys, fwd), "backward_max_rel": _rel(back_gpu, back),
                              "sample": "every view"}}


def _report_lf(geo, prm, beta, threads):
    """Configuration 2's Long & Fessler projector on the GPU against the
    FP64 oracle over every view (the LF reference needs yapi.h and is not
    built here: the oracle's OpenMP loop is the timed CPU port)."""
    import torch

    from oracle.oracle import LF2DParams, Oracle
    from yoricktrtt_amd.projectors import LFProjector

    rng = np.random.default_rng(22)
    x = rng.uniform(size=(geo.ny, geo.nx))
    lf = LFProjector(prm, beta, np.zeros(beta.size), precision="f32")
    fms, bms, sino, vol = _gpu_pair(lf, torch.tensor(x, dtype=torch.float32, device="cuda"))
    ys = sino.double().cpu().numpy()
    back_gpu = vol.double().cpu().numpy()
    lf.close()
    o = Oracle()
    po = LF2DParams(*(getattr(prm, f) for f, _ in prm._fields_))
    t0 = time.perf_counter()
    fwd = o.lf2d_views(po, beta, np.zeros(beta.size), x, 0)
    t1 = time.perf_counter()
    back = o.lf2d_views(po, beta, np.zeros(beta.size), ys, 1)
    t2 = time.perf_counter()
    return {"config": 2, "workload": "Long & Fessler SF projector on configuration 2's geometry",
            "gpu": {"fwd_ms": round(fms, 3), "bwd_ms": round(bms, 3),
                    "pairs_per_s": round(1e3 / (fms + bms), 2), "dtype": "f32"},
            "cpu_reference": {"pairs_per_s": round(1.0 / (t2 - t0), 3), "wall_s": round(t2 - t0, 2),
                              "cores": int(os.environ.get("OMP_NUM_THREADS", threads)),
                              "kind": "port",
                              "sample": "every view, forward and back (oracle lf2d loop, FP64, OpenMP over views)"},
            "gpu_over_cpu": round((t2 - t0) * 1e3 / (fms + bms), 1),
            "error_vs_fp64": {"forward_max_rel": _rel(ys, fwd), "backward_max_rel": _rel(back_gpu, back),
                              "sample": "every view"}}


def _report_4d(ref, threads):
    """Configuration 4 (10 phases, 516 views): the GPU plan over every view;
    the reference on one view pair per host thread (forward of the view's
    phase, back projection of its row), scaled to 516 views; the GPU
    against those FP64 outputs (a plan over the sampled views for the back
    projection)."""
    import torch

    from yoricktrtt_amd.configs import _tuned, config4_setup
    from yoricktrtt_amd.projectors import Projector

    geo, phase, _ = config4_setup()
    nt = int(phase.max()) + 1
    one = np.ones(phase.size)
    rng = np.random.default_rng(4)
    x = rng.uniform(size=(nt, geo.nz, geo.ny, geo.nx))
    xd = torch.tensor(x, dtype=torch.float32, device="cuda")
    P = _tuned(Projector(geo, model="cubic", precision="f32", nt=nt, t0=phase, t1=phase, w0=one,
                         w1=np.zeros(phase.size)))
    fms, bms, sino, _ = _gpu_pair(P, xd)
    P.close()
    views = np.linspace(0, phase.size - 1, threads).astype(int)
    ys = sino[views].double().cpu().numpy()
    sino_gpu = ys.copy()
    fwd = np.zeros((views.size, geo.nv, geo.nu))
    bk = np.zeros((views.size, geo.nz, geo.ny, geo.nx))

    def pair(j):
        v = views[j]
        ref.CB3D_cubic(fwd[j], x[phase[v]], geo.nx, geo.ny, geo.nz, geo.s_scl, geo.nu, geo.nv,
                       geo.u_scl, geo.v_scl, geo.xos[v], geo.xsd[v], 0)
        ref.CB3D_cubic(bk[j], ys[j], geo.nx, geo.ny, geo.nz, geo.s_scl, geo.nu, geo.nv, geo.u_scl,
                       geo.v_scl, geo.xos[v], geo.xsd[v], 1)

    wall = _pool_time(pair, views.size, threads)
    back = np.zeros_like(x)
    for j, v in enumerate(views):
        back[phase[v]] += bk[j]
    sub = geo.subset(views)
    Ps = Projector(sub, model="cubic", precision="f32", nt=nt, t0=phase[views], t1=phase[views],
                   w0=np.ones(views.size), w1=np.zeros(views.size))
    vol = Ps.backward(torch.tensor(ys, dtype=torch.float32, device="cuda"))
    back_gpu = vol.double().cpu().numpy()
    Ps.close()
    rate = views.size / wall / phase.size
    return {"config": 4, "workload": "4-D cone, 10 phases x 128x128x64 x 2 mm, 516 views, 256x192 "
                                     "detector 2 mm, Rsc 1000, Rsd 1500",
            "gpu": {"fwd_ms": round(fms, 3), "bwd_ms": round(bms, 3),
                    "pairs_per_s": round(1e3 / (fms + bms), 2), "dtype": "f32"},
            "cpu_reference": {"pairs_per_s": round(rate, 4), "wall_s": round(wall, 2),
                              "cores": threads, "kind": "reference",
                              "sample": f"{views.size} threads x (1 forward + 1 back view of "
                                        "trtt_CB3D_cubic, FP64), scaled to 516 views"},
            "gpu_over_cpu": round(1e3 / (fms + bms) / rate, 1),
            "error_vs_fp64": {"forward_max_rel": _rel(sino_gpu, fwd),
                              "backward_max_rel": _rel(back_gpu, back),
                              "sample": f"the {views.size} timed views (forward of the full plan; "
                                        "back projection by a plan over those views)"}}


def configs_report(threads):
    """Configurations 1, 2 (and its Long & Fessler projector) and 4 beside
    the reference CPU path: GPU projection times, the reference timed on the
    host cores, and the GPU's max error against the FP64 results."""
    from oracle.oracle import REF_O3_SO, REF_SO, Reference
    from yoricktrtt_amd.configs import config1_setup, config2_setup

    path = REF_O3_SO if Reference.available(REF_O3_SO) else REF_SO
    ref = Reference(path)
    build = "O3_x86-64-v3" if path == REF_O3_SO else "O2"
    out = []
    g1 = config1_setup()
    out.append(_report_2d(1, g1, False, "f64", "2-D parallel 256^2, 360 views / 180 deg, 367 bins",
                          ref, threads))
    g2, prm, beta = config2_setup()
    out.append(_report_2d(2, g2, True, "f32", "2-D fan 512^2 x 0.5 mm, 720 views / 360 deg, 736 "
                          "flat bins 0.75 mm, Rsc 1000, Rsd 1500", ref, threads))
    out.append(_report_lf(g2, prm, beta, threads))
    out.append(_report_4d(ref, threads))
    for r in out:
        if r["cpu_reference"]["kind"] == "reference":
            r["cpu_reference"]["build"] = build
    return out


def table_model_rate(x, y, steps=3):
    """Throughput of the tabulated footprint (the model the Yorick scripts
    call, Ytrtt_3D_projectors.i:124-128; 10^5-entry table as
    Ytrtt3D_script_reconstructions.i builds it) on the same workload."""
    import torch

    from yoricktrtt_amd.projectors import Projector

    P = Projector(geometry(), model="table", precision="f32", size_footprint=100000)
    if P.path > 0:
        from yoricktrtt_amd.autotune import autotune

        autotune(P, cache=os.path.join(HERE, "yoricktrtt_amd", "tuning", "tuning.json"))
    path = P.path
    sino = torch.zeros((NVIEWS, NV, NU), device=x.device, dtype=torch.float32)
    vol = torch.zeros_like(x)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    P.forward(x, sino)
    P.backward(y, vol)
    fw, bw = [], []
    for _ in range(steps):
        sino.zero_()
        vol.zero_()
        ev[0].record()
        P.forward(x, sino)
        ev[1].record()
        P.backward(y, vol)
        ev[2].record()
        torch.cuda.synchronize()
        fw.append(ev[0].elapsed_time(ev[1]))
        bw.append(ev[1].elapsed_time(ev[2]))
    P.close()
    f, b = float(np.mean(fw)), float(np.mean(bw))
    kern = ("axial kernels, FP64 table indices, cubic CDF at the table nodes" if path > 0
            else "general footprint kernels (FP64 table coordinates, FP32 data)")
    return {"pairs_per_s": round(1e3 / (f + b), 3), "forward_ms": round(f, 2),
            "backward_ms": round(b, 2), "kernels": kern, "steps": steps}


def pcie_rates(P, x_host, y_host, views=4):
    """Rates with host arrays, PCIe transfers included (not the bench value):
    the per-view drop-in symbol trtt_CB3D_cubic as the unchanged Yorick loop
    calls it (FP64 host volume up and projection down per view), and the
    batched host-array evaluation trtt_hip_eval of the whole data term."""
    import ctypes as C

    from yoricktrtt_amd import _lib

    lib = _lib.load()
    geo = P.geom
    dp = _lib.dp
    x64 = np.ascontiguousarray(x_host, dtype=np.float64)
    y0 = np.ascontiguousarray(y_host[0], dtype=np.float64)
    out_f = np.zeros((NV, NU))
    out_b = np.zeros((NZ, NY, NX))

    def pair(v):
        xos, xsd = np.ascontiguousarray(geo.xos[v]), np.ascontiguousarray(geo.xsd[v])
        for out, inp, job in ((out_f, x64, 0), (out_b, y0, 1)):
            if lib.trtt_CB3D_cubic(out.ctypes.data_as(dp), inp.ctypes.data_as(dp), NX, NY, NZ, 1.0,
                                   3, NU, NV, 1.0, 1.0, xos.ctypes.data_as(dp),
                                   xsd.ctypes.data_as(dp), job) != 0:
                raise RuntimeError(_lib.last_error())

    pair(0)  # builds the cached one-view plan
    t0 = time.perf_counter()
    for v in range(1, views + 1):
        pair(v * (NVIEWS // (views + 1)))
    per_view_pair = (time.perf_counter() - t0) / views
    P.eval_set_data(y_host)
    P.eval_host(x64)
    t0 = time.perf_counter()
    reps = 2
    for _ in range(reps):
        P.eval_host(x64)
    t_eval = (time.perf_counter() - t0) / reps
    return {"dropin_pairs_per_s": round(1.0 / (per_view_pair * NVIEWS), 4),
            "dropin_sample": f"{views} views, forward + back each, FP64 host arrays, FP32 compute",
            "eval_host_pairs_per_s": round(1.0 / t_eval, 3),
            "eval_host_sample": f"{reps} evaluations of trtt_hip_eval (x up, gx down, FP64 host)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--model", default="cubic", choices=["cubic", "table"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU reference baseline")
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("BENCH_CPU_THREADS", len(os.sched_getaffinity(0)))))
    ap.add_argument("--no-check", action="store_true", help="skip the error statement vs FP64")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the table-model and host-array (PCIe) figures")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes, no autotuner")
    ap.add_argument("--configs-report", default=None, metavar="PATH",
                    help="after the benchmark line, write configurations 1, 2 (with Long & "
                         "Fessler) and 4 beside the reference CPU path to PATH (JSON lines)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run an elementwise copy of 1 GiB (HBM counter calibration)")
    ap.add_argument("--tune-cache", default=os.path.join(HERE, "yoricktrtt_amd", "tuning",
                                                         "tuning.json"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch

    from yoricktrtt_amd.projectors import Projector

    dev = torch.device("cuda:0")
    if args.copy_probe:
        # reads 1 GiB and writes 1 GiB: the byte counters of this kernel
        # calibrate FETCH_SIZE / WRITE_SIZE of the projector kernels
        a = torch.rand(1 << 28, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            torch.add(a, 0.0, out=b)
        torch.cuda.synchronize()
        del a, b
    geo = geometry()
    P = Projector(geo, model=args.model, precision="f32")
    print(f"bench: plan ready ({NVIEWS} views)", file=sys.stderr, flush=True)
    tuning = None
    if not args.no_tune and P.path > 0:
        from yoricktrtt_amd.autotune import autotune

        tr = autotune(P, cache=args.tune_cache,
                      log=lambda m: print("bench: tune " + m, file=sys.stderr, flush=True))
        tuning = dict(tr["tuning"], cached=tr["cached"])
        print(f"bench: tuning {tuning}", file=sys.stderr, flush=True)
    timed_kernels = P.path > 0
    if timed_kernels:
        P.set_timing(True)
    x = torch.from_numpy(phantom()).to(dev)
    g = torch.Generator(device=dev).manual_seed(7)
    y = torch.rand((NVIEWS, NV, NU), generator=g, device=dev, dtype=torch.float32)
    sino = torch.zeros((NVIEWS, NV, NU), device=dev, dtype=torch.float32)
    vol = torch.zeros((NZ, NY, NX), device=dev, dtype=torch.float32)

    def step(evs=None):
        sino.zero_()
        vol.zero_()
        if evs:
            evs[0].record()
        P.forward(x, sino)
        if evs:
            evs[1].record()
        P.backward(y, vol)
        if evs:
            evs[2].record()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print("bench: warmup done", file=sys.stderr, flush=True)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    kt = []  # (forward main kernel ms, launches, backward kernel ms) per step
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(ev[k])
        if timed_kernels:
            # the plan's own HIP events on the launch stream, around the
            # main forward kernel of each view chunk and the backward kernel
            fms, nl = P.kernel_time(0)
            bms, _ = P.kernel_time(1)
            kt.append((fms, nl, bms))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd_ms = np.mean([e[0].elapsed_time(e[1]) for e in ev])
    bwd_ms = np.mean([e[1].elapsed_time(e[2]) for e in ev])
    if P.error_flag() != 0:
        raise RuntimeError("a voxel fell in the source plane")

    ms_per_step = wall * 1e3 / args.steps
    pairs_per_s = args.steps / wall
    fwd_b, bwd_b, comp_b = algorithmic_bytes()
    if kt:
        kf = float(np.mean([t[0] for t in kt]))
        nlaunch = kt[0][1]
        kb = float(np.mean([t[2] for t in kt]))
    else:
        kf, nlaunch, kb = fwd_ms, 1, bwd_ms
    dom_is_bwd = kb >= kf
    # per launch: the forward runs once per chunk of views (bytes and time
    # both split by chunk), the backward once
    dom_ms = kb if dom_is_bwd else kf / nlaunch
    dom_bytes = bwd_b if dom_is_bwd else fwd_b / nlaunch
    achieved = dom_bytes / (dom_ms * 1e-3)
    result = {
        "metric": "cone 256^3 fwd+back pairs/s",
        "value": round(pairs_per_s, 4),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (30 seeded ellipsoids; uniform random sinogram)",
        "config": {"workload": "3D cone beam 256^3 x 1 mm, 360 views / 360 deg, 512x384 flat "
                               "detector 1 mm, Rsc 1000 mm, Rsd 1500 mm, cubic footprint",
                   "model": args.model},
        "roofline": {
            "bound": "hbm",
            "kernel": ("axial::bwd" if dom_is_bwd else "axial::fwd") if kt else
                      ("fp_backward" if dom_is_bwd else "fp_forward"),
            "launch_ms": round(dom_ms, 3),
            "bytes_per_launch": int(dom_bytes),
            "achieved": round(achieved / 1e9, 1),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4),
            "traffic": None,  # filled below from the rocprofv3 counters
            "bytes_model": "per-view-operator bytes (each view reads the volume; transpose "
                           "read-modify-writes it)",
        },
        "kernels_ms": {"forward_op": round(fwd_ms, 3), "backward_op": round(bwd_ms, 3),
                       "fwd_main": round(kf, 3), "fwd_launches": nlaunch,
                       "bwd_main": round(kb, 3)},
        "tuning": tuning,
        "pair_fracs": {
            "hbm_per_view_bytes": round((fwd_b + bwd_b) * pairs_per_s / HBM_PEAK, 4),
            "hbm_compulsory_bytes": round(comp_b * pairs_per_s / HBM_PEAK, 6),
            # 4 flops per footprint weight (both models use the same windows)
            "fp32_footprint_flops": round(4 * FOOTPRINT_WEIGHTS * pairs_per_s / FP32_PEAK, 4),
        },
    }
    result["roofline"]["traffic"] = measured_traffic(tuning, result["roofline"]["kernel"])
    # the axial kernels are bound by instruction issue, not by HBM
    result["roofline"]["issue"] = measured_issue(tuning, result["roofline"]["kernel"], dom_ms)
    x_host = x.cpu().numpy()
    y_host = y.cpu().numpy()
    if not args.no_check:
        print("bench: error against the FP64 oracle", file=sys.stderr, flush=True)
        result["error_vs_fp64"] = error_vs_fp64(geo, x_host, y_host, sino, vol)
    if not args.no_extra and args.model == "cubic":
        print("bench: table model", file=sys.stderr, flush=True)
        result["table_model"] = table_model_rate(x, y)
        print("bench: host-array rates", file=sys.stderr, flush=True)
        P.set_timing(False)
        result["pcie"] = pcie_rates(P, x_host, y_host)
    if not args.no_cpu:
        print("bench: timing the reference C plugin on the host", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_reference_baseline(geo, args.cpu_threads)
    else:
        result["cpu_baseline"] = None
    P.close()
    print(json.dumps(result), flush=True)
    if args.configs_report:
        print("bench: configurations 1, 2, 4 beside the reference", file=sys.stderr, flush=True)
        with open(args.configs_report, "w") as f:
            for r in configs_report(args.cpu_threads):
                print("bench: " + json.dumps(r), file=sys.stderr, flush=True)
                f.write(json.dumps(r) + "\n")


if __name__ == "__main__":
    main()
