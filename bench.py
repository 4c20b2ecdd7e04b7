"""Benchmark of the SALSA coagulation + condensation step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
    python bench.py --counter-run {mixed,fp64,copy} [--lanes 16,16,16] [--steps 3]

The second form is the program profiled by the hardware-counter runs: it
runs K steps at a fixed launch shape (no autotune, no timing, no CPU leg), or
a 512 MiB device-to-device copy whose FETCH_SIZE / WRITE_SIZE readings scale
the counters to bytes.

Workload: 64x64x120 cells (491,520) at the reference-legal bin layout
nbin = (3,4) -> aerosol 11 (1a 3, 2a 4, 2b 4), cloud 8, precipitation 7; SO4 + H2O;
dt = 1 s; seeded synthetic fields with a cloudy boundary-layer slab z in [60, 90).
A step is one run_SALSA-style pass (tendency mode) over every interior cell
with all inputs resident in HBM.  The FP32-mixed path is the headline value;
the plain FP64 path is timed on the same inputs for the speedup, and the two
are compared on every bin.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "uclales-salsa_amd"))

METRIC = "cell-steps/s SALSA coag+cond: FP32-mixed vs FP64 GPU at ≤1e-5 rel err"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def _cpu_worker(args):
    """Time one slice of the packed cells on one core (spawned process): the
    compiled reference (oracle/_ref harness, its own clock around the
    coagulation + condensation calls) or, where that is not built, the FP64 C
    restatement in oracle/."""
    shm_name, shapes, lo, hi, kind = args
    import numpy as np
    from multiprocessing import shared_memory
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    shm = shared_memory.SharedMemory(name=shm_name)
    off = 0
    arrs = []
    for shp in shapes:
        n = int(np.prod(shp))
        arrs.append(np.ndarray(shp, dtype=np.float64, buffer=shm.buf, offset=off))
        off += 8 * n
    L = O.make_layout()
    P = O.make_params(level=4)
    sl = slice(lo, hi)
    a = [np.array(x[sl], copy=True) for x in arrs]     # own copies: the shared block is unmapped next
    del arrs
    shm.close()
    if kind == "reference":
        import refio
        dt = refio.run(L, P, 4, *a)["compute_s"]
    else:
        t0 = time.perf_counter()
        O.step_batch(L, P, *a)
        dt = time.perf_counter() - t0
    return hi - lo, dt


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(desc, f, max_cores=16):
    """CPU leg on the host cores: one process per core over contiguous slices
    of the same 491,520 cells, one step each.  Throughput = cells / the
    slowest process's compute time (the processes run concurrently).  The
    process count is the CPU affinity of this process, capped at max_cores:
    a one-GPU share of the benchmark host is 16 cores, whatever the machine
    itself has (both counts are recorded)."""
    import multiprocessing as mp
    from multiprocessing import shared_memory

    import numpy as np
    from salsa_amd import workload as W
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refio
    kind = "reference" if refio.available() else "port"
    T, p, rs, rsi, rv, numc, volc = W.to_salsa_units(desc, f, level=4)
    n, nt0 = numc.shape
    ntot = nt0 + desc.nice                      # the oracle layout keeps the 7 inert ice bins
    nc = np.zeros((n, ntot))
    vc = np.zeros((n, ntot, 9))
    nc[:, :nt0] = numc
    vc[:, :nt0] = volc
    del numc, volc
    arrays = [T, p, rs, rsi, rv, nc, vc]
    size = sum(a.nbytes for a in arrays)
    shm = shared_memory.SharedMemory(create=True, size=size)
    off = 0
    shapes = []
    for a in arrays:
        np.ndarray(a.shape, dtype=np.float64, buffer=shm.buf, offset=off)[...] = a
        shapes.append(a.shape)
        off += a.nbytes
    del arrays, nc, vc
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    cores = min(max_cores, affinity)
    bounds = np.linspace(0, n, cores + 1).astype(int)
    ctx = mp.get_context("spawn")
    try:
        with ctx.Pool(cores) as pool:
            res = pool.map(_cpu_worker, [(shm.name, shapes, bounds[i], bounds[i + 1], kind) for i in range(cores)])
    finally:
        shm.close()
        shm.unlink()
    slowest = max(r[1] for r in res)
    per_core = sum(r[0] for r in res) / sum(r[1] for r in res)
    what = ("the compiled reference (oracle/_ref/ref_harness: coagulation + condensation of "
            "src/src_salsa built with amdflang -O2; its own clock, file I/O excluded)"
            if kind == "reference" else "the FP64 C restatement (oracle/salsa_oracle.c, gcc -O2)")
    return {"value": n / slowest, "unit": "cell-steps/s", "cores": cores, "kind": kind, "per_core": per_core,
            "sample": "all %d cells of the 64x64x120 workload, 1 step, %s, one process per core" % (n, what),
            "host_cpu_count": os.cpu_count(), "host_affinity": affinity, "cpu_model": _cpu_model()}


def counter_run(kind, lanes, steps):
    """Fixed-shape steps (or the copy calibration) for rocprofv3 --pmc runs."""
    import numpy as np
    import torch
    from salsa_amd import capi, workload as W
    dev = torch.device("cuda")
    if kind == "copy":
        a = torch.empty(512 << 20, dtype=torch.uint8, device=dev)
        a.fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(steps):
            b.copy_(a)
        torch.cuda.synchronize()
        print("counter run: copy of 512 MiB x %d done" % steps, flush=True)
        return
    desc = capi.bin_desc()
    f = W.config_fields(desc, 2)
    g = f.grid
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fd = dict(press=T(f.press), temp=T(f.temp), rsl=T(f.rsl), rsi=T(f.rsi), dn=T(f.dn), rv=T(f.rv))
    n_p = [T(x) for x in f.n[:3]]
    m_p = [T(x) for x in f.m[:3]]
    n_t = [torch.zeros_like(x) for x in n_p]
    m_t = [torch.zeros_like(x) for x in m_p]
    rv_t = torch.zeros_like(fd["rv"])
    prec = capi.SALSA_FP32_MIXED if kind == "mixed" else capi.SALSA_FP64
    ctx = capi.Context(desc, capi.params(level=4, precision=prec), g.nzp, g.nxp, g.nyp)
    for k in range(3):
        ctx.set_shape(k, lanes[k])
    for _ in range(steps):
        ctx.step_device(fd, n_p, m_p, n_t, m_t, rv_t)
    ctx.sync()
    ctx.close()
    print("counter run: %s, lanes %s, %d steps done" % (kind, lanes, steps), flush=True)


def gradient_leg(capi, W, desc, nsteps=20):
    """Forward-mode gradient of the precipitation volume over a 20-step
    trajectory of the 32x32x80 workload (81,920 cells, 30 % cloudy): dJ/dN0
    for every bin of every cell and dJ/dE, timed per precision path."""
    import numpy as np
    import torch
    f = W.config_fields(desc, 3)
    g = f.grid
    dev = torch.device("cuda")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fd = dict(press=T(f.press), temp=T(f.temp), rsl=T(f.rsl), rsi=T(f.rsi), dn=T(f.dn), rv=T(f.rv))
    n_p = [T(x) for x in f.n[:3]]
    m_p = [T(x) for x in f.m[:3]]
    djdn = [torch.zeros_like(x) for x in n_p]
    res = {"workload": "32x32x80 = %d cells, bins 11/8/7, %d steps, J = precipitation volume, "
                       "%d directions per cell (every bin + E)" % (g.ncells, nsteps, 27),
           "unit": "cell-steps/s"}
    for prec, name in ((capi.SALSA_FP64, "fp64"), (capi.SALSA_FP32_MIXED, "mixed")):
        ctx = capi.Context(desc, capi.params(level=4, precision=prec), g.nzp, g.nxp, g.nyp)
        ctx.gradient_device(fd, n_p, m_p, djdn, nsteps=1)        # warm-up: code load, buffers
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J, dJdE = ctx.gradient_device(fd, n_p, m_p, djdn, nsteps=nsteps)
        dt = time.perf_counter() - t0
        ctx.close()
        res[name] = {"value": g.ncells * nsteps / dt, "seconds": dt, "J": J, "dJdE": dJdE}
    return res


def adjoint_leg(capi, W, desc, nsteps=20, every=5):
    """Reverse-mode gradient (salsa_adjoint_device) of the same objective on
    the same workload: checkpointed forward pass (every 5 steps), the stretch
    between checkpoints recomputed, one adjoint sweep; timed per precision
    path around the whole call (inputs resident in HBM, result totals read
    back)."""
    import numpy as np
    import torch
    f = W.config_fields(desc, 3)
    g = f.grid
    dev = torch.device("cuda")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fd = dict(press=T(f.press), temp=T(f.temp), rsl=T(f.rsl), rsi=T(f.rsi), dn=T(f.dn), rv=T(f.rv))
    n_p = [T(x) for x in f.n[:3]]
    m_p = [T(x) for x in f.m[:3]]
    djdn = [torch.zeros_like(x) for x in n_p]
    res = {"workload": "32x32x80 = %d cells, bins 11/8/7, %d steps, checkpoints every %d steps, "
                       "J = precipitation volume, dJ/dN0 of every bin and dJ/dE" % (g.ncells, nsteps, every),
           "unit": "cell-steps/s"}
    for prec, name in ((capi.SALSA_FP64, "fp64"), (capi.SALSA_FP32_MIXED, "mixed")):
        ctx = capi.Context(desc, capi.params(level=4, precision=prec), g.nzp, g.nxp, g.nyp)
        ctx.adjoint_device(fd, n_p, m_p, djdn, nsteps=nsteps, checkpoint_every=every)    # warm-up, buffers
        best = None
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            J, dJdE = ctx.adjoint_device(fd, n_p, m_p, djdn, nsteps=nsteps, checkpoint_every=every)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        ctx.close()
        res[name] = {"value": g.ncells * nsteps / best, "seconds": best, "J": J, "dJdE": dJdE}
    return res


def config_leg(capi, W, cfg, nsteps=10, gradient=False):
    """The two 128x128x100 configurations (1,638,400 cells), FP32-mixed,
    `nsteps` in-place steps after the autotuner and one warm-up step: config 4
    (level 4, 15 % cloudy cells with every cloud and precipitation bin
    active, the rest aerosol-only: the cost-class compaction at work) and
    config 5 (level 5, SO4 + dust, ice bins with rime, T 240-275 K).  Timed
    around the steps (wall clock, inputs resident in HBM) and on the GPU (HIP
    events around each step's launches).  With `gradient`, config 5 also runs
    the forward-mode derivative over the same 10 steps (level 5 has no
    reverse sweep)."""
    import numpy as np
    import torch
    level = 5 if cfg == 5 else 4
    desc = capi.bin_desc(species=("SO4", "DU") if cfg == 5 else ("SO4",))
    t0 = time.perf_counter()
    f = W.config_fields(desc, cfg)
    gen = time.perf_counter() - t0
    g = f.grid
    ncat = 4 if level == 5 else 3
    dev = torch.device("cuda")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fd = dict(press=T(f.press), temp=T(f.temp), rsl=T(f.rsl), rsi=T(f.rsi), dn=T(f.dn), rv=T(f.rv))
    n_p = [T(x) for x in f.n[:ncat]]
    m_p = [T(x) for x in f.m[:ncat]]
    ctx = capi.Context(desc, capi.params(level=level, precision=capi.SALSA_FP32_MIXED), g.nzp, g.nxp, g.nyp)
    ctx.autotune(fd, n_p, m_p)
    st = dict(fd)
    st["rv"] = fd["rv"].clone()
    nn = [x.clone() for x in n_p]
    mm = [x.clone() for x in m_p]
    ctx.step_device(st, nn, mm, flags=capi.STEP_INPLACE)       # warm-up on a copy
    st["rv"].copy_(fd["rv"])
    for a, b in zip(nn + mm, n_p + m_p):
        a.copy_(b)
    ctx.sync()
    torch.cuda.synchronize()
    ctx.timing(True)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        ctx.step_device(st, nn, mm, flags=capi.STEP_INPLACE)
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms, ksteps = ctx.kernel_time()
    ms, cells, n = ctx.phase_time()
    ctx.timing(False)
    shapes = [ctx.shape_of(k) for k in range(3)]
    res = {"workload": "config %d: 128x128x100 = %d cells, level %d, bins %d/%d/%d%s, FP32-mixed, %d in-place steps"
                       % (cfg, g.ncells, level, desc.nbins, desc.ncld, desc.nprc,
                          "/%d ice, SO4+DU" % desc.nice if level == 5 else ", SO4", nsteps),
           "value": g.ncells * nsteps / wall, "unit": "cell-steps/s", "ms_per_step": wall / nsteps * 1e3,
           "gpu_ms_per_step": kms / max(ksteps, 1), "cells_per_class": [c / max(n, 1) for c in cells],
           "launch_lanes_per_cell": shapes, "field_generation_s": gen}
    if level == 5:
        res["sedimentation"] = sedim5_leg(capi, desc, g, fd, nn, mm)
    if gradient:
        djdn = [torch.zeros_like(x) for x in n_p]
        ctx.gradient_device(fd, n_p, m_p, djdn, nsteps=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J, dJdE = ctx.gradient_device(fd, n_p, m_p, djdn, nsteps=nsteps)
        dt = time.perf_counter() - t0
        fl = ctx.cell_flags()
        res["gradient"] = {"value": g.ncells * nsteps / dt, "seconds": dt, "J": J, "dJdE": dJdE,
                           "unstable_cells": int(np.count_nonzero(fl & capi.CELL_GRAD_UNSTABLE))}
    ctx.close()
    return res


def sedim5_leg(capi, desc, g, fd, n_p, m_p, calls=20):
    """Level-5 sedimentation on device (salsa_sedim_device: aerosol, cloud,
    precipitation and ice, DepositionFast flag 4 for the ice) of the config-5
    state after its steps, on the 128x128x100 grid, timed with HIP events on
    the context stream.  Algorithmic bytes per call: every element of the
    four categories read with its tendency read and written (24 B), fall
    speeds written (8 B per precipitation and ice bin and cell), per cell tk,
    th, adn read, tlt read and written, rrate and irate written (56 B)."""
    import numpy as np
    import torch
    n3 = g.nzp * g.nxp * g.nyp
    dev = fd["temp"].device
    nt = [torch.zeros_like(x) for x in n_p]
    mt = [torch.zeros_like(x) for x in m_p]
    rng = np.random.default_rng(0)
    sf = dict(tk=fd["temp"], th=fd["temp"] * 1.02, adn=fd["dn"],
              ustar=torch.from_numpy(0.1 + 0.4 * rng.random(g.nxp * g.nyp)).to(dev))
    z = lambda *sh: torch.zeros(sh, dtype=torch.float64, device=dev)
    tlt, rr, sr, ir, si = z(n3), z(n3), z(g.nxp * g.nyp), z(n3), z(g.nxp * g.nyp)
    vt, vi = z(desc.nprc, n3), z(desc.nice, n3)
    dzt = 1. / np.full(g.nzp, 20.)
    ctx = capi.Context(desc, capi.params(level=5), g.nzp, g.nxp, g.nyp)
    call = lambda: ctx.sedim_device(1.0, dzt, sf, n_p, m_p, nt, mt, tlt, rr, sr, vt, sed=(1, 1, 1, 1), irate=ir,
                                    sfcirate=si, vt_ice=vi)
    call()
    ctx.sync()
    stream = torch.cuda.ExternalStream(ctx.stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(calls):
        call()
    e1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / calls
    ctx.close()
    elems = sum(x.numel() for x in n_p + m_p)
    nbytes = 24 * elems + 8 * (desc.nprc + desc.nice) * n3 + 56 * n3
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"ms_per_call": ms, "cells": n3, "bytes_per_call": nbytes, "achieved_GBs": gbs,
            "peak_GBs": HBM_PEAK_GBS, "frac": gbs / HBM_PEAK_GBS, "bound": "hbm",
            "max_ice_fall_speed_m_s": float(vi.max().item()), "ice_cells": int((n_p[3].sum(dim=0) > 1e-6).sum().item())}


def les_leg(capi, desc, f, fd, n_p, m_p, calls=20):
    """The LES side after a step on device (salsa_les_update_device:
    tend_constrain2, update_sclrs, tend0, SALSA_diagnostics) on the same
    fields: one step's tendencies, then `calls` updates timed with HIP events
    on the context stream.  Algorithmic bytes per call: every tracer element
    read and written with its tendency in the update pass (32 B) and read in
    the diagnostics pass (8 B); per cell the temperature, humidity and vapour
    read (24 B) and four diagnostic sums written (32 B)."""
    import numpy as np
    import torch
    g = f.grid
    n3 = g.nzp * g.nxp * g.nyp
    np_ = [x.clone() for x in n_p]
    mp_ = [x.clone() for x in m_p]
    nt = [torch.zeros_like(x) for x in np_]
    mt = [torch.zeros_like(x) for x in mp_]
    rvt = torch.zeros_like(fd["rv"])
    ctx = capi.Context(desc, capi.params(level=4), g.nzp, g.nxp, g.nyp)
    ctx.step_device(fd, np_, mp_, nt, mt, rvt)
    lf = dict(temp=fd["temp"], rh=fd["rv"] / fd["rsl"], rp=fd["rv"])
    diag = torch.zeros((6, n3), dtype=torch.float64, device=fd["temp"].device)
    dzt = 1. / np.full(g.nzp, 20.)
    ctx.les_update_device(1.0, dzt, lf, np_, nt, mp_, mt, diag)
    ctx.sync()
    stream = torch.cuda.ExternalStream(ctx.stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(calls):
        ctx.les_update_device(1.0, dzt, lf, np_, nt, mp_, mt, diag)
    e1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / calls
    ctx.close()
    elems = sum(x.numel() for x in np_ + mp_)
    nbytes = 40 * elems + 56 * n3
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"ms_per_call": ms, "cells": n3, "tracer_elements": elems, "bytes_per_call": nbytes,
            "achieved_GBs": gbs, "peak_GBs": HBM_PEAK_GBS, "frac": gbs / HBM_PEAK_GBS, "bound": "hbm"}


def sedim_leg(capi, desc, f, fd, n_p, m_p, calls=20):
    """Level-4 sedimentation on device (salsa_sedim_device) on the same grid,
    with 10 drops/kg per precipitation bin at its mid diameter in every
    cell (the workload carries no precipitation), timed with HIP events on
    the context stream.  Algorithmic bytes per call: every aerosol, cloud and
    precipitation element read with its tendency read and written (24 B),
    fall speeds written (8 B per precipitation bin and cell), per cell tk,
    th, adn read, tlt read and written, rrate written (48 B)."""
    import numpy as np
    import torch
    g = f.grid
    n3 = g.nzp * g.nxp * g.nyp
    dev = fd["temp"].device
    np_ = [x.clone() for x in n_p]
    mp_ = [x.clone() for x in m_p]
    nb = desc.nprc
    dmid = np.array(desc.dmid[desc.nbins + desc.ncld:desc.nbins + desc.ncld + nb])
    np_[2][:] = 10.
    mp_[2][:] = 0.
    for b in range(nb):
        mp_[2][desc.iwa * nb + b] = 10. * 1000. * 0.5235988 * dmid[b] ** 3
        mp_[2][b] = 1e-3 * mp_[2][desc.iwa * nb + b]
    nt = [torch.zeros_like(x) for x in np_]
    mt = [torch.zeros_like(x) for x in mp_]
    rng = np.random.default_rng(0)
    sf = dict(tk=fd["temp"], th=fd["temp"] * 1.02, adn=fd["dn"],
              ustar=torch.from_numpy(0.1 + 0.4 * rng.random(g.nxp * g.nyp)).to(dev))
    tlt = torch.zeros(n3, dtype=torch.float64, device=dev)
    rr = torch.zeros(n3, dtype=torch.float64, device=dev)
    sr = torch.zeros(g.nxp * g.nyp, dtype=torch.float64, device=dev)
    vt = torch.zeros((nb, n3), dtype=torch.float64, device=dev)
    dzt = 1. / np.full(g.nzp, 20.)
    ctx = capi.Context(desc, capi.params(level=4), g.nzp, g.nxp, g.nyp)
    ctx.sedim_device(1.0, dzt, sf, np_, mp_, nt, mt, tlt, rr, sr, vt)
    ctx.sync()
    stream = torch.cuda.ExternalStream(ctx.stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(calls):
        ctx.sedim_device(1.0, dzt, sf, np_, mp_, nt, mt, tlt, rr, sr, vt)
    e1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / calls
    ctx.close()
    slow = sum(x.numel() for x in np_[:2] + mp_[:2])
    prec = np_[2].numel() + mp_[2].numel()
    nbytes = 24 * slow + 24 * prec + 8 * nb * n3 + 48 * n3
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"ms_per_call": ms, "cells": n3, "bytes_per_call": nbytes, "achieved_GBs": gbs, "peak_GBs": HBM_PEAK_GBS,
            "frac": gbs / HBM_PEAK_GBS, "bound": "hbm",
            "max_fall_speed_m_s": float(vt.max().item())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-gradient", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 128x128x100 legs (configs 4 and 5)")
    ap.add_argument("--config5-gradient", action="store_true",
                    help="also the forward-mode derivative at config 5 full size (about a minute)")
    ap.add_argument("--no-les", action="store_true")
    ap.add_argument("--counter-run", choices=("mixed", "fp64", "copy"))
    ap.add_argument("--lanes", default="16,16,16")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the SALSA step has no cross-GPU work "
              "(cells are independent; a multi-GPU run would shard the LES column blocks). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.counter_run:
        counter_run(args.counter_run, [int(x) for x in args.lanes.split(",")], args.steps)
        return

    import numpy as np
    from salsa_amd import capi, opcount, workload as W

    desc = capi.bin_desc()
    t0 = time.perf_counter()
    f = W.config_fields(desc, 2)
    g = f.grid
    gen_s = time.perf_counter() - t0

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(desc, f)

    import torch
    dev = torch.device("cuda")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fd = dict(press=T(f.press), temp=T(f.temp), rsl=T(f.rsl), rsi=T(f.rsi), dn=T(f.dn), rv=T(f.rv))
    n_p = [T(x) for x in f.n[:3]]
    m_p = [T(x) for x in f.m[:3]]
    n_t = [torch.zeros_like(x) for x in n_p]
    m_t = [torch.zeros_like(x) for x in m_p]
    rv_t = torch.zeros_like(fd["rv"])
    torch.cuda.synchronize()

    def run(prec):
        ctx = capi.Context(desc, capi.params(level=4, precision=prec), g.nzp, g.nxp, g.nyp)
        ctx.autotune(fd, n_p, m_p)
        for _ in range(args.warmup):
            ctx.step_device(fd, n_p, m_p, n_t, m_t, rv_t)
        ctx.sync()
        torch.cuda.synchronize()
        ctx.timing(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.step_device(fd, n_p, m_p, n_t, m_t, rv_t)
        ctx.sync()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kms, ksteps = ctx.kernel_time()
        phases = ctx.phase_time()
        ctx.timing(False)
        shapes = [ctx.shape_of(k) for k in range(3)]
        # one in-place step for the accuracy comparison
        rv = fd["rv"].clone()
        fdd = dict(fd)
        fdd["rv"] = rv
        np_ = [x.clone() for x in n_p]
        mp_ = [x.clone() for x in m_p]
        ctx.step_device(fdd, np_, mp_, flags=capi.STEP_INPLACE)
        ctx.sync()
        ctx.close()
        st = (rv.cpu().numpy(), [x.cpu().numpy() for x in np_], [x.cpu().numpy() for x in mp_])
        return wall / args.steps, kms / max(ksteps, 1), shapes, st, phases

    step_mixed, kern_mixed, shapes_mixed, st_mixed, ph_mixed = run(capi.SALSA_FP32_MIXED)
    res64 = None if args.no_fp64 else run(capi.SALSA_FP64)

    ncell = g.ncells
    value = ncell / step_mixed
    flops = opcount.flops_per_cell_step(desc, 4) * ncell
    bytes_ = opcount.bytes_per_cell_step(desc, 4, tendency=True) * ncell
    fpc = opcount.flops_per_cell_step(desc, 4)
    step_tf = flops / (kern_mixed * 1e-3) / 1e12

    def kernels(ph):
        """Per launch: the class kernels' average duration and cells, and
        the dominant one (largest GPU time per step)."""
        ms, cells, n = ph
        n = max(n, 1)
        per = {"gather_classify_ms": ms[0] / n, "scatter_ms": ms[4] / n}
        cls = []
        for k in range(3):
            c = cells[k] / n
            t = ms[1 + k] / n
            per["class%d" % k] = {"ms": t, "cells": c,
                                  "TFLOPs": fpc * c / (t * 1e-3) / 1e12 if t > 0 and c > 0 else 0.}
            cls.append((t, k))
        dom = max(cls)[1]
        return per, dom

    per_mixed, dom = kernels(ph_mixed)
    d = per_mixed["class%d" % dom]
    achieved_tf = d["TFLOPs"]
    prof_traffic = None
    prof = os.path.join(ROOT, "profiles", "traffic_k_salsa_step_64x64x120.json")
    if os.path.exists(prof):
        with open(prof) as fh:
            tr = json.load(fh)
        prof_traffic = tr.get("dominant_kernel_bytes_per_launch")
    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": step_mixed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "SALSA coagulation+condensation step, 64x64x120 = %d cells, bins 11/8/7 "
                               "(nbin=(3,4)), SO4+H2O, dt=1 s, tendency mode, cloudy slab z in [60,90)" % ncell,
                   "cells": ncell, "layout": "11/8/7", "precision": "FP32-mixed",
                   "launch_lanes_per_cell": shapes_mixed},
        "roofline": {"bound": "valu", "engine": "VALU (no MFMA: not a dense contraction)",
                     "achieved": achieved_tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tf / FP32_PEAK_TFLOPS, "traffic": prof_traffic,
                     "kernel": "k_salsa_step, cost class %d (%s), %d lanes per cell"
                               % (dom, ("clear, RH <= 0.98", "clear, RH > 0.98", "cloudy")[dom],
                                  shapes_mixed[dom]),
                     "kernel_ms_per_launch": d["ms"], "cells_per_launch": d["cells"],
                     "flop_per_cell_step": fpc,
                     "step_achieved": step_tf, "step_frac": step_tf / FP32_PEAK_TFLOPS,
                     "gpu_ms_per_step": kern_mixed, "launches": per_mixed,
                     "hbm_GBs_algorithmic": bytes_ / (kern_mixed * 1e-3) / 1e9,
                     "hbm_frac": bytes_ / (kern_mixed * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample",
                                                                      "host_cpu_count", "host_affinity",
                                                                      "cpu_model")},
    }
    if cpu is not None:
        out["cpu_baseline"]["per_core"] = cpu["per_core"]
        out["speedup_vs_cpu"] = value / cpu["value"]
    if res64 is not None:
        step64, kern64, shapes64, st64, ph64 = res64
        per64, dom64 = kernels(ph64)
        out["fp64"] = {"value": ncell / step64, "ms_per_step": step64 * 1e3, "gpu_ms_per_step": kern64,
                       "launch_lanes_per_cell": shapes64, "launches": per64,
                       "dominant_kernel_frac": per64["class%d" % dom64]["TFLOPs"] / 78.6,
                       "step_frac": flops / (kern64 * 1e-3) / 1e12 / 78.6, "peak_TFLOPs": 78.6}
        out["speedup_vs_fp64"] = step64 / step_mixed
        # accuracy of the mixed path against the FP64 path, every bin of every cell
        fo = lambda st: W.to_salsa_units(desc, W.Fields(grid=g, level=4, press=f.press, temp=f.temp, rsl=f.rsl,
                                                         rsi=f.rsi, dn=f.dn, rv=st[0], n=st[1] + [None],
                                                         m=st[2] + [None]))
        a, b = fo(st_mixed), fo(st64)
        rel = lambda x, y, fl: np.abs(x - y) / np.maximum(np.abs(y), fl)
        e = {"rv": rel(a[4], b[4], 1e-12), "N": rel(a[5], b[5], 1e-3),
             "V_SO4": rel(a[6][:, :, 0], b[6][:, :, 0], 1e-24), "V_H2O": rel(a[6][:, :, 1], b[6][:, :, 1], 1e-24)}
        out["accuracy_vs_fp64"] = {k: {"max": float(v.max()), "frac_gt_1e-5": float(np.mean(v > 1e-5))}
                                   for k, v in e.items()}
    if not args.no_les:
        out["les_update"] = les_leg(capi, desc, f, fd, n_p, m_p)
        out["sedimentation"] = sedim_leg(capi, desc, f, fd, n_p, m_p)
    if not args.no_gradient:
        out["gradient"] = gradient_leg(capi, W, desc)
        out["adjoint"] = adjoint_leg(capi, W, desc)
        for name in ("fp64", "mixed"):
            out["adjoint"][name]["vs_forward_tangents"] = (out["adjoint"][name]["value"] /
                                                           out["gradient"][name]["value"])
    if not args.no_configs:
        out["config4"] = config_leg(capi, W, 4)
        out["config5"] = config_leg(capi, W, 5, gradient=args.config5_gradient)
    out["setup_s"] = gen_s
    print(json.dumps(out))


if __name__ == "__main__":
    main()
