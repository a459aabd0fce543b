#!/usr/bin/env python3
"""Benchmark: ijdDS neighbour list of 1,048,576 fcc atoms at rc = 5 A on one GPU.

A step is one pass of the hot path over the workload with positions already
resident in HBM: bin -> count -> scan -> fill of i, j, d, D, S (float64, int32
indices and shifts) into pre-allocated outputs (steady state, as an MD loop
runs it). The K timed steps are bracketed by torch.cuda.synchronize() and HIP
events on the stream the kernels run on; each fill launch is bracketed by its
own events for the kernel-level roofline.

Prints one JSON line. `value` is pairs/s of the whole pipeline; the
`roofline` object is for the dominant kernel (the fill); `cpu_baseline` is the
reference's OpenMP CPU path (compiled from its sources into oracle/_ref) on
the host cores, on the same input.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "matscipy-neighbours_amd"))
sys.path.insert(0, ROOT)

METRIC = "pairs/s at 1M atoms rc=5Å (%HBM peak); batched systems/s vs per-system"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_fill_fcc1M.json")
KERNEL_SOURCE = os.path.join(ROOT, "matscipy-neighbours_amd", "csrc", "mnl_kernels.hip")
FILL_PREFIX = "k_fill<1, 0, false, "


def _fill_kernel(chunks):
    """The float64 uniform-cutoff fill kernel of a given window size."""
    return f"{FILL_PREFIX}{chunks}>"


def _autotune(lib, _lib, plan, sp):
    """The engine's autotuner on the plan's last count: the launch shape it
    keeps for this configuration class."""
    import ctypes
    split, chunks, ms = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_float(0)
    _lib.check(lib.mnl_plan_autotune(plan.handle, sp, ctypes.byref(split), ctypes.byref(chunks),
                                     ctypes.byref(ms)))
    return {"split": split.value, "chunks": chunks.value, "tuned_count_fill_ms": ms.value,
            "source": "autotune"}


def _source_sha():
    import hashlib
    with open(KERNEL_SOURCE, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def _counter_means(path):
    """Mean counter value (KiB) per kernel name of a rocprofv3 --pmc CSV."""
    import csv
    import re
    agg = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            k = re.sub(r"\(.*", "", r["Kernel_Name"].replace("mnl::(anonymous namespace)::", "")
                       .replace("void ", ""))
            agg.setdefault(k, []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in agg.items()}


def traffic_summary(fetch_csv, write_csv, cfetch_csv, cwrite_csv, commands):
    """HBM bytes per fill launch from FETCH_SIZE / WRITE_SIZE runs, each scaled
    by what the same counter reads for a plain 1 GiB copy (--calibrate-copy).
    The summary keeps one entry per fill kernel (window size) measured on the
    current kernel source; entries of an older source are dropped."""
    copy_bytes = float(1 << 30)
    ff = copy_bytes / (_counter_means(cfetch_csv)["__amd_rocclr_copyBuffer"] * 1024)
    wf = copy_bytes / (_counter_means(cwrite_csv)["__amd_rocclr_copyBuffer"] * 1024)
    fetches = _counter_means(fetch_csv)
    kernel = next(k for k in fetches if k.startswith(FILL_PREFIX))
    fetch = fetches[kernel] * 1024
    write = _counter_means(write_csv)[kernel] * 1024
    entry = {
        "fetch_size_bytes_raw": fetch, "write_size_bytes_raw": write,
        "copy_calibration": {"fetch_factor": ff, "write_factor": wf,
                             "copy_bytes_read": copy_bytes, "copy_bytes_written": copy_bytes},
        "bytes_read_per_launch": fetch * ff, "bytes_written_per_launch": write * wf,
        "bytes_per_launch": fetch * ff + write * wf,
    }
    out = {"workload": "fcc 64^3 cells (1048576 atoms), rc 5.0 A, ijdDS f64",
           "kernel_source_sha256": _source_sha(), "kernels": {}, "commands": commands}
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            old = json.load(f)
        if old.get("kernel_source_sha256") == out["kernel_source_sha256"]:
            out["kernels"].update(old.get("kernels", {}))
    out["kernels"][kernel] = entry
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


def copy_rate(torch, stream, reps=5):
    """Measured device-to-device copy rate (GB/s, bytes read plus bytes
    written) of a 1 GiB buffer: the best of a few copies, HIP events on the
    stream the copies run on."""
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    best = float("inf")
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        b.copy_(a)
        e1.record(stream)
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del a, b
    return 2.0 * (1 << 30) / (best * 1e-3) / 1e9


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """The host the CPU baseline ran on: CPU model, logical CPUs, and the CPUs
    this process may use (its affinity mask)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return dict(cpu_model=model, nproc=os.cpu_count() or 1, usable_cpus=usable)


def cpu_baseline(w, cell, inv, quantities, reps, single_reps=3):
    """Time the reference CPU core, compiled from its sources into
    oracle/_ref, on the same input: `reps` calls after one warm-up with
    OMP_NUM_THREADS threads (the usable CPUs when unset), then `single_reps`
    calls after one warm-up on one thread. Without the compiled reference
    the oracle's restatement is timed instead, and the result says so
    (kind "port", reference_missing) besides a warning on stderr."""
    import numpy as np

    import oracle
    host = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", host["usable_cpus"]))
    args = (quantities, np.zeros(3), cell, inv, w["pbc"], w["positions"], w["cutoff"])
    extra = {}
    if oracle.ref_available():
        secs, npairs, threads = oracle.ref_time(*args, reps + 1, threads=threads)
        secs1, _, _ = oracle.ref_time(*args, single_reps + 1, threads=1)
        oracle.ref_time(*args, 1, threads=threads)  # back to the parallel setting
        kind = "reference"
        what = "reference CPU core compiled from its sources (OpenMP, -O3)"
    else:
        _log("WARNING: oracle/_ref/libref_neighbours.so is missing: the CPU baseline times "
             "the oracle restatement, not the compiled reference")
        def run(nrep):
            t = []
            out = None
            for _ in range(nrep):
                t0 = time.perf_counter()
                out = oracle.neighbour_list(*args)
                t.append(time.perf_counter() - t0)
            return np.array(t), len(out["i"])
        secs, npairs = run(reps + 1)
        secs1 = None
        kind = "port"
        what = "oracle restatement of the reference CPU path (OpenMP, -O3)"
        extra["reference_missing"] = "oracle/_ref/libref_neighbours.so not built in this tree"
    med = float(np.median(secs[1:]))
    res = dict(value=npairs / med, unit="pairs/s", cores=int(threads), kind=kind,
               sample=f"full workload, {reps} timed calls after 1 warm-up on {threads} "
                      f"threads, median {med * 1e3:.1f} ms/call; {what}", **host, **extra)
    if secs1 is not None:
        med1 = float(np.median(secs1[1:]))
        res["single_thread"] = dict(value=npairs / med1, unit="pairs/s", ms_per_call=med1 * 1e3,
                                    calls=single_reps)
    return res


def batched_leg(lib, _lib, torch, dev, steps, warmup, tune=True):
    """1024 uneven periodic systems (8-512 atoms, 0.05 atoms/A^3, tilted
    cells, rc 5 A): one captured graph replay per batch versus one plain
    single-system build (count + fill) per system, outputs compared."""
    import ctypes

    import numpy as np

    from matscipy_neighbours_amd import workloads
    from matscipy_neighbours_amd.neighbours import _batch_geometry

    systems = workloads.uneven_batch()
    pos = np.ascontiguousarray(np.concatenate([s["positions"] for s in systems]))
    cells = np.stack([s["cell"] for s in systems])
    nat = np.array([len(s["positions"]) for s in systems])
    nsys, offs, co, ce, inv, pb, rcs = _batch_geometry(cells, True, 5.0, nat, None)
    P_ = _lib.ptr
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    pos_t = torch.from_numpy(pos).to(dev)
    plan = _lib.Plan(0)
    _lib.check(lib.mnl_plan_set_batch(plan.handle, sp, nsys, P_(offs), P_(co), P_(ce),
                                      P_(inv), P_(pb), P_(rcs)))
    n0 = ctypes.c_int64(0)
    _lib.check(lib.mnl_plan_batch_count(plan.handle, sp, P_(pos_t), ctypes.byref(n0)))
    P = int(n0.value)
    shape = _autotune(lib, _lib, plan, sp) if tune else {"source": "built-in"}

    def bufs():
        return dict(i=torch.empty(P, dtype=torch.int32, device=dev),
                    j=torch.empty(P, dtype=torch.int32, device=dev),
                    D=torch.empty((P, 3), dtype=torch.float64, device=dev),
                    d=torch.empty(P, dtype=torch.float64, device=dev),
                    S=torch.empty((P, 3), dtype=torch.int32, device=dev))

    b = bufs()
    sys_off = torch.empty(nsys + 1, dtype=torch.int64, device=dev)
    np_dev = torch.zeros(1, dtype=torch.int64, device=dev)
    g = ctypes.c_void_p()
    _lib.check(lib.mnl_plan_batch_graph_create(plan.handle, P_(pos_t), P, P_(b["i"]),
                                               P_(b["j"]), P_(b["D"]), P_(b["d"]), P_(b["S"]),
                                               0, P_(sys_off), P_(np_dev), ctypes.byref(g)))
    try:
        for _ in range(warmup):
            _lib.check(lib.mnl_graph_launch(g, sp))
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            _lib.check(lib.mnl_graph_launch(g, sp))
        e1.record(stream)
        torch.cuda.synchronize()
        t_batch = e0.elapsed_time(e1) * 1e-3 / steps
    finally:
        lib.mnl_graph_destroy(g)
    assert int(np_dev.item()) == P

    # per-system plain builds into slices of one buffer set
    ref = bufs()
    soff = sys_off.cpu().numpy()
    sp_plan = _lib.Plan(0)
    zero3 = np.zeros(3)

    def per_system():
        npairs = ctypes.c_int64(0)
        for k in range(nsys):
            a0, a1 = int(offs[k]), int(offs[k + 1])
            p0 = int(soff[k])
            _lib.check(lib.mnl_plan_count(
                sp_plan.handle, sp, P_(zero3), P_(ce[k]), P_(inv[k]), P_(pb[k]), a1 - a0,
                ctypes.c_void_p(pos_t.data_ptr() + 24 * a0), float(rcs[k]), None, None, 0,
                None, -1, ctypes.byref(npairs)))
            _lib.check(lib.mnl_plan_fill(
                sp_plan.handle, sp, ctypes.c_void_p(ref["i"].data_ptr() + 4 * p0),
                ctypes.c_void_p(ref["j"].data_ptr() + 4 * p0),
                ctypes.c_void_p(ref["D"].data_ptr() + 24 * p0),
                ctypes.c_void_p(ref["d"].data_ptr() + 8 * p0),
                ctypes.c_void_p(ref["S"].data_ptr() + 12 * p0)))
        torch.cuda.synchronize()

    per_system()  # warm-up (sizes the plan's workspaces)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        per_system()
    t_single = (time.perf_counter() - t0) / reps
    identical = all(torch.equal(b[q], ref[q]) for q in "ijDdS")
    return {
        "workload": "1024 periodic systems, 8-512 atoms (uniform), cubic cells at 0.05 "
                    "atoms/A^3 with off-diagonal tilt up to 10%, uniform positions (seed 4), "
                    f"rc 5.0 A, ijdDS f64; {len(pos)} atoms, {P} pairs",
        "systems": nsys,
        "atoms": int(len(pos)),
        "pairs": P,
        "batched_ms": t_batch * 1e3,
        "batched_systems_per_s": nsys / t_batch,
        "per_system_ms": t_single * 1e3,
        "per_system_systems_per_s": nsys / t_single,
        "speedup": t_single / t_batch,
        "identical_pairs": bool(identical),
        "launch_shape": shape,
        "how": "batched: one HIP graph replay (count + fill + per-system offsets) timed "
               "with HIP events over the replays; per-system: mnl_plan_count (waits for the "
               "pair total) + mnl_plan_fill per system into pre-allocated slices, wall clock",
    }


def _plan_step_fn(lib, _lib, torch, dev, w, f32=False, flags="ijdDS", backend=-1):
    """A steady-state build of one workload: count (async) + fill into
    pre-allocated outputs, as bench steps run it (backend -1: the built-in
    dense/hashed rule, 0 dense, 1 hashed). Returns (step, npairs)."""
    import ctypes

    import numpy as np

    from matscipy_neighbours_amd.neighbours import _parse_cutoff
    P_ = _lib.ptr
    N = len(w["positions"])
    cell = np.ascontiguousarray(w["cell"], dtype=np.float64)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    pbc = np.ascontiguousarray(np.asarray(w["pbc"], bool), dtype=np.int32)
    co = np.zeros(3)
    rc, per_atom, pt, nc = _parse_cutoff(w["cutoff"], N)
    types = None
    if pt is not None:
        types = torch.as_tensor(np.asarray(w["numbers"], np.int32), device=dev)
    pos = torch.from_numpy(np.ascontiguousarray(w["positions"])).to(dev)
    plan = _lib.Plan(0)
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    n = ctypes.c_int64(0)
    _lib.check(lib.mnl_plan_count(plan.handle, sp, P_(co), P_(cell), P_(inv), P_(pbc), N,
                                  P_(pos), rc, None, P_(pt), nc, P_(types), backend,
                                  ctypes.byref(n)))
    P = int(n.value)
    fdt = torch.float32 if f32 else torch.float64
    out = dict(i=torch.empty(P, dtype=torch.int32, device=dev),
               j=torch.empty(P, dtype=torch.int32, device=dev),
               D=torch.empty((P, 3), dtype=fdt, device=dev),
               d=torch.empty(P, dtype=fdt, device=dev),
               S=torch.empty((P, 3), dtype=torch.int32, device=dev))
    np_dev = torch.zeros(1, dtype=torch.int64, device=dev)

    def step():
        _lib.check(lib.mnl_plan_count_async(plan.handle, sp, P_(co), P_(cell), P_(inv), P_(pbc),
                                            N, P_(pos), rc, None, P_(pt), nc, P_(types),
                                            backend, P_(np_dev)))
        _lib.check(lib.mnl_plan_fill_async(plan.handle, sp, P, *(P_(out[q]) if q in flags else
                                                                None for q in "ijDdS"),
                                           1 if f32 else 0))
    step.plan, step.out, step.sp, step.N = plan, out, sp, N
    return step, P


def _time_steps(torch, stream, step, steps, warmup):
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def leg_latency(lib, _lib, torch, dev, steps, warmup, with_cpu):
    """1000 atoms, 10 A box, rc 2.5: latency of one full build (count, host
    read of the pair total, fill) on the host clock."""
    import ctypes

    import numpy as np

    from matscipy_neighbours_amd import workloads
    w = workloads.uniform_box()
    step, P = _plan_step_fn(lib, _lib, torch, dev, w)
    cell = np.ascontiguousarray(w["cell"])
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    pbc = np.ones(3, np.int32)
    pos = torch.from_numpy(w["positions"]).to(dev)
    o = step.out
    n = ctypes.c_int64(0)

    def build():
        _lib.check(lib.mnl_plan_count(step.plan.handle, step.sp, _lib.ptr(np.zeros(3)),
                                      _lib.ptr(cell), _lib.ptr(inv), _lib.ptr(pbc), 1000,
                                      _lib.ptr(pos), 2.5, None, None, 0, None, -1,
                                      ctypes.byref(n)))
        _lib.check(lib.mnl_plan_fill(step.plan.handle, step.sp, *(_lib.ptr(o[q])
                                                                  for q in "ijDdS")))
        torch.cuda.synchronize()
    for _ in range(20):
        build()
    t = []
    for _ in range(200):
        t0 = time.perf_counter()
        build()
        t.append(time.perf_counter() - t0)
    entry = {"atoms": 1000, "pairs": P, "us_per_build": float(np.median(t)) * 1e6,
             "how": "count (waits for the pair total) + fill, host clock, median of 200"}
    if with_cpu:
        import oracle
        if oracle.ref_available():
            secs, npairs, threads = oracle.ref_time("ijdDS", np.zeros(3), cell, inv,
                                                    w["pbc"], w["positions"], 2.5, 51)
            entry["cpu_reference_us"] = float(np.median(secs[1:])) * 1e6
            entry["cpu_threads"] = int(threads)
    return "uniform_1000_rc2.5", entry


def leg_slab(lib, _lib, torch, dev, steps, warmup, with_cpu):
    """The slab: 200,000 atoms, two species, per-type cutoffs, pbc TTF."""
    from matscipy_neighbours_amd import workloads
    stream = torch.cuda.current_stream(dev)
    w = workloads.triclinic_slab()
    step, P = _plan_step_fn(lib, _lib, torch, dev, w)
    ms = _time_steps(torch, stream, step, steps, warmup)
    N = step.N
    b = 24 * N + 4 * N + 52 * P
    backend = "dense" if lib.mnl_plan_backend(step.plan.handle) == 0 else "hashed"
    other = 1 if backend == "dense" else 0
    step2, P2 = _plan_step_fn(lib, _lib, torch, dev, w, backend=other)
    assert P2 == P
    ms2 = _time_steps(torch, stream, step2, steps, warmup)
    return "slab_200k_pertype", {
        "atoms": N, "pairs": P, "ms_per_step": ms, "pairs_per_s": P / (ms * 1e-3),
        "hbm_frac": b / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), "backend": backend,
        f"ms_per_step_{'hashed' if other else 'dense'}_grid": ms2}


def leg_vacuum(lib, _lib, torch, dev, steps, warmup, with_cpu):
    """131,072 atoms in a corner of a 20x wider periodic box: the built-in rule
    takes the hashed grid (about 10^8 dense cells); the dense grid is timed
    beside it."""
    from matscipy_neighbours_amd import workloads
    stream = torch.cuda.current_stream(dev)
    w = workloads.vacuum_cluster()
    step, P = _plan_step_fn(lib, _lib, torch, dev, w)
    ms = _time_steps(torch, stream, step, steps, warmup)
    N = step.N
    backend = "dense" if lib.mnl_plan_backend(step.plan.handle) == 0 else "hashed"
    other = 1 if backend == "dense" else 0
    step2, P2 = _plan_step_fn(lib, _lib, torch, dev, w, backend=other)
    assert P2 == P
    ms2 = _time_steps(torch, stream, step2, steps, warmup)
    del step2
    return "vacuum_cluster_131k", {
        "atoms": N, "pairs": P, "ms_per_step": ms, "pairs_per_s": P / (ms * 1e-3),
        "hbm_frac": (24 * N + 52 * P) / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "backend": backend, f"ms_per_step_{'hashed' if other else 'dense'}_grid": ms2}


def leg_fcc_f32(lib, _lib, torch, dev, steps, warmup, with_cpu):
    """The headline fcc workload with the float32 output option (D, d
    float32; pairs still decided in float64): 36 B per pair instead of 52."""
    from matscipy_neighbours_amd import workloads
    stream = torch.cuda.current_stream(dev)
    w = workloads.fcc_bulk()
    step, P = _plan_step_fn(lib, _lib, torch, dev, w, f32=True)
    ms = _time_steps(torch, stream, step, steps, warmup)
    return "fcc_1M_ijdDS_f32", {
        "atoms": step.N, "pairs": P, "ms_per_step": ms, "pairs_per_s": P / (ms * 1e-3),
        "how": "count + fill of i, j, D (float32), d (float32), S into pre-allocated outputs"}


def leg_autograd(lib, _lib, torch, dev, steps, warmup, with_cpu):
    """Differentiable bcc: forward ijdD (f64, f32) and the backward kernel."""
    from matscipy_neighbours_amd import workloads
    stream = torch.cuda.current_stream(dev)
    w = workloads.bcc_sheared()
    fwd = {}
    for f32 in (False, True):
        step, P = _plan_step_fn(lib, _lib, torch, dev, w, f32=f32, flags="ijDd")
        fwd["f32" if f32 else "f64"] = (_time_steps(torch, stream, step, steps, warmup), step)
    step = fwd["f64"][1]
    o = step.out
    step()
    N = step.N
    gD = torch.randn((P, 3), dtype=torch.float64, device=dev)
    gd = torch.randn(P, dtype=torch.float64, device=dev)
    gpos = torch.empty((N, 3), dtype=torch.float64, device=dev)
    gcell = torch.empty((3, 3), dtype=torch.float64, device=dev)
    # S is needed for the cell gradient: one more fill with S
    S = torch.empty((P, 3), dtype=torch.int32, device=dev)
    _lib.check(lib.mnl_plan_fill_async(step.plan.handle, step.sp, P, None, None, None, None,
                                       _lib.ptr(S), 0))

    def backward(det):
        def run():
            _lib.check(lib.mnl_plan_pair_backward(step.plan.handle, step.sp, P, N,
                                                  _lib.ptr(o["i"]), _lib.ptr(o["j"]),
                                                  _lib.ptr(S), _lib.ptr(o["D"]), _lib.ptr(o["d"]),
                                                  _lib.ptr(gD), _lib.ptr(gd), 0, None, det,
                                                  _lib.ptr(gpos), _lib.ptr(gcell)))
        return run
    ms_b = _time_steps(torch, stream, backward(0), steps, warmup)
    ms_bd = _time_steps(torch, stream, backward(1), steps, warmup)
    return "bcc_131k_autograd", {
        "atoms": N, "pairs": P,
        "forward_ijdD_f64_ms": fwd["f64"][0], "forward_ijdD_f32_ms": fwd["f32"][0],
        "backward_ms": ms_b, "backward_deterministic_ms": ms_bd,
        "forward_f64_hbm_frac": (24 * N + 40 * P) / (fwd["f64"][0] * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "forward_f32_hbm_frac": (24 * N + 24 * P) / (fwd["f32"][0] * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "backward_hbm_frac": (108 * P + 24 * N) / (ms_b * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "how": "forward = count + fill of i, j, D, d; backward reads i, j, S, D, d, dL/dD, "
               "dL/dd (108 B/pair) and writes dL/dr (24 B/atom) and dL/dcell"}


def leg_lj(lib, _lib, torch, dev, steps, warmup, with_cpu):
    """The fused Lennard-Jones consumer on the fcc 1M workload (eps 0.4,
    sigma 2.3, rc 5 A): forces, energy and virial straight from the cell
    grid, against building the ijdD list (count + fill of i, D, d) and
    summing the same forces over it with torch."""
    import ctypes

    import numpy as np

    from matscipy_neighbours_amd import workloads
    stream = torch.cuda.current_stream(dev)
    w = workloads.fcc_bulk()
    N = len(w["positions"])
    cell = np.ascontiguousarray(w["cell"], dtype=np.float64)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    pbc = np.ones(3, np.int32)
    co = np.zeros(3)
    pos = torch.from_numpy(w["positions"]).to(dev)
    eps, sig, rc = 0.4, 2.3, 5.0
    plan = _lib.Plan(0)
    sp = ctypes.c_void_p(stream.cuda_stream)
    P_ = _lib.ptr
    forces = torch.empty((N, 3), dtype=torch.float64, device=dev)
    energy = torch.empty(1, dtype=torch.float64, device=dev)
    virial = torch.empty((3, 3), dtype=torch.float64, device=dev)

    def fused():
        _lib.check(lib.mnl_plan_lj(plan.handle, sp, P_(co), P_(cell), P_(inv), P_(pbc), N,
                                   P_(pos), rc, eps, sig, P_(forces), P_(energy), P_(virial)))
    ms_fused = _time_steps(torch, stream, fused, steps, warmup)

    step, P = _plan_step_fn(lib, _lib, torch, dev, w, flags="iDd")
    o = step.out
    f_list = torch.empty_like(forces)
    counts = torch.empty(N, dtype=torch.int32, device=dev)
    # the rows are i-sorted: a segmented sum over the plan's per-atom counts
    # (torch.segment_reduce), or index_add_ where that is not available
    try:
        torch.segment_reduce(torch.ones((4, 3), dtype=torch.float64, device=dev), "sum",
                             lengths=torch.tensor([1, 3], device=dev))
        consumer = "segment_reduce"
    except (RuntimeError, NotImplementedError):
        consumer = "index_add_"

    def via_list():
        step()
        r2 = o["d"] * o["d"]
        s6 = (sig * sig / r2) ** 3
        coef = -24.0 * eps * (2.0 * s6 * s6 - s6) / r2
        terms = coef[:, None] * o["D"]
        if consumer == "segment_reduce":
            _lib.check(lib.mnl_plan_counts(step.plan.handle, step.sp, P_(counts)))
            f_list.copy_(torch.segment_reduce(terms, "sum", lengths=counts))
        else:
            f_list.zero_().index_add_(0, o["i"], terms)
    ms_list = _time_steps(torch, stream, via_list, steps, warmup)
    fused()
    torch.cuda.synchronize()
    scale = float(forces.abs().max())
    return "fcc_1M_lj_fused", {
        "atoms": N, "pairs": P, "fused_ms": ms_fused, "via_list_ms": ms_list,
        "speedup": ms_list / ms_fused,
        "max_force_difference_rel": float((forces - f_list).abs().max()) / scale,
        "how": "fused: mnl_plan_lj (binning + one traversal accumulating forces, energy and "
               "virial, no pair list written); via list: count + fill of i, D, d "
               f"into pre-allocated outputs, then the forces summed with torch {consumer} "
               "over the i-sorted rows; HIP events over the steps"}


def leg_post(lib, _lib, torch, dev, steps, warmup, with_cpu):
    """The consumers beside the list on the fcc 1M workload, through the
    public functions with device-resident inputs and outputs: coordination
    (count pass only, rc 5 A), neighbour_matrix (rc 5 A, 64 slots per atom),
    first_neighbours over the rc 5 A i list (46.6 M pairs) and triplet_list
    over the rc 3 A list (12 neighbours per atom, 132 triplets per atom),
    and with the distance cutoff 2.55 A (about half the pairs kept).
    HIP events on torch's stream around each call, which includes the host
    reads of totals the functions make."""
    import numpy as np

    from matscipy_neighbours_amd import (coordination, first_neighbours, neighbour_list,
                                         neighbour_matrix, triplet_list, workloads)
    stream = torch.cuda.current_stream(dev)
    w = workloads.fcc_bulk()
    N = len(w["positions"])
    cell = np.ascontiguousarray(w["cell"], dtype=np.float64)
    pos = torch.from_numpy(w["positions"]).to(dev)
    K = 64
    geo = dict(positions=pos, cell=cell, pbc=True)
    out = {}

    def rate(ms, nbytes):
        return {"ms": ms, "algorithmic_GB": nbytes / 1e9,
                "achieved_GBps": nbytes / (ms * 1e-3) / 1e9,
                "hbm_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}

    ms = _time_steps(torch, stream, lambda: coordination(cutoff=5.0, **geo), steps, warmup)
    out["coordination_rc5"] = dict(rate(ms, 28 * N), bytes_per_atom="24 read + 4 written")

    m = {}

    def matrix():
        m["r"] = neighbour_matrix(cutoff=5.0, max_neighbours=K, **geo)
    ms = _time_steps(torch, stream, matrix, steps, warmup)
    cnt = m["r"][2]
    out["neighbour_matrix_rc5_K64"] = dict(
        rate(ms, 24 * N + (4 + 24) * K * N + 4 * N), max_count=int(cnt.max()),
        bytes_per_atom="24 read + 28 per slot (idx, D) + 4 (count) written")
    del m["r"], cnt

    i5 = neighbour_list("i", cutoff=5.0, **geo)
    P5 = int(i5.numel())
    ms = _time_steps(torch, stream, lambda: first_neighbours(N, i5), steps, warmup)
    out["first_neighbours_rc5"] = dict(rate(ms, 4 * P5 + 4 * (N + 1)), pairs=P5,
                                       bytes="4 per pair read + 4 per atom written")
    del i5

    i3 = neighbour_list("i", cutoff=3.0, **geo)
    seed = first_neighbours(N, i3)
    P3 = int(i3.numel())
    t = {}

    def triplets():
        t["r"] = triplet_list(seed)
    ms = _time_steps(torch, stream, triplets, steps, warmup)
    T = int(t["r"][0].numel())
    out["triplet_list_rc3"] = dict(rate(ms, 4 * (N + 1) + 8 * T), pairs=P3, triplets=T,
                                   bytes="4 per seed read + 8 per triplet (ij, ik) written")
    del t["r"]
    d3 = neighbour_list("d", cutoff=3.0, **geo)

    def triplets_cut():
        t["r"] = triplet_list(seed, d3, 2.55)
    ms = _time_steps(torch, stream, triplets_cut, steps, warmup)
    Tc = int(t["r"][0].numel())
    out["triplet_list_rc3_cutoff2.55"] = dict(
        rate(ms, 4 * (N + 1) + 8 * P3 + 8 * Tc), pairs=P3, triplets=Tc,
        bytes="4 per seed + 8 per pair (d) read + 8 per triplet written")
    out["how"] = ("public functions on device tensors, outputs allocated by torch's caching "
                  "allocator inside the timed call; HIP events on torch's stream, mean over "
                  "the steps")
    return "fcc_1M_consumers", out


def leg_alloc(lib, _lib, torch, dev, steps, warmup, with_cpu):
    """The fcc 1M ijdDS build with the output allocation inside the timed
    region: the one-shot C entry point (a device result owning its buffers,
    taken per call from the library's kept blocks and given back by
    mnl_result_free), on the same positions every call and on four copies
    jittered by 0.01 A in turn, whose pair totals differ as an MD loop's do;
    and the Python neighbour_list with device positions (outputs from torch's
    caching allocator). Host clock, median over the calls."""
    import ctypes

    import numpy as np

    from matscipy_neighbours_amd import neighbour_list, workloads
    w = workloads.fcc_bulk()
    N = len(w["positions"])
    cell = np.ascontiguousarray(w["cell"], dtype=np.float64)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    pbc = np.ones(3, np.int32)
    co = np.zeros(3)
    P_ = _lib.ptr
    pos = torch.from_numpy(w["positions"]).to(dev)
    rng = np.random.default_rng(11)
    jittered = [torch.from_numpy(w["positions"] + rng.normal(0.0, 0.01, (N, 3))).to(dev)
                for _ in range(4)]
    res = ctypes.c_void_p()
    npairs = []

    def oneshot_on(p):
        _lib.check(lib.mnl_neighbour_list_device(31, P_(co), P_(cell), P_(inv), P_(pbc), N,
                                                 P_(p), 1, w["cutoff"], None, None, 0, None,
                                                 dev.index or 0, ctypes.byref(res)))
        torch.cuda.synchronize()
        npairs.append(lib.mnl_result_npairs(res))
        lib.mnl_result_free(res)

    def oneshot():
        oneshot_on(pos)

    calls = [0]

    def oneshot_jittered():
        oneshot_on(jittered[calls[0] % 4])
        calls[0] += 1

    def pyapi():
        out = neighbour_list("ijdDS", positions=pos, cell=cell, pbc=True, cutoff=w["cutoff"])
        torch.cuda.synchronize()
        del out

    def host_ms(fn):
        for _ in range(warmup):
            fn()
        t = []
        for _ in range(steps):
            t0 = time.perf_counter()
            fn()
            t.append(time.perf_counter() - t0)
        return float(np.median(t)) * 1e3

    ms_one = host_ms(oneshot)
    P = int(npairs[-1])
    npairs.clear()
    ms_jit = host_ms(oneshot_jittered)
    P_jit = sorted(set(int(x) for x in npairs[-4:]))
    lib.mnl_release_cached_memory()
    ms_py = host_ms(pyapi)
    return "fcc_1M_with_allocation", {
        "atoms": N, "pairs": P, "oneshot_ms": ms_one, "oneshot_jittered_ms": ms_jit,
        "jittered_pairs": P_jit, "python_api_ms": ms_py,
        "oneshot_pairs_per_s": P / (ms_one * 1e-3),
        "how": "one-shot: mnl_neighbour_list_device (count, host read of the total, output "
               "allocation, fill) + mnl_result_free per call, on the same positions and on "
               "four copies jittered by 0.01 A (normal) in turn; Python: "
               "neighbour_list('ijdDS') with device positions; host clock around each call "
               "including a device synchronisation, median"}


OTHER_LEGS = {"latency": leg_latency, "slab": leg_slab, "vacuum": leg_vacuum,
              "fcc_f32": leg_fcc_f32,
              "autograd": leg_autograd, "lj": leg_lj, "alloc": leg_alloc,
              "post": leg_post}


def other_workloads(lib, _lib, torch, dev, steps, warmup, with_cpu):
    """The smaller workloads: the 1000-atom latency case, the per-type slab,
    the float32 output option on fcc 1M, and the differentiable bcc case
    (forward f64 / f32 and backward)."""
    res = {}
    for leg in OTHER_LEGS.values():
        name, entry = leg(lib, _lib, torch, dev, steps, warmup, with_cpu)
        res[name] = entry
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-reps", type=int, default=30,
                    help="timed calls of the CPU baseline (after one warm-up); the median "
                         "is reported")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-systems leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other workloads (1000-atom latency, slab, fcc 1M with "
                         "float32 outputs, bcc autograd)")
    ap.add_argument("--reps", type=int, default=64, help="fcc unit cells per edge")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the engine's built-in launch shapes instead of tuning them first "
                         "(profiling runs: the tuner's candidate launches would enter the "
                         "per-kernel statistics)")
    ap.add_argument("--leg", choices=["batch"] + list(OTHER_LEGS),
                    help="run only this secondary workload (for profiling it on its own) "
                         "and print its result as one JSON line")
    ap.add_argument("--chunks", type=int, choices=[4, 5],
                    help="run the fcc fill with this candidate window (split 1) instead of "
                         "the autotuned one (profiling runs of either fill kernel)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (to scale HBM "
                         "traffic counters) and exit")
    ap.add_argument("--traffic-from", nargs=4,
                    metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="write the fill kernel's HBM traffic summary from rocprofv3 --pmc "
                         "counter CSVs and exit (no GPU needed)")
    args = ap.parse_args()
    if args.traffic_from:
        traffic_summary(*args.traffic_from, commands=[
            "rocprofv3 --pmc FETCH_SIZE -f csv -d <dir> -o run -- python3 bench.py --steps 3 "
            "--warmup 1 --no-cpu --no-batch --no-extra --chunks <4 or 5>",
            "rocprofv3 --pmc WRITE_SIZE -f csv -d <dir> -o run -- python3 bench.py --steps 3 "
            "--warmup 1 --no-cpu --no-batch --no-extra --chunks <4 or 5>",
            "rocprofv3 --pmc FETCH_SIZE -f csv -d <dir> -o run -- python3 bench.py "
            "--calibrate-copy --steps 3",
            "rocprofv3 --pmc WRITE_SIZE -f csv -d <dir> -o run -- python3 bench.py "
            "--calibrate-copy --steps 3",
            "python3 bench.py --traffic-from <fetch csv> <write csv> <copy fetch csv> "
            "<copy write csv>"])
        return
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(max(args.steps, 1)):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read_per_launch": a.numel() * 8,
                          "bytes_written_per_launch": b.numel() * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the neighbour-list engine runs on a single device; splitting a build "
              "over several GPUs by atom range is not implemented)", file=sys.stderr)
        sys.exit(2)

    import ctypes

    import numpy as np
    import torch

    from matscipy_neighbours_amd import _lib, workloads

    lib = _lib.load()
    _lib.require_gpu()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    if args.leg:
        _log(f"leg {args.leg}")
        if args.leg == "batch":
            name, entry = "batched", batched_leg(lib, _lib, torch, dev, args.steps, args.warmup,
                                                 tune=not args.no_autotune)
        else:
            name, entry = OTHER_LEGS[args.leg](lib, _lib, torch, dev, args.steps, args.warmup,
                                               not args.no_cpu)
        print(json.dumps({"leg": name, "steps": args.steps, "warmup": args.warmup,
                          "result": entry}), flush=True)
        return

    _log("generating workload")
    w = workloads.fcc_bulk(reps=args.reps)
    N = w["positions"].shape[0]
    cell = np.ascontiguousarray(w["cell"], dtype=np.float64)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    co = np.zeros(3)
    pbc = np.ascontiguousarray(w["pbc"], dtype=np.int32)
    pos = torch.from_numpy(w["positions"]).to(dev)
    plan = _lib.Plan(0)
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    P_ = _lib.ptr

    npairs = ctypes.c_int64(0)
    _lib.check(lib.mnl_plan_count(plan.handle, sp, P_(co), P_(cell), P_(inv), P_(pbc), N,
                                  P_(pos), w["cutoff"], None, None, 0, None, -1,
                                  ctypes.byref(npairs)))
    P = int(npairs.value)
    _log(f"N={N} pairs={P}")
    if args.chunks:
        _lib.check(lib.mnl_plan_set_launch_shape(plan.handle, 1, args.chunks))
        shape = {"source": "forced", "split": 1, "chunks": args.chunks}
    elif args.no_autotune:
        # the engine's built-in window: 5 chunks when a home cell's candidates
        # (27 neighbour cells at the mean density) fit 4 < c/64 <= 5
        ncells = int(np.prod(np.floor(np.diag(cell) / w["cutoff"])))
        cand = 27.0 * N / ncells
        shape = {"source": "built-in", "split": 1, "chunks": 5 if 256 < cand <= 320 else 4}
    else:
        shape = _autotune(lib, _lib, plan, sp)
        _log(f"autotuned launch shape: {shape}")
    fill_kernel = _fill_kernel(shape["chunks"])
    out_i = torch.empty(P, dtype=torch.int32, device=dev)
    out_j = torch.empty(P, dtype=torch.int32, device=dev)
    out_d = torch.empty(P, dtype=torch.float64, device=dev)
    out_D = torch.empty((P, 3), dtype=torch.float64, device=dev)
    out_S = torch.empty((P, 3), dtype=torch.int32, device=dev)
    np_dev = torch.zeros(1, dtype=torch.int64, device=dev)

    def step(ev=None):
        _lib.check(lib.mnl_plan_count_async(plan.handle, sp, P_(co), P_(cell), P_(inv),
                                            P_(pbc), N, P_(pos), w["cutoff"], None, None, 0,
                                            None, -1, P_(np_dev)))
        if ev is not None:
            ev[0].record(stream)
        _lib.check(lib.mnl_plan_fill_async(plan.handle, sp, P, P_(out_i), P_(out_j),
                                           P_(out_D), P_(out_d), P_(out_S), 0))
        if ev is not None:
            ev[1].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _log(f"warmup done ({args.warmup} steps)")

    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t0.record(stream)
    for k in range(args.steps):
        step(evs[k])
    t1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_step = t0.elapsed_time(t1) / args.steps
    fill_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    assert int(np_dev.item()) == P, "pair count changed between steps"
    _log(f"timed {args.steps} steps: {ms_step:.4f} ms/step (wall {wall * 1e3 / args.steps:.4f}),"
         f" fill {fill_ms:.4f} ms")

    # Algorithmic bytes. Pipeline: read positions 24 B/atom, write 52 B/pair.
    # Fill kernel: write 52 B/pair, read each atom's cell-sorted record
    # (32 B position + 16 B image/index) and its int64 offset once: 56 B/atom.
    pipe_bytes = 24 * N + 52 * P
    fill_bytes = 52 * P + 56 * N
    achieved = fill_bytes / (fill_ms * 1e-3) / 1e9
    copy_gbs = copy_rate(torch, stream)
    _log(f"device-to-device copy: {copy_gbs:.0f} GB/s (bytes read + written)")
    # HBM traffic of the fill from the committed counter summary, valid only
    # while the kernel source is the one it was measured on.
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tf = json.load(f)
        if tf.get("kernel_source_sha256") != _source_sha():
            _log("traffic summary is stale (kernel source changed): traffic = null")
        elif fill_kernel not in tf.get("kernels", {}):
            _log(f"traffic summary has no entry for {fill_kernel}: traffic = null")
        else:
            traffic = tf["kernels"][fill_kernel]["bytes_per_launch"]

    batched = None
    if not args.no_batch:
        _log("batched systems leg")
        batched = batched_leg(lib, _lib, torch, dev, args.steps, args.warmup,
                              tune=not args.no_autotune)
        _log(f"batched: {batched}")

    extra = None
    if not args.no_extra:
        _log("other workloads")
        extra = other_workloads(lib, _lib, torch, dev, args.steps, args.warmup,
                                not args.no_cpu)
        _log(f"other workloads: {extra}")

    cpu = None
    if not args.no_cpu:
        _log("timing the CPU baseline")
        cpu = cpu_baseline(w, cell, inv, "ijdDS", args.cpu_reps)
        _log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC,
        "value": P / (ms_step * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"fcc a=3.615 A, {args.reps}^3 cells ({N} atoms), jitter 0.1 A, "
                        f"periodic cube {args.reps * 3.615:.2f} A, rc 5.0 A, ijdDS f64, "
                        "i-sorted, outputs pre-allocated",
            "atoms": N,
            "pairs": P,
        },
        "hbm_frac_pipeline": pipe_bytes / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "roofline": {
            "bound": "hbm",
            "kernel": fill_kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "fill_ms": fill_ms,
            "copy_measured": copy_gbs,
            "frac_of_copy": achieved / copy_gbs,
        },
        "launch_shape": shape,
        "cpu_baseline": cpu,
        "batched": batched,
        "other_workloads": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
