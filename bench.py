#!/usr/bin/env python3
"""POLFED hot-path benchmark on one MI355X.

Default workload (the precision configuration): random-field XXZ chain, L=22
periodic, Sz=0 sector (D=705432), J=1, Delta=1, fields h_i ~ U[-2,2] from
numpy seed 2, block size 16, nev=200, target = middle of the spectrum,
reference defaults (cutoff 0.17, DoS N=250, R=300 Jackson), eigentol 1e-10,
inputs resident on the GPU.  In one run:

* time to nev eigenpairs (||H psi - E psi|| <= 1e-10) of the plain FP64-filter
  solve and of the FP32-filter form (mixed FP64/FP32 filters + FP64 polish),
  same start block and KPM probes, every returned pair's eigenvalue error
  against the FP64 run and its residual (``time_to_nev``);
* the timed steps: each step is one FP32 Clenshaw filter application of the
  solve's polynomial (order K, K+1 fused step launches) on the 16-column
  block; ``value`` is their algorithmic bandwidth (4 x 16 x D x 4 bytes per
  launch), the roofline entry that of the FP32 step kernel from HIP events;
* FP32 against FP64 KPM moments on the same probes;
* the CPU baseline (the oracle's C restatement of the reference's CPU filter,
  a bounded sample projected to the FP64 solve) and the reference's
  CPU-runnable case (XXZ L=14, single vector) solved in full on CPU and GPU.

``--workload xxz-l20`` keeps the former bench line (one FP64 solve per step).
Prints one JSON line.  --gpus N > 1 is not supported: this build targets a
single GPU.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import pathlib
import sys
import time

REPO = pathlib.Path(__file__).resolve().parent
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))

# the bench's pure-FP32 solve starts only before this many seconds of the run
# (the whole default run must stay well inside a 600-second limit)
PURE_FP32_BEFORE_S = 360.0
METRIC = "time to nev eigenpairs (res≤1e-10), FP32-filter vs FP64; filter HBM GB/s"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = REPO / "profiles" / "pmc_traffic.json"


def _load_package():
    if "polfed_amd" in sys.modules:
        return sys.modules["polfed_amd"]
    pkg_dir = REPO / "polfed.jl_amd"
    spec = importlib.util.spec_from_file_location(
        "polfed_amd", pkg_dir / "__init__.py", submodule_search_locations=[str(pkg_dir)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["polfed_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class FilterTimer:
    """Wraps op.filter: HIP events (on the launch stream) around each call."""

    def __init__(self, op):
        import torch
        self.torch = torch
        self.op = op
        self.inner = op.filter
        self.events = []
        self.ncols = []
        self.orders = []
        op.filter = self

    def __call__(self, coeffs, X, Y=None, dtype="f64", polytype="Chebyshev"):
        t = self.torch
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        e0.record()
        out = self.inner(coeffs, X, Y, dtype=dtype, polytype=polytype)
        e1.record()
        self.events.append((e0, e1))
        self.ncols.append(X.shape[0] if X.dim() == 2 else 1)
        self.orders.append(len(coeffs) - 1)
        return out

    def reset(self):
        self.events, self.ncols, self.orders = [], [], []

    def summary(self, D, elem_bytes):
        """(avg ms per recurrence step, algorithmic bytes per step, steps)."""
        self.torch.cuda.synchronize()
        ms = sum(a.elapsed_time(b) for a, b in self.events)
        steps = sum(k + 1 for k in self.orders)          # K recurrence steps + final sum
        bytes_total = sum((k + 1) * 4 * s * D * elem_bytes for k, s in zip(self.orders, self.ncols))
        return ms / max(steps, 1), bytes_total / max(steps, 1), steps, ms


def build_workload(dev, fp="f64"):
    import numpy as np
    from polfed_amd import models
    L = 20
    op = models.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 3.0, rng=np.random.default_rng(1),
                                device=dev)
    q, _ = np.linalg.qr(np.random.default_rng(101).standard_normal((op.D, 8)))
    x0 = dev.to_device(q)
    return op, x0


def progress_filter(op, label, every=5):
    """Log one line per `every` filter applications (long solves stay audible)."""
    inner = op.filter
    state = {"n": 0, "t0": time.perf_counter()}

    def wrapped(coeffs, X, Y=None, dtype="f64", polytype="Chebyshev"):
        out = inner(coeffs, X, Y, dtype=dtype, polytype=polytype)
        state["n"] += 1
        if state["n"] % every == 0:
            log(f"{label}: filter application {state['n']} ({dtype}), "
                f"{time.perf_counter() - state['t0']:.1f} s")
        return out

    op.filter = wrapped
    return inner


def kernel_name(op, dtype="f64", s=8):
    """Template instance of the spin-chain step kernel the drivers launch for
    (dtype, s) -- the autotuned layout -- and that layout as the --layout
    string (V,PBLOCKS,WIDE,PARTITION) of the filter-only profiling driver."""
    t = op.tuning(s, dtype)
    T = "double" if dtype == "f64" else "float"
    V, G = t["vec"], s // t["vec"]
    if t["occupancy"]:
        name = f"chain_step_kernel_occ<{T}, {V}, {G}>"
    else:
        name = f"chain_step_kernel<{T}, {V}, {G}, true, {2 if t['wide'] else 1}>"
    return name, f"{V},{t['pblocks']},{t['wide']},{t['row_partition']}"


def measured_traffic(kernel):
    """HBM-side bytes per launch of the step kernel instance `kernel` (full
    template name) from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (each counter scaled by its plain-copy calibration factor), or
    None when that instance was not counted."""
    try:
        table = json.loads(TRAFFIC_FILE.read_text())["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    entry = table.get(kernel)
    return float(entry["bytes_per_launch"]) if entry else None


def filter_only(args):
    """Profiling driver: `--filter-only` Clenshaw steps of a fixed filter on the
    bench workload's operator and start block (XXZ L=20, FP64 block 8, order
    200), with ``--workload xxz-l22`` on the L=22 operator and 16-column start
    block of that workload (order 100), or with ``--workload j1j2-l24`` on the
    J1-J2 L=24 operator with a random 32-column block (order 20); the
    autotuned step kernel unless
    ``--layout`` fixes it.  Prints one JSON line with the average step time
    (HIP events)."""
    import numpy as np
    import torch
    from polfed_amd import models
    from polfed_amd.device import Device
    dev = Device(0)
    if args.workload == "qsun-batch":
        return batch_filter_only(args, dev)
    if args.workload == "j1j2-l24":
        op = models.j1j2_hamiltonian(24, 12, 1.0, 0.5, 1.0, 1.0, 0.0, device=dev)
        op.set_rescale(-9.0, 9.0)
        s, K = 32, 20
        x0 = dev.to_device(np.random.default_rng(0).standard_normal((op.D, s)))
    elif args.workload == "xxz-l22":
        op = models.xxz_hamiltonian(22, 11, 1.0, 1.0, 2.0, rng=np.random.default_rng(2),
                                    device=dev)
        op.set_rescale(-18.0, 14.0)
        s, K = 16, 100
        x0 = dev.to_device(np.linalg.qr(np.random.default_rng(102).standard_normal((op.D, s)))[0])
    else:
        op, x0 = build_workload(dev)
        op.set_rescale(-17.2, 13.5)
        s, K = 8, 200
    dt = args.dtype
    if args.layout:  # a fixed launch configuration (no autotuner trials in the profile)
        V, pb, wide, part = (int(v) for v in args.layout.split(","))
        op.set_tuning(vec=V, pblocks=pb, wide=wide, row_partition=part)
    else:
        op.ensure_tuned(s, dt)
    if args.row_partition >= 0:  # force the row order over the autotuned choice
        op.set_tuning(row_partition=args.row_partition)
    c = np.cos(np.arange(K + 1) * 0.3) * 0.01
    y = torch.empty_like(x0)
    reps = max(1, args.filter_only // (K + 1))
    op.filter(c, x0, y, dtype=dt)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        op.filter(c, x0, y, dtype=dt)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / (reps * (K + 1))
    esz = 8 if dt == "f64" else 4
    kname, klayout = kernel_name(op, dt, s)
    print(json.dumps({"mode": "filter-only", "kernel": kname,
                      "layout": klayout, "forced": bool(args.layout), "dtype": dt,
                      "D": op.D, "s": s, "steps": reps * (K + 1), "us_per_step": us,
                      "algorithmic_GBps": 4 * s * op.D * esz / (us * 1e-6) / 1e9}), flush=True)


def batch_filter_only(args, dev):
    """Profiling driver for the multi-instance filter: the 64 Quantum-Sun
    instances of the qsun-batch workload, each rescaled by the fixed bounds
    [-40, 40] (wider than every instance's spectrum) with a 401-coefficient
    filter, one column each; the autotuned kernel (``--dtype``), then about
    ``--filter-only`` batched steps in stream mode (one kernel per step, so a
    profiler sees every step).  ``--batch-form`` other than auto replaces the
    autotuned form by a fixed one: rows from memory or block rows in LDS with
    4 (FP64) / 8 (FP32) rows per lane, or row vectors one at a time with one
    row group per lane."""
    import numpy as np
    import torch
    from polfed_amd.batch import QSunBatch
    ops, x0s = qsun_batch_instances(dev, args.instances)
    for op in ops:
        op.set_rescale(-40.0, 40.0)
    K = 400
    b = QSunBatch(ops, 1)
    b.set_filter([np.cos(np.arange(K + 1) * 0.3) * 0.01 for _ in ops])
    dt = args.dtype
    tuned_us = b.autotune(dt)
    b.set_tuning(graph=0)
    rpl = 4 if dt == "f64" else 8
    forms = {"plain": dict(vec=1, rows_per_lane=rpl, lds_rows=0, row_vectors=0),
             "lds": dict(vec=1, rows_per_lane=rpl, lds_rows=1, row_vectors=0),
             "rowvec": dict(vec=1, rows_per_lane=1, lds_rows=0, row_vectors=1)}
    if args.batch_form != "auto":
        b.set_tuning(**forms[args.batch_form])
    X = [dev.to_device(x) for x in x0s]
    Y = [torch.empty_like(x) for x in X]
    reqs = [(x, y, dt) for x, y in zip(X, Y)]
    b.filter(reqs)
    torch.cuda.synchronize()
    reps = max(1, args.filter_only // (K + 1))
    t0 = time.perf_counter()
    for _ in range(reps):
        b.filter(reqs)
    torch.cuda.synchronize()
    us = (time.perf_counter() - t0) / (reps * (K + 1)) * 1e6
    rows = sum(op.D for op in ops)
    esz = 8 if dt == "f64" else 4
    print(json.dumps({"mode": "batch-filter-only", "dtype": dt, "instances": len(ops),
                      "rows": rows, "tuning": b.tuning(), "autotune_us_per_step": tuned_us,
                      "stream_us_per_step": us, "steps": reps * (K + 1),
                      "algorithmic_GBps": 4 * rows * esz / (us * 1e-6) / 1e9}), flush=True)


def packed_compare(args):
    """Profiling driver: the reference's own GPU data structure -- the sparse
    matrix packed by value (``polfed(mat, ...)``) -- against the matrix-free
    kernel, on the same XXZ chains (L=20 FP64 block 8, L=22 FP32 block 16):
    the matrix is read back from the matrix-free operator's connectivity,
    packed and uploaded; both are autotuned and time 3 order-200 filters.
    Prints one JSON line per (chain, operator)."""
    import numpy as np
    import scipy.sparse as sp
    import torch
    from polfed_amd import models
    from polfed_amd.device import Device
    from polfed_amd.operators import PackedOperator
    dev = Device(0)
    for L, s, dt in ((20, 8, "f64"), (22, 16, "f32")):
        mf = models.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 3.0, rng=np.random.default_rng(1),
                                    device=dev)
        D = mf.D
        cols, vals, counts, diag = mf.connectivity(max_per_row=64)
        keep = np.arange(64)[None, :] < counts[:, None]
        H = (sp.coo_matrix((vals[keep], (np.repeat(np.arange(D), counts), cols[keep])),
                           shape=(D, D)) + sp.diags(diag)).tocsr()
        pk = PackedOperator.from_matrix(H, device=dev)
        X = dev.to_device(np.random.default_rng(0).standard_normal((D, s)))
        c = np.cos(np.arange(201) * 0.3) * 0.01
        ys = {}
        for name, op in (("matrix-free", mf), ("packed", pk)):
            op.set_rescale(-20.0, 20.0)
            op.autotune(s, dt, 30)
            Y = op.filter(c, X, dtype=dt)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(3):
                op.filter(c, X, Y, dtype=dt)
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / (3 * 201)
            ys[name] = op.filter(c, X, dtype="f64")
            esz = 8 if dt == "f64" else 4
            print(json.dumps({"mode": "packed-compare", "L": L, "D": D, "s": s, "dtype": dt,
                              "operator": name, "nnz_offdiag": int(H.nnz - D),
                              "us_per_step": us,
                              "algorithmic_GBps": 4 * s * D * esz / (us * 1e-6) / 1e9}), flush=True)
        torch.cuda.synchronize()
        diff = float((ys["packed"] - ys["matrix-free"]).abs().max() / ys["matrix-free"].abs().max())
        print(json.dumps({"mode": "packed-compare", "L": L, "fp64_filter_rel_diff": diff}), flush=True)


# (D, k, s) of the reorthogonalization profiles: the last one of each solve
# (XXZ L=20 block 8, XXZ L=22 block 16, J1-J2 L=24 block 32)
REORTH_SHAPES = {"l20": (184756, 416, 8), "l22": (705432, 832, 16), "l24": (2704156, 1440, 32)}


def reorth_only(args):
    """Profiling driver: CGS2 reorthogonalization W -= V (V^T W), twice, of an
    s-column block against k basis vectors of length D (``--reorth-shape``:
    the last reorthogonalization of the XXZ L=20, XXZ L=22 or J1-J2 L=24
    solve), timed with HIP events; prints FP64 MFMA TFLOP/s and algorithmic
    GB/s (per pass: 4 D k s flops, 8 (2 D k + 3 D s) bytes)."""
    import torch
    from polfed_amd import blockops as bo
    from polfed_amd.device import Device
    dev = Device(0)
    D, k, s = REORTH_SHAPES[args.reorth_shape]
    g = torch.Generator(device="cuda").manual_seed(7)
    V = torch.randn((k, D), dtype=torch.float64, device="cuda", generator=g)
    W = torch.randn((s, D), dtype=torch.float64, device="cuda", generator=g)
    reps = max(1, args.reorth_only)
    bo.reorth(dev, V, W, passes=2)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        bo.reorth(dev, V, W, passes=2)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    flops = 2 * 4.0 * D * k * s
    nbytes = 2 * 8.0 * (2 * D * k + 3 * D * s)
    print(json.dumps({"mode": "reorth-only", "D": D, "k": k, "s": s, "passes": 2,
                      "us_per_call": us, "TFLOPs": flops / (us * 1e-6) / 1e12,
                      "mfma_f64_frac": flops / (us * 1e-6) / 78.6e12,
                      "algorithmic_GBps": nbytes / (us * 1e-6) / 1e9,
                      "hbm_frac": nbytes / (us * 1e-6) / 8e12}), flush=True)


def qsun_batch_instances(dev, n=64):
    """64 Quantum-Sun disorder realizations: L_grain=3 (8x8 GOE grain scaled
    by 1/sqrt(9)), L_loc = 10..14 round-robin (D = 8192..131072), alpha=0.75,
    zeta=0.2, g0=1, fields U[0.5,1.5]; numpy seeds 1000+i for the draws and
    2000+i for the start vectors."""
    import numpy as np
    from polfed_amd import models
    ops, x0s = [], []
    for i in range(n):
        L_loc = 10 + i % 5
        op = models.qsun_hamiltonian(L_loc, 3, 1.0, 0.75, rng=np.random.default_rng(1000 + i),
                                     device=dev)
        ops.append(op)
        x0s.append(np.random.default_rng(2000 + i).standard_normal(op.D))
    return ops, x0s


def qsun_batch(args):
    """Batched workload: nev=32 pairs of each of 64 Quantum-Sun instances at
    the middle of the spectrum (single-vector Lanczos, eigentol 1e-10).
    Timed, as selected by --modes: the batched mixed-precision solve (FP32
    filters once an instance's wanted pairs are nearly converged), the batched
    FP64 solve, and the plain FP64 path (one polfed call per instance)."""
    import numpy as np
    import torch
    from polfed_amd import FactorizationConfig, TransformConfig, polfed, polfed_batch
    from polfed_amd.device import Device
    dev = Device(0)
    ops, x0s = qsun_batch_instances(dev, args.instances)
    fact = FactorizationConfig(eigentol=1e-10)
    out = {"metric": METRIC, "workload": f"{len(ops)} Quantum-Sun instances, L_grain=3, "
           "L_loc=10..14 round-robin (D=8192..131072), nev=32 each, target middle, "
           "eigentol 1e-10", "unit": "s", "higher_is_better": False}

    def summarize(res_list, t):
        conv = [int(len(v)) for v, _, _ in res_list]
        maxres = max(float(r.lanczos.all_residuals.max()) if r.lanczos.all_residuals is not None
                     and r.lanczos.all_residuals.size else 0.0 for _, _, r in res_list)
        # instances that stop short of nev: the reference's Krylov budget
        # (expectedkrylovdim, maxiter iterations) ran out first
        short = [i for i, c in enumerate(conv) if c < 32]
        return {"time_s": t, "pairs_total": int(sum(conv)), "instances_complete":
                int(sum(c == 32 for c in conv)), "instances_short_of_nev": short,
                "pairs_of_short_instances": [conv[i] for i in short],
                "short_at_iteration_budget": all(res_list[i][2].iterations
                                                 >= res_list[i][2].lanczos.maxiter
                                                 for i in short),
                "max_residual": maxres,
                "iterations_max": int(max(r.iterations for _, _, r in res_list)),
                "orders": [int(r.order) for _, _, r in res_list]}

    modes = [m for m in args.modes.split(",") if m]
    results = {}
    for mode in [m for m in ("mixed", "f64") if m in modes]:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res, brep = polfed_batch(ops, x0s, 32, "middle", produce_report=True, fact=fact,
                                 transform=TransformConfig(dtype=mode, **_switch(args)), progress=log)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        results[mode] = res
        e = summarize(res, t)
        e.update({"setup_s": brep.setup_s, "lanczos_s": brep.lanczos_s,
                  "filter_s": brep.filter_s, "batched_steps": brep.batched_steps,
                  "filter_calls": brep.filter_calls, "autotune_us_per_step": brep.autotune_us,
                  "fp32_iterations": int(sum(r.lanczos.fast_iterations for _, _, r in res))})
        e.pop("orders")
        out[f"batched_{mode}"] = e
        log(f"batched {mode}: {t:.1f} s pairs={e['pairs_total']} maxres={e['max_residual']:.2e}")
    if "mixed" in results and "f64" in results:
        errs = []
        for (v1, _, r1), (v2, _, r2) in zip(results["mixed"], results["f64"]):
            n = min(len(v1), len(v2))
            if n:
                errs.append(float(np.max(np.abs(v1[:n] - v2[:n])) / (r2.Emax - r2.Emin)))
        out["batched_mixed"]["max_eig_err_vs_fp64_over_width"] = max(errs) if errs else None
        out["batched_mixed"]["speedup_vs_batched_fp64"] = (out["batched_f64"]["time_s"]
                                                           / out["batched_mixed"]["time_s"])
    if "plain" in modes:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain = []
        for i, (op, x0) in enumerate(zip(ops, x0s)):
            plain.append(polfed(op, x0, 32, "middle", produce_report=True, fact=fact))
            log(f"plain f64 instance {i + 1}/{len(ops)}: D={op.D} "
                f"{time.perf_counter() - t0:.1f} s")
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        e = summarize(plain, t)
        e.pop("orders")
        out["plain_f64"] = e
        for m in ("mixed", "f64"):
            if f"batched_{m}" in out:
                out[f"batched_{m}"]["speedup_vs_plain_fp64"] = t / out[f"batched_{m}"]["time_s"]
        log(f"plain f64: {t:.1f} s pairs={e['pairs_total']}")
    print(json.dumps(out), flush=True)


def cpu_info():
    """Host CPU model, logical CPUs and the CPUs this process may use."""
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
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def precision_workload(dev):
    """XXZ L=22 Sz=0 (D=705432), J=1, Delta=1, fields h_i = 2(2u-1) with u from
    numpy seed 2, start block = QR of a numpy seed-102 normal D x 16 block."""
    import numpy as np
    from polfed_amd import models
    op = models.xxz_hamiltonian(22, 11, 1.0, 1.0, 2.0, rng=np.random.default_rng(2), device=dev)
    q, _ = np.linalg.qr(np.random.default_rng(102).standard_normal((op.D, 16)))
    return op, dev.to_device(q), q


def _stage_times(rep):
    """Seconds per stage of a solve: bounds probes, DoS, order, and the
    Lanczos stages (filter = mapping; the host projected eigensolver sits in
    convergence)."""
    keys = ("bounds", "dos", "transform", "lanczos.mapping", "lanczos.tridiagonalization",
            "lanczos.orthogonalization", "lanczos.reorthogonalization", "lanczos.convergence",
            "lanczos.switch_test", "lanczos.polish")
    return {k.replace("lanczos.", ""): round(rep.times[k], 3) for k in keys if k in rep.times}


def _pair_errors(vals, res, ref_vals, width):
    """Each returned pair's eigenvalue error against the nearest FP64-run
    eigenvalue (over the spectral width) and its residual."""
    import numpy as np
    ref = np.sort(ref_vals)
    err = []
    for v in vals:
        j = int(np.searchsorted(ref, v))
        cand = [ref[i] for i in (j - 1, j) if 0 <= i < ref.size]
        err.append(min(abs(v - c) for c in cand) / width if cand else float("nan"))
    return np.array(err), np.asarray(res)


def cpu_baseline_l22(op, rep, x0_host, budget_s=8.0):
    """The oracle's C restatement of the reference's packed CPU filter on a
    bounded prefix of one filter application of the solve's polynomial on the
    16-column block, projected to the FP64 solve's (K+1) x iterations: once
    with one thread per column (MulColsParallel(1), the reference's default
    for a block, ParallelizationClenshaw.jl:77-95) and once with every usable
    core splitting each column's rows (MulColsParallel(nt), make_loop.jl)."""
    import numpy as np
    from oracle import cpu as ocpu
    from oracle import models as om
    from oracle import packing as opk
    from polfed_amd import spectral as spc
    H = om.spin_half_hamiltonian(op.L, op.bonds, op.couplings, op.anisotropies, op.fields,
                                 op.Lup)
    d, bk = opk.get_diags_and_offdiagonals_by_value(H)
    dr, bkr = opk.rescaled_packing(d, bk, rep.Emin, rep.Emax)
    P = ocpu.PackedCPU(dr, bkr)
    s = x0_host.shape[1]
    info = cpu_info()
    cores = info["usable_cpus"]
    c = spc.normalized_coefficients(spc.PolfedDefaults.coefficients, rep.target, rep.order)
    steps_app = rep.order + 1
    out = dict(info)
    # (a) one thread per column
    thr = min(s, cores)
    P.filter(c[:21], x0_host, nthreads=thr)  # warm-up: threads, first touch
    t0 = time.perf_counter()
    P.filter(c[:101], x0_host, nthreads=thr)
    est = (time.perf_counter() - t0) / 101
    n = int(min(steps_app, max(101, budget_s / est)))
    t0 = time.perf_counter()
    P.filter(c[:n], x0_host, nthreads=thr)
    per_step_cols = (time.perf_counter() - t0) / n
    # (b) all usable cores on each column's rows, columns in turn
    x1 = np.ascontiguousarray(x0_host[:, 0])
    P.filter_rows(c[:21], x1, nthreads=cores)
    t0 = time.perf_counter()
    P.filter_rows(c[:101], x1, nthreads=cores)
    est = (time.perf_counter() - t0) / 101 * s
    n2 = int(min(steps_app, max(101, budget_s / est)))
    t0 = time.perf_counter()
    P.filter_rows(c[:n2], x1, nthreads=cores)
    per_step_rows = (time.perf_counter() - t0) / n2 * s
    proj = lambda t: t * steps_app * rep.iterations
    out.update({
        "value": proj(per_step_cols), "unit": "s", "cores": thr, "kind": "port",
        "sample": (f"oracle/cpu_ref.c packed-bucket Clenshaw filter (FP64, the reference's CPU "
                   f"operator layout), one thread per column of the 16-column block: the first "
                   f"{n} of {steps_app} steps of one filter application, {per_step_cols * 1e3:.3f}"
                   f" ms/step, projected x (K+1) x {rep.iterations} Lanczos iterations of the "
                   f"FP64 solve (filter only; projected, not run)"),
        "all_cores": {"value": proj(per_step_rows), "cores": cores,
                      "sample": (f"the same filter with all {cores} usable cores splitting each "
                                 f"column's rows (columns in turn): the first {n2} steps of one "
                                 f"column, {per_step_rows * 1e3:.3f} ms per 16-column step, "
                                 f"projected the same way")}})
    return out


def xxz_l14_full_solves(dev):
    """The reference's CPU-runnable case, run in full on both sides: XXZ L=14
    Sz=0 (D=3432), W=5 seed 0, single vector, nev=50, target middle, reference
    defaults, eigentol 1e-10: the oracle's complete CPU solve (one thread for
    the sparse products) and the GPU FP64 solve, same start vector and probes."""
    import numpy as np
    import torch
    from oracle import models as om
    from oracle import polfed as opf
    from polfed_amd import FactorizationConfig, models, polfed
    from polfed_amd.blockops import fill_probes
    L, nev = 14, 50
    h = om.resolve_fields(L, 0.0, 5.0, rng=np.random.default_rng(0))
    op = models.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 5.0, fields=h, device=dev)
    H = om.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 5.0, fields=h)
    x0 = np.random.default_rng(100).standard_normal(op.D)
    x0 /= np.linalg.norm(x0)
    P = fill_probes(dev, op.D, 300, seed=0x5EED)

    def gpu():
        return polfed(op, x0, nev, ("rescaled", 0.0), produce_report=True,
                      fact=FactorizationConfig(eigentol=1e-10), probes=P)
    gpu()  # autotuning and first launches
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vals, _, rep = gpu()
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    ref = opf.polfed(lambda X: H @ X, x0, nev, ("rescaled", 0.0), eigentol=1e-10,
                     probes=dev.to_host(P))
    t_cpu = time.perf_counter() - t0
    same = ref.plan.order == rep.order and len(vals) == len(ref.vals)
    return {"workload": "XXZ L=14 Sz=0 D=3432 W=5 seed 0, single vector, nev=50, target "
                        "middle, eigentol 1e-10 (full solves, run)",
            "cpu_oracle_s": t_cpu, "cpu_threads": 1, "gpu_f64_s": t_gpu,
            "order": rep.order, "cpu_order": ref.plan.order, "pairs": int(len(vals)),
            "cpu_pairs": int(len(ref.vals)),
            "max_eig_diff_over_width": (float(np.max(np.abs(np.sort(vals) - ref.vals))
                                              / (rep.Emax - rep.Emin)) if same else None)}


def precision_bench(args):
    """The default bench: time to nev eigenpairs on the precision workload
    (XXZ L=22, block 16, nev=200) with the FP64 filter (the plain drop-in) and
    with the FP32-filter form, both complete solves in this run, every pair
    of the FP32-filter form against the FP64 run; the timed steps are whole
    FP32 filter applications of the solve's polynomial on the 16-column block
    (the hot path), which give `value` (algorithmic GB/s) and the roofline of
    the FP32 step kernel."""
    import numpy as np
    import torch
    from polfed_amd import FactorizationConfig, TransformConfig, polfed
    from polfed_amd.blockops import fill_probes
    from polfed_amd.device import Device
    t_start = time.perf_counter()
    dev = Device(0)
    op, x0, x0_host = precision_workload(dev)
    D, s, nev = op.D, 16, 200
    log(f"workload XXZ L=22 D={D} s={s} nev={nev}")
    probes = fill_probes(dev, D, 300, seed=0x22)  # the same DoS (hence order) for both solves
    fact = FactorizationConfig(eigentol=1e-10)
    sw = args.switch_tol if args.switch_tol is not None else TransformConfig.switch_tol
    solves = {}
    progress_filter(op, "l22")
    for label, tc in (("fp64", TransformConfig(dtype="f64")),
                      ("fp32_filter", TransformConfig(dtype="mixed", switch_tol=sw))):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vals, vecs, rep = polfed(op, x0, nev, "middle", produce_report=True, transform=tc,
                                 fact=fact, probes=probes)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        lz = rep.lanczos
        solves[label] = (t, vals, lz.residuals, rep)
        log(f"{label}: {t:.1f} s K={rep.order} iters={rep.iterations} pairs={len(vals)} "
            f"fp32 iters={lz.fast_iterations}")
        del vecs
    t64, v64, r64, rep64 = solves["fp64"]
    t32, v32, r32, rep32 = solves["fp32_filter"]
    width = rep64.Emax - rep64.Emin
    eerr, res32 = _pair_errors(v32, r32, v64, width)
    lz32 = rep32.lanczos
    ok32 = int(np.sum(res32 <= 1e-10))
    time_to_nev = {
        "workload": "XXZ L=22 Sz=0 D=705432 W=2 seed 2, block 16, nev=200, target middle, "
                    "eigentol 1e-10, same start block and KPM probes for both solves",
        "fp64": {"time_s": t64, "order": rep64.order, "iterations": rep64.iterations,
                 "pairs": int(len(v64)), "max_residual": float(r64.max()) if r64.size else None,
                 "stage_s": _stage_times(rep64)},
        "fp32_filter": {"time_s": t32, "mode": f"mixed (switch_tol {sw:g}) + FP64 polish",
                        "order": rep32.order, "iterations": rep32.iterations,
                        "fp32_iterations": int(lz32.fast_iterations),
                        "switch_iteration": int(lz32.switch_iteration),
                        "fp64_fallback": bool(lz32.fallback),
                        "polish_s": float(lz32.times.get("polish", 0.0)),
                        "pairs": int(len(v32)), "pairs_residual_le_1e-10": ok32,
                        "max_residual": float(res32.max()) if res32.size else None,
                        "stage_s": _stage_times(rep32)},
        "speedup_fp32_filter_vs_fp64": t64 / t32,
        "per_pair": {"eig_err_over_width": [float(f"{e:.2g}") for e in eerr],
                     "residual": [float(f"{r:.2g}") for r in res32]},
    }
    log(f"speed-up {t64 / t32:.3f}x, pairs {len(v32)} (<=1e-10: {ok32}) vs {len(v64)}")

    # ---- FP32 filters in every iteration (+ polish): the speed of the pure
    # FP32 filter and the residual level it leaves, when the run has time
    if time.perf_counter() - t_start < PURE_FP32_BEFORE_S:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vf, _, repf = polfed(op, x0, nev, "middle", produce_report=True,
                             transform=TransformConfig(dtype="f32"), fact=fact, probes=probes)
        torch.cuda.synchronize()
        tf = time.perf_counter() - t0
        lzf = repf.lanczos
        allv = lzf.all_vals if lzf.all_vals is not None else np.zeros(0)
        allr = lzf.all_residuals if lzf.all_residuals is not None else np.zeros(0)
        ef, _ = _pair_errors(allv, allr, v64, width)
        time_to_nev["fp32_throughout"] = {
            "time_s": tf, "speedup_vs_fp64": t64 / tf, "fp32_iterations": int(lzf.fast_iterations),
            "pairs_returned": int(len(vf)), "lanczos_converged": int(allv.size),
            "pairs_residual_le_1e-10": int(np.sum(allr <= 1e-10)),
            "residual_max": float(allr.max()) if allr.size else None,
            "residual_median": float(np.median(allr)) if allr.size else None,
            "eig_err_over_width_max": float(ef.max()) if ef.size else None,
            "eig_err_over_width_median": float(np.median(ef)) if ef.size else None}
        log(f"fp32 throughout: {tf:.1f} s, residuals {time_to_nev['fp32_throughout']['residual_max']}")
    else:
        time_to_nev["fp32_throughout"] = "skipped: the run had used its time budget"

    # ---- timed hot-path steps: FP32 filter applications of the solve's polynomial
    from polfed_amd import spectral as spc
    c = spc.normalized_coefficients(spc.PolfedDefaults.coefficients, rep64.target, rep64.order)
    y = torch.empty_like(x0)
    for _ in range(args.warmup):
        op.filter(c, x0, y, dtype="f32")
    torch.cuda.synchronize()
    events = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        op.filter(c, x0, y, dtype="f32")
        e1.record()
        events.append((e0, e1))
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t0
    launches = rep64.order + 1
    per_launch = 4 * s * D * 4
    ms_app = sum(a.elapsed_time(b) for a, b in events) / len(events)
    avg_launch_ms = ms_app / launches
    kname, klayout = kernel_name(op, "f32", s)
    traffic = measured_traffic(kname)
    value = per_launch * launches * args.steps / t_steps / 1e9
    achieved = per_launch / (avg_launch_ms * 1e-3) / 1e9
    log(f"steps: {args.steps} x {t_steps / args.steps:.3f} s, {value:.0f} GB/s, kernel {kname}")

    # ---- KPM moments in FP32 against FP64 on the same probes
    mu64 = op.kpm_moments(250, probes, dtype="f64")
    mu32 = op.kpm_moments(250, probes, dtype="f32")
    kerr = np.abs(mu32 - mu64) / D

    line = {"metric": METRIC, "value": value, "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": t_steps / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (numpy seeded fields and start block)",
            "config": {"workload": "XXZ L=22 Sz=0 D=705432 W=2 seed 2, block 16, nev=200, "
                                   "target middle; step = one FP32 Clenshaw filter application "
                                   f"(order {rep64.order}) on the 16-column block"},
            "time_to_nev": time_to_nev,
            "kpm_fp32_vs_fp64": {"moments": 250, "probes": 300,
                                 "max_err_over_D": float(kerr.max()),
                                 "median_err_over_D": float(np.median(kerr)),
                                 "per_moment_err_over_D": [float(f"{e:.1g}") for e in kerr]},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": traffic, "kernel": kname, "layout": klayout,
                         "per_launch_bytes": per_launch, "avg_launch_ms": avg_launch_ms}}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_l22(op, rep64, x0_host)
        line["cpu_baseline"]["xxz_l14_full_solve"] = xxz_l14_full_solves(dev)
    print(json.dumps(line, separators=(",", ":")), flush=True)


def xxz_l14(args):
    """The reference's CPU-runnable case: random-field XXZ L=14, Sz=0 (D=3432),
    W=5 seed 0, single-vector Lanczos, nev=50, target middle, reference
    defaults, eigentol 1e-10.  The GPU FP64 solve is timed against the
    oracle's full CPU solve on the same start vector and KPM probes (the
    eigenvalues are compared when both pick the same order), and the oracle's C
    restatement of the reference's CPU filter is timed on one thread, the way
    the reference runs single-vector filters (ParallelizationClenshaw.jl:55-70)."""
    import numpy as np
    import torch
    from oracle import cpu as ocpu
    from oracle import models as om
    from oracle import packing as opk
    from oracle import polfed as opf
    from polfed_amd import FactorizationConfig, models, polfed
    from polfed_amd import spectral as spc
    from polfed_amd.blockops import fill_probes
    from polfed_amd.device import Device
    dev = Device(0)
    L, nev = 14, 50
    h = om.resolve_fields(L, 0.0, 5.0, rng=np.random.default_rng(0))
    op = models.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 5.0, fields=h, device=dev)
    H = om.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 5.0, fields=h)
    x0 = np.random.default_rng(100).standard_normal(op.D)
    x0 /= np.linalg.norm(x0)
    P = fill_probes(dev, op.D, 300, seed=0x5EED)

    def run():
        return polfed(op, x0, nev, ("rescaled", 0.0), produce_report=True,
                      fact=FactorizationConfig(eigentol=1e-10), probes=P)
    run()  # warm-up: autotuning and first launches
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vals, vecs, rep = run()
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    psi = dev.to_host(vecs)
    width = rep.Emax - rep.Emin
    out = {"metric": METRIC, "workload": "XXZ L=14 Sz=0 D=3432 W=5 seed 0, single vector, "
                                         "nev=50, target middle, eigentol 1e-10",
           "unit": "s", "higher_is_better": False,
           "gpu_f64": {"time_s": t_gpu, "order": rep.order, "iterations": rep.iterations,
                       "returned_pairs": int(len(vals)), "max_residual": rep.eigenresidual,
                       "max_residual_vs_host_matrix": float(
                           np.linalg.norm(H @ psi - psi * vals, axis=0).max()) if len(vals) else None,
                       "times_s": rep.times}}
    log(f"gpu: {t_gpu:.3f} s K={rep.order} iters={rep.iterations} pairs={len(vals)}")
    t0 = time.perf_counter()
    ref = opf.polfed(lambda X: H @ X, x0, nev, ("rescaled", 0.0), eigentol=1e-10,
                     probes=dev.to_host(P))
    t_cpu = time.perf_counter() - t0
    same_k = ref.plan.order == rep.order
    out["cpu_oracle_full_solve"] = {
        "time_s": t_cpu, "kind": "port", "order": ref.plan.order,
        "iterations": ref.lanczos.numiter, "returned_pairs": int(len(ref.vals)),
        "threads": "1 for the sparse products (scipy CSR), numpy BLAS for the small dense algebra",
        "max_eig_diff_vs_gpu_over_width": (float(np.max(np.abs(np.sort(vals) - ref.vals)) / width)
                                           if same_k and len(vals) == len(ref.vals) else None),
        "speedup_gpu_vs_cpu": t_cpu / t_gpu}
    log(f"oracle full solve: {t_cpu:.2f} s K={ref.plan.order} pairs={len(ref.vals)}")
    d, bk = opk.get_diags_and_offdiagonals_by_value(H)
    dr, bkr = opk.rescaled_packing(d, bk, rep.Emin, rep.Emax)
    Pc = ocpu.PackedCPU(dr, bkr)
    c = spc.normalized_coefficients(spc.PolfedDefaults.coefficients, rep.target, rep.order)
    X1 = x0.reshape(-1, 1).copy()
    Pc.filter(c, X1, nthreads=1)  # warm-up
    t0 = time.perf_counter()
    napps = 5
    for _ in range(napps):
        Pc.filter(c, X1, nthreads=1)
    t_app = (time.perf_counter() - t0) / napps
    out["cpu_baseline"] = {
        "value": t_app * rep.iterations, "unit": "s", "cores": 1, "kind": "port",
        "sample": (f"oracle/cpu_ref.c packed-bucket Clenshaw filter on one thread (the "
                   f"reference's single-vector CPU filter is single-threaded): {napps} whole "
                   f"applications of {rep.order + 1} steps, {t_app * 1e3:.2f} ms each, x "
                   f"{rep.iterations} Lanczos iterations (filter only)")}
    print(json.dumps(out), flush=True)


def j1j2_l24(args):
    """J1-J2 L=24 Sz=0 (D=2704156), J1=1, J2=0.5: KPM with 32 random-phase
    probes and 2048 Chebyshev moments (Jackson), FP32 against FP64 on the same
    probes; then the DoS-chosen target (maximal density), the order for nev=300
    and one filter application on a 32-column block in FP32 and FP64, timed
    (the whole solve, 46 iterations of such filters, is projected from it)."""
    import numpy as np
    import torch
    from polfed_amd import TransformConfig, models
    from polfed_amd import blockops as bo
    from polfed_amd import spectral as spc
    from polfed_amd.configs import PolfedDefaults
    from polfed_amd.device import Device
    dev = Device(0)
    op = models.j1j2_hamiltonian(24, 12, 1.0, 0.5, 1.0, 1.0, 0.0, device=dev)
    # spectral bounds from the reference's Lanczos probe (unscaled operator)
    from polfed_amd.core import _extremal
    v0 = dev.to_device(np.random.default_rng(5).standard_normal(op.D))
    v0 = v0 / torch.linalg.vector_norm(v0)
    inner_apply = op.apply
    napply = {"n": 0}

    def apply_logged(X, Y=None, rescaled=False, dtype="f64"):
        napply["n"] += 1
        if napply["n"] % 100 == 0:
            log(f"l24 bounds: {napply['n']} operator applications")
        return inner_apply(X, Y, rescaled=rescaled, dtype=dtype)

    op.apply = apply_logged
    t0 = time.perf_counter()
    Emin, Emax = _extremal(dev, op, v0, "SR"), _extremal(dev, op, v0, "LR")
    t_bounds = time.perf_counter() - t0
    op.apply = inner_apply
    log(f"l24 bounds [{Emin:.6f}, {Emax:.6f}] in {t_bounds:.1f} s")
    op.set_rescale(Emin, Emax)
    P = bo.fill_probes(dev, op.D, 32, seed=24, kind="phase")
    out = {"metric": METRIC, "workload": "J1-J2 L=24 Sz=0 D=2704156 J1=1 J2=0.5; KPM R=32 "
           "random-phase probes, N=2048 Jackson; nev=300 at the maximal-density target, block 32",
           "Emin": Emin, "Emax": Emax, "bounds_s": t_bounds}
    mus = {}
    for dt in ("f64", "f32"):
        op.ensure_tuned(32, dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mus[dt] = op.kpm_moments(2048, P, dtype=dt)
        torch.cuda.synchronize()
        out[f"kpm_{dt}_s"] = time.perf_counter() - t0
        log(f"kpm {dt}: {out[f'kpm_{dt}_s']:.2f} s")
    d = np.abs(mus["f32"] - mus["f64"]) / op.D
    out["kpm_moment_err_max_over_D"] = float(d.max())
    out["kpm_moment_err_median_over_D"] = float(np.median(d))
    out["kpm_per_moment_err_over_D"] = [float(f"{e:.1g}") for e in d]
    a, b = (Emax - Emin) / 2, (Emax + Emin) / 2
    rho = spc.DensityOfStates(mus["f64"], "Jackson")
    plan = spc.build_transform_plan(TransformConfig(), a, b, 300, "maxdos")
    spc.getbounds(plan, a, b, rho, "mean")
    spc.getorderofexpansion(plan)
    coeffs = spc.normalized_coefficients(plan.coefficients, plan.target, plan.order)
    out.update({"target": plan.target, "order": plan.order,
                "maxdim": PolfedDefaults.expectedkrylovdim(300, 32, 1.25)})
    q = torch.randn((32, op.D), dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(9))
    y = torch.empty_like(q)
    iters = -(-out["maxdim"] // 32)
    if args.j1j2_iters > 0:
        out["partial_solve"] = j1j2_partial_solve(dev, op, coeffs, args.j1j2_iters, iters)
        print(json.dumps(out), flush=True)
        return
    for dt in ("f32", "f64"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.filter(coeffs, q, y, dtype=dt)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        steps = plan.order + 1
        out[f"filter_{dt}"] = {"time_s": t, "us_per_step": t / steps * 1e6,
                               "algorithmic_GBps": 4 * 32 * op.D * (4 if dt == "f32" else 8)
                               * steps / t / 1e9,
                               "projected_solve_filter_s": t * iters}
        log(f"filter {dt}: {t:.1f} s ({t / steps * 1e6:.1f} us/step)")
    print(json.dumps(out), flush=True)


def j1j2_partial_solve(dev, op, coeffs, n_iter, full_iters):
    """The first ``n_iter`` iterations of the nev=300 block-32 Lanczos run on
    the FP32 filter (FP64 three-term step, CGS2 reorthogonalization and
    convergence checks as in polfed), timed per stage; the whole run of
    ``full_iters`` iterations is projected by scaling the measured time per
    iteration (the reorthogonalization, which grows with the basis, is under
    1 % of it)."""
    import numpy as np
    import torch
    from polfed_amd.lanczos import lanczos
    x0 = torch.randn((32, op.D), dtype=torch.float64, device="cuda",
                     generator=torch.Generator(device="cuda").manual_seed(11))
    x0 = torch.linalg.qr(x0.T)[0].T.contiguous()
    done = {"n": 0}

    def f32_filter(X, Y):
        op.filter(coeffs, X, Y, dtype="f32")
        done["n"] += 1
        log(f"l24 partial solve: filter {done['n']}/{n_iter}")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = lanczos(dev, f32_filter, x0, 300, maxdim=32 * n_iter, which="LR", tol=1e-14,
                  eigentol=1e-10, mapvals=op)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    per_filter = res.times.get("mapping", 0.0) / max(res.numiter, 1)
    return {"iterations": int(res.numiter), "krylov_dim": int(res.krylovdim), "time_s": t,
            "stage_s": {k: float(v) for k, v in res.times.items()},
            "filter_s_per_iteration": per_filter,
            "lanczos_converged_at_stop": int(res.converged),
            "projected_full_solve_s": t / max(res.numiter, 1) * full_iters,
            "note": (f"{res.numiter} of {full_iters} iterations run; the projection scales "
                     "the measured time per iteration (filter, reorthogonalization at the "
                     "sizes reached, checks) to the whole run")}


def copy_calibration():
    """Plain device copy of a known size (256 MiB) for calibrating the
    FETCH_SIZE / WRITE_SIZE byte counters."""
    import torch
    a = torch.ones(64 * 2**20, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"mode": "copy-calibration", "bytes_per_copy": a.numel() * 4}), flush=True)


def _switch(args):
    """TransformConfig keywords for --switch-tol / --polish-fraction overrides."""
    kw = {}
    if getattr(args, "switch_tol", None) is not None:
        kw["switch_tol"] = args.switch_tol
    if getattr(args, "polish_fraction", None) is not None:
        kw["polish_fraction"] = args.polish_fraction
    return kw


def solve(op, x0, fp="f64", basistype=None, switch_tol=None):
    from polfed_amd import DoSConfig, FactorizationConfig, TransformConfig, polfed
    tc = TransformConfig(dtype=fp) if switch_tol is None else TransformConfig(dtype=fp, switch_tol=switch_tol)
    return polfed(op, x0, 100, "middle", produce_report=True,
                  transform=tc,
                  fact=FactorizationConfig(eigentol=1e-10, basistype=basistype), dos=DoSConfig())


def cpu_baseline(op, report, x0_host, budget_s=15.0):
    """The oracle's C restatement of the reference's packed CPU filter, timed on
    a bounded sample of the solve's filter work -- whole filter applications
    (K+1 steps on the block), as many as fit about ``budget_s`` seconds, or a
    prefix of one when a single application would take longer -- projected
    to the full solve: t_step * (K+1) * iterations."""
    import numpy as np
    import scipy.sparse as sp
    from oracle import cpu as ocpu
    from oracle import models as om
    from oracle import packing as opk
    from polfed_amd import spectral as spc
    H = om.spin_half_hamiltonian(op.L, op.bonds, op.couplings, op.anisotropies, op.fields,
                                 op.Lup)
    d, bk = opk.get_diags_and_offdiagonals_by_value(H)
    dr, bkr = opk.rescaled_packing(d, bk, report.Emin, report.Emax)
    P = ocpu.PackedCPU(dr, bkr)
    s = x0_host.shape[1]
    threads = min(s, os.cpu_count() or 1)
    c = spc.normalized_coefficients(spc.PolfedDefaults.coefficients, report.target, report.order)
    steps_app = report.order + 1
    P.filter(c[:51], x0_host, nthreads=threads)  # warm-up (threads, first touch)
    t0 = time.perf_counter()
    P.filter(c[:201], x0_host, nthreads=threads)  # estimate on a 201-step prefix
    est_app = (time.perf_counter() - t0) / 201 * steps_app
    if est_app <= budget_s:
        napps = max(1, min(4, int(np.ceil(budget_s / est_app))))
        t0 = time.perf_counter()
        for _ in range(napps):
            P.filter(c, x0_host, nthreads=threads)
        t = time.perf_counter() - t0
        nsteps = napps * steps_app
        what = f"{napps} whole filter application(s) ({napps} x {steps_app} steps)"
    else:
        nsteps = int(max(201, budget_s / est_app * steps_app))
        t0 = time.perf_counter()
        P.filter(c[:nsteps], x0_host, nthreads=threads)
        t = time.perf_counter() - t0
        what = f"the first {nsteps} of {steps_app} steps of one filter application"
    per_step = t / nsteps
    projected = per_step * steps_app * report.iterations
    return {"value": projected, "unit": "s", "cores": threads, "kind": "port",
            "sample": (f"oracle/cpu_ref.c packed-bucket Clenshaw filter (MulColsParallel: one "
                       f"thread per column) on the {s}-column block: {what} in {t:.2f} s = "
                       f"{per_step * 1e3:.3f} ms/step, projected x (K+1) x "
                       f"{report.iterations} Lanczos iterations (filter only)")}


def plain_bench(args):
    """``--workload xxz-l20``: one complete FP64 solve of XXZ L=20 (D=184756,
    W=3 seed 1, block 8, nev=100) per step, with FP32-filter forms beside."""
    import numpy as np
    import torch
    from polfed_amd.device import Device

    dev = Device(0)
    op, x0 = build_workload(dev)
    log(f"workload XXZ L=20 D={op.D} s=8 nev=100")
    timer = FilterTimer(op)
    for w in range(args.warmup):
        t0 = time.perf_counter()
        solve(op, x0)
        log(f"warmup {w} {time.perf_counter() - t0:.2f} s")
    timer.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reports = []
    for i in range(args.steps):
        vals, vecs, rep = solve(op, x0)
        reports.append(rep)
        log(f"step {i}: K={rep.order} iters={rep.iterations} conv={rep.eigenconverged} "
            f"maxres={rep.eigenresidual:.2e}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step, bytes_step, nsteps, filter_ms = timer.summary(op.D, 8)
    rep = reports[-1]
    achieved = bytes_step / (ms_step * 1e-3) / 1e9
    kname, klayout = kernel_name(op)
    traffic = measured_traffic(kname)
    out = {
        "metric": METRIC,
        "value": elapsed / args.steps,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (numpy seeded fields and start block)",
        "config": {"workload": "XXZ L=20 Sz=0 D=184756 W=3 seed 1, block 8, nev=100, "
                               "target middle, eigentol 1e-10, FP64 filter (plain drop-in)"},
        "solve": {"order": rep.order, "iterations": rep.iterations,
                  "eigenconverged": rep.eigenconverged, "max_residual": rep.eigenresidual,
                  "Emin": rep.Emin, "Emax": rep.Emax, "times_s": rep.times},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "kernel": kname, "layout": klayout,
                     "per_launch_bytes": bytes_step, "avg_launch_ms": ms_step,
                     "filter_share": filter_ms * 1e-3 / elapsed},
    }
    if not args.no_fp32:
        import numpy as np
        width = rep.Emax - rep.Emin
        ref_vals = np.sort(vals)
        for mode in args.modes_l20.split(","):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            vm, _, rm = solve(op, x0, mode, switch_tol=args.switch_tol)
            torch.cuda.synchronize()
            em = time.perf_counter() - t1
            lz = rm.lanczos
            allv = np.sort(lz.all_vals) if lz.all_vals is not None else np.zeros(0)
            n = min(len(allv), len(ref_vals))
            entry = {
                "time_s": em, "speedup_vs_fp64": (elapsed / args.steps) / em,
                "returned_pairs": int(len(vm)), "lanczos_converged": int(lz.converged),
                "max_residual_all": float(lz.all_residuals.max()) if lz.all_residuals is not None else None,
                "max_eig_err_vs_fp64_over_width": float(np.max(np.abs(allv[:n] - ref_vals[:n])) / width) if n else None,
                "iterations": int(lz.numiter), "fp32_iterations": int(lz.fast_iterations),
                "switch_iteration": int(lz.switch_iteration), "fp64_fallback": bool(lz.fallback),
            }
            out[f"fp32_filter_{mode}"] = entry
            log(f"{mode}: {em:.2f} s pairs={len(vm)} fp32 iters={lz.fast_iterations}/{lz.numiter} "
                f"maxres={entry['max_residual_all']}")
    if args.hybrid_basis >= 0:
        import numpy as np
        from polfed_amd import HybridMatrixBasis
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        vh, _, rh = solve(op, x0, basistype=HybridMatrixBasis(max_device_vectors=args.hybrid_basis))
        torch.cuda.synchronize()
        eh = time.perf_counter() - t1
        lz = rh.lanczos
        n = min(len(vh), len(vals))
        out["hybrid_basis"] = {
            "time_s": eh, "slowdown_vs_device_basis": eh / (elapsed / args.steps),
            "device_vectors": int(lz.basis_device_vectors), "host_vectors": int(lz.basis_host_vectors),
            "krylovdim": int(lz.krylovdim), "returned_pairs": int(len(vh)),
            "max_residual": float(rh.eigenresidual),
            "max_eig_diff_over_width": float(np.max(np.abs(np.sort(vh)[:n] - np.sort(vals)[:n]))
                                             / (rep.Emax - rep.Emin)) if n else None,
            "reorth_s": rh.times.get("lanczos.reorthogonalization"),
            "device_basis_reorth_s": rep.times.get("lanczos.reorthogonalization"),
        }
        log(f"hybrid basis: {eh:.2f} s ({lz.basis_device_vectors} device + "
            f"{lz.basis_host_vectors} host columns)")
    if not args.no_cpu:
        q = dev.to_host(x0)
        out["cpu_baseline"] = cpu_baseline(op, rep, q)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp32", action="store_true",
                    help="skip the FP32-filter solves (mixed and pure FP32)")
    ap.add_argument("--filter-only", type=int, default=0, metavar="STEPS",
                    help="profiling driver: time about STEPS filter steps only")
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"],
                    help="filter arithmetic for --filter-only")
    ap.add_argument("--packed-compare", action="store_true",
                    help="profiling driver: packed sparse operator vs matrix-free filter steps")
    ap.add_argument("--layout", default="", metavar="V,PBLOCKS,WIDE,PARTITION",
                    help="--filter-only: fixed step-kernel launch configuration instead of "
                         "the autotuner's (e.g. 2,8,0,1)")
    ap.add_argument("--j1j2-iters", type=int, default=0, metavar="N",
                    help="--workload j1j2-l24: run the first N iterations of the nev=300 "
                         "solve on the FP32 filter (timed per stage) instead of the FP32 and "
                         "FP64 single filter applications")
    ap.add_argument("--batch-form", default="auto", choices=["auto", "plain", "lds", "rowvec"],
                    help="--workload qsun-batch --filter-only: the multi-instance kernel form "
                         "(autotuned, or rows from memory / block rows in LDS / row vectors)")
    ap.add_argument("--row-partition", type=int, default=-1, choices=[-1, 0, 1],
                    help="--filter-only: spin-chain row order (-1 autotuned, 0 reference, "
                         "1 XCD partition)")
    ap.add_argument("--reorth-only", type=int, default=0, metavar="CALLS",
                    help="profiling driver: time CALLS CGS2 reorthogonalizations")
    ap.add_argument("--reorth-shape", default="l20", choices=sorted(REORTH_SHAPES),
                    help="--reorth-only: (D, k, s) of the XXZ L=20, XXZ L=22 or J1-J2 L=24 "
                         "solve's last reorthogonalization")
    ap.add_argument("--workload", default="xxz-l22",
                    choices=["xxz-l22", "xxz-l20", "xxz-l14", "qsun-batch", "j1j2-l24"],
                    help="xxz-l22: the bench line (FP64 and FP32-filter solves, FP32 filter "
                         "steps); xxz-l20: the plain FP64 solve as the step; xxz-l14: the "
                         "CPU-runnable case against the oracle's full CPU solve; qsun-batch: 64 "
                         "Quantum-Sun instances; j1j2-l24: KPM and filter at L=24")
    ap.add_argument("--instances", type=int, default=64, help="qsun-batch: number of instances")
    ap.add_argument("--modes", default="mixed,f64",
                    help="qsun-batch: comma list of mixed, f64 (batched) and plain (FP64, "
                         "one polfed call per instance)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="profiling driver: one 256 MiB device copy, five times")
    ap.add_argument("--switch-tol", type=float, default=None,
                    help="mixed-precision solves: residual estimate at which the filter "
                         "switches to FP32 (default: TransformConfig's)")
    ap.add_argument("--polish-fraction", type=float, default=None,
                    help="FP32-filter solves: degree of the FP64 polish window as a fraction "
                         "of the filter order (default: TransformConfig's)")
    ap.add_argument("--modes-l20", default="mixed,f32",
                    help="FP32-filter solves reported beside the FP64 bench solve")
    ap.add_argument("--hybrid-basis", type=int, default=-1, metavar="DEVICE_VECTORS",
                    help="also solve with a hybrid basis of this many device columns "
                         "(the rest in pinned host memory) and report its cost")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "single-GPU build: only --gpus 1 is supported; "
                                   "the multi-GPU sharding is described, not built"}))
        sys.exit(2)

    _load_package()
    if args.copy_calibration:
        return copy_calibration()
    if args.filter_only:
        return filter_only(args)
    if args.reorth_only:
        return reorth_only(args)
    if args.packed_compare:
        return packed_compare(args)
    if args.workload == "qsun-batch":
        return qsun_batch(args)
    if args.workload == "xxz-l22":
        return precision_bench(args)
    if args.workload == "j1j2-l24":
        return j1j2_l24(args)
    if args.workload == "xxz-l14":
        return xxz_l14(args)
    return plain_bench(args)


if __name__ == "__main__":
    main()
