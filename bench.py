#!/usr/bin/env python3
"""Benchmark: Lanczos SpMV on one MI355X (XXZ Heisenberg chain N=24).

Workload: XXZ chain, N=24 sites, periodic, Jxy = Jz = 1, Sz = 0 sector
(n = 2,704,156 states, nnz = 36,564,892), complex128 CSR built on the GPU.
A step is one product y = H v of the Lanczos iteration over that matrix,
with v resident in HBM.  Algorithmic bytes per product (int32 indices,
complex128 values, x read once, y written once):
    B = 20 nnz + 4 (n + 1) + 32 n = 828,647,460 bytes.
value = B * steps / wall time of the timed steps (GB/s, 1e9 bytes).
roofline.achieved = B / average kernel duration from HIP events on the
stream the kernel runs on; roofline.traffic = HBM bytes per launch from the
committed rocprofv3 counter summary of the same kernel source (null if the
source changed since).  The line also carries the full Lanczos ground
state (k = 150, full reorthogonalisation) against the Bethe-ansatz energy,
and the oracle's single-thread CPU product timed on a bounded sample (with
the host's CPU model and thread count), plus the other CPU rows
(cpu_baseline_more: XXZ N=24 ground state, Hubbard matrix-free product, on
bounded samples with labelled projections; --cpu-full times them in full).
Further sections (each can be skipped): the Hubbard 4x4 quench (ground state
then real-time steps with the matrix-free Kronecker product), 4096 batched
quantum-jump trajectories of the dissipative Hubbard chain L=8 against the
one-at-a-time loop and the oracle (the reference's loop, operators rebuilt
every step, and a precomputed-operator variant), the served request stream
(four workers, one worker, one request at a time with the same rules, full
reorthogonalisation) against the plain one-at-a-time path, and extras:
a basis filtered by double occupancy and a qubits-times-mode matrix (build
and product).  The Lanczos section also runs the reference's two-vector
solver (Hermitian::getSolver) on the same operator and start, and the quench
section the same for its ground state.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
E_BETHE_XXX24 = -10.67001451653724  # Bethe ansatz, N=24 PBC (S=1/2 XXX)
SPMV_DEFAULT_MAPPING = -13  # the row mapping the library picks for this matrix (qd_linop_get_lanes)
TRAFFIC_PROFILE = os.path.join("profiles", "spmv_xxz24_traffic.txt")


def spmv_traffic():
    """HBM bytes per launch of the default product kernel at XXZ N=24 from
    the committed counter summary (two rocprofv3 --pmc runs of this script:
    FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a plain 1 GiB copy).
    The summary's record line names the sha256 of the kernel source it was
    measured on; the figure is used only while spmv.hip is that source, else
    the field is null.  Returns (bytes or None, note)."""
    import hashlib
    src = os.path.join(ROOT, "qudrip_amd", "csrc", "spmv.hip")
    try:
        sha = hashlib.sha256(open(src, "rb").read()).hexdigest()
        rec = [ln for ln in open(os.path.join(ROOT, TRAFFIC_PROFILE)) if ln.startswith("traffic-record:")]
    except OSError as e:
        return None, f"no counter summary ({e.strerror})"
    if not rec:
        return None, f"{TRAFFIC_PROFILE} has no traffic-record line"
    kv = dict(f.split("=", 1) for f in rec[-1].split()[1:])
    if kv.get("spmv.hip-sha256") != sha:
        return None, f"{TRAFFIC_PROFILE} was measured on another spmv.hip (sha256 {kv.get('spmv.hip-sha256', '?')[:12]})"
    return float(kv["bytes_per_launch"]), (f"{TRAFFIC_PROFILE}: rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of "
                                           f"{kv.get('kernel', '?')}, same spmv.hip (sha256 {sha[:12]})")


def host_cpu():
    """CPU model (lscpu's 'Model name', read from /proc/cpuinfo), nproc and
    the threads this process may use."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.lower().startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_threads": usable}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=24)
    ap.add_argument("--krylov", type=int, default=150)
    ap.add_argument("--no-lanczos", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-quench", action="store_true", help="skip the Hubbard 4x4 quench workload")
    ap.add_argument("--quench-steps", type=int, default=200)
    ap.add_argument("--no-traj", action="store_true", help="skip the dissipative trajectory workload")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the Nd-filtered basis and qubits-times-mode measurements")
    ap.add_argument("--serve-requests", type=int, default=48, help="requests in the Poisson stream")
    ap.add_argument("--serve-workers", type=int, default=6)
    ap.add_argument("--traj", type=int, default=4096, help="number of trajectories")
    ap.add_argument("--traj-cpu", type=int, default=64,
                    help="trajectories in the oracle's CPU sample (a quarter of them in the reference-loop sample)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--lanes", type=int, default=None, help="SpMV row mapping override (<= 0: CSR-stream variant)")
    ap.add_argument("--serve-profile", action="store_true",
                    help="serve the request stream once (default scheme) and exit: a run to profile the served path")
    ap.add_argument("--hbm-probe-mb", type=int, default=0,
                    help="measure the streaming read / copy bandwidth over this many MiB and exit")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration) and exit")
    ap.add_argument("--kron-products", type=int, default=0,
                    help="only apply the Hubbard 4x4 two-block matrix-free product this many times and exit "
                         "(a run to collect the product's counters)")
    ap.add_argument("--traj-steps-only", type=int, default=0,
                    help="only run the batched dissipative trajectory workload for this many time steps (T = steps "
                         "x 0.05) and exit (a run to collect the trajectory engine's counters)")
    ap.add_argument("--trace-summary", default="",
                    help="summarise a rocprofv3 --kernel-trace output directory (dispatches, span, GPU busy as "
                         "the union of the kernel intervals), print one JSON line and exit (no GPU)")
    ap.add_argument("--trace-window", default="",
                    help="with --trace-summary: only from the first to the last dispatch of a kernel whose name "
                         "starts with this prefix (FIRST..LAST: two prefixes), with the kernel time per kernel in "
                         "that window")
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time the CPU restatement on the full-size workloads: the XXZ N=24 Lanczos "
                         "ground state (k = 150) and one Hubbard 4x4 matrix-free product (minutes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for {args.gpus}); "
              "sharding over several GPUs is designed, not implemented", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.hbm_probe_mb:
        # streaming ceiling: non-temporal read and copy kernels of known size
        import ctypes

        from qudrip_amd.runtime import Context
        ctx = Context(0)
        rd, cp = ctypes.c_double(), ctypes.c_double()
        ctx.check(ctx.lib.qd_stream_bandwidth(ctx.ptr, args.hbm_probe_mb << 20, 20, ctypes.byref(rd),
                                              ctypes.byref(cp)), "stream_bandwidth")
        print(json.dumps({"probe_bytes": args.hbm_probe_mb << 20, "read_GBs": round(rd.value, 1),
                          "copy_GBs": round(cp.value, 1), "peak_GBs": 8000.0}), flush=True)
        return

    if args.serve_profile:
        from qudrip_amd import serving as SV
        reqs = SV.request_stream(args.serve_requests, rate=200.0, seed=4)
        out = SV.run_served(reqs, workers=args.serve_workers)
        print(json.dumps({"served_requests": len(reqs), "workers": args.serve_workers,
                          "solves_per_s": round(out["solves_per_s"], 3), "span_s": round(out["span_s"], 3),
                          "failed": sum(1 for x in out["results"].values() if x["status"] != 0)}), flush=True)
        return

    if args.trace_summary:
        print(json.dumps(trace_summary(args.trace_summary, args.trace_window)), flush=True)
        return

    if args.kron_products:
        kron_products(args.kron_products)
        return

    if args.traj_steps_only:
        from qudrip_amd import workloads as WL
        from qudrip_amd.runtime import Context
        log = lambda m: print("[bench] " + m, file=sys.stderr, flush=True)
        r = WL.hubbard_trajectories(Context(0), ntraj=args.traj, T=args.traj_steps_only * 0.05, log=log)
        print(json.dumps({"trajectories": args.traj, "steps": r["nsteps"], "run_s": round(r["run_s"], 3),
                          "jumps": r["jumps"], "losses": r["losses"],
                          "krylov_form": "lanczos" if r["lanczos"] else "arnoldi"}))
        return

    if args.calib_copy_mb:
        # plain copy kernel of known size: reads and writes calib_copy_mb MiB each
        a = torch.ones(args.calib_copy_mb * (1 << 20) // 8, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": a.numel() * 8}), flush=True)
        return

    from qudrip_amd import operators as Q
    from qudrip_amd.runtime import CSR, Basis, Context, Krylov, LinOp, TermProgram

    ctx = Context(0)
    N = args.N
    idx, ham = Q.xxz_chain(N, 1.0, 1.0, pbc=True)
    sec = Q.Constrain(idx, Q.getNset(idx) == N // 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    basis = Basis(ctx, sec)
    prog = TermProgram(ctx, ham, N)
    csr = CSR(ctx, basis, prog)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    n, nnz = csr.n, csr.nnz
    op = LinOp.from_csr(ctx, csr)
    if args.lanes is not None:
        op.set_lanes(args.lanes)
    B = 20 * nnz + 4 * (n + 1) + 32 * n

    x = ctx.seeded_state(n, 2)
    y = ctx.zeros(n)
    for _ in range(args.warmup):
        op.apply(x, y)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        op.apply(x, y)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = B * args.steps / wall / 1e9
    achieved = B / (kern_ms * 1e-3) / 1e9
    print(f"[bench] N={N} n={n} nnz={nnz} build={build_s*1e3:.1f} ms  spmv {kern_ms*1e3:.1f} us/launch "
          f"= {achieved:.0f} GB/s (events), wall {value:.0f} GB/s", file=sys.stderr, flush=True)

    result = {
        "metric": "Lanczos SpMV GB/s vs HBM peak; solves/sec + p99 latency, served vs plain",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded splitmix64 start vector, seed 2)",
        "config": {"workload": f"XXZ chain N={N} PBC Jxy=Jz=1 Sz=0, complex128 CSR SpMV in Lanczos",
                   "n": n, "nnz": nnz, "bytes_per_spmv": B,
                   "spmv_row_mapping": op.lanes()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "csr_build_ms": round(build_s * 1e3, 2),
    }
    if N == 24 and op.lanes() == SPMV_DEFAULT_MAPPING:
        traffic, note = spmv_traffic()
        result["roofline"]["traffic"] = traffic
        result["roofline_traffic_source"] = note

    if not args.no_lanczos:
        kr = Krylov(ctx, n, args.krylov, herm=True)
        psi = ctx.seeded_state(n, 2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0 = kr.ground_state(op, psi)
        torch.cuda.synchronize()
        solve_s = time.perf_counter() - t0
        ref = E_BETHE_XXX24 if N == 24 else None
        result["lanczos"] = {
            "k": args.krylov, "E0": e0, "E_ref": ref,
            "rel_err": (abs(e0 - ref) / abs(ref)) if ref else None,
            "solve_s": round(solve_s, 4), "solves_per_sec": round(1.0 / solve_s, 3),
            "reorth": "delayed CGS2 (full)",
        }
        print(f"[bench] lanczos k={args.krylov}: E0={e0:.14f} solve {solve_s:.3f} s", file=sys.stderr, flush=True)
        kr.close()
        # Hermitian::getSolver's two-vector Lanczos (no reorthogonalisation,
        # two passes of the recurrence) on the same operator and start
        from qudrip_amd.runtime import HermitianLanczos
        hl = HermitianLanczos(ctx, n, args.krylov)
        psi = ctx.seeded_state(n, 2)
        hl.ground_state(op, psi)  # warm-up (first launches of its kernels)
        psi = ctx.seeded_state(n, 2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0h = hl.ground_state(op, psi)
        torch.cuda.synchronize()
        hs = time.perf_counter() - t0
        inf = hl.info()
        steps = 2 * inf["dim"] - 1  # recurrence steps over both passes (each one product)
        result["hermitian_lanczos"] = {
            "k": args.krylov, "E0": e0h, "rel_err": (abs(e0h - ref) / abs(ref)) if ref else None,
            "coefficients": inf["dim"], "solve_s": round(hs, 4), "ms_per_step": round(hs * 1e3 / steps, 4),
            "bytes_per_step": B + 6 * 16 * n,
            "step_GBs": round((B + 6 * 16 * n) * steps / hs / 1e9, 1),
            "reorth": "none (Hermitian::LanczosSolver, two passes)"}
        print(f"[bench] hermitian lanczos k={args.krylov}: E0={e0h:.14f} solve {hs:.3f} s "
              f"({hs * 1e3 / steps:.3f} ms/step)", file=sys.stderr, flush=True)
        hl.close()

    # sections start from a clean device heap: cached blocks of the earlier
    # sections are returned first (a server that starts after heavy
    # allocation churn in its process serves measurably slower)
    if not args.no_serve:
        torch.cuda.empty_cache()
        result["serving"] = serving(ctx, args)

    if not args.no_quench:
        torch.cuda.empty_cache()
        from qudrip_amd import workloads as WL
        log = lambda m: print("[bench] " + m, file=sys.stderr, flush=True)
        q = WL.hubbard_quench(ctx, 4, 4, 5, 5, U0=4.0, U1=2.0, k_gs=120, k_rt=30, dt=0.05,
                              steps=args.quench_steps, matfree=True, log=log, two_vector_gs=True)
        d = q["docc"]
        result["hubbard4x4_quench"] = {
            "n": q["n"], "E0_U4": q["E0"], "gs_s": round(q["gs_s"], 3), "k_gs": 120,
            "rt_steps": q["steps"], "k_rt": 30, "dt": 0.05, "rt_s": round(q["rt_s"], 3),
            "rt_step_ms": round(q["rt_step_ms"], 2),
            "docc_t": {f"{t:.2f}": round(d[int(round(t / 0.05))], 10) for t in (0.0, 2.5, 5.0, 7.5, 10.0)
                       if int(round(t / 0.05)) < len(d)},
            "energy_drift": abs(q["energy_end"] - q["energy_start"]), "norm_end": q["norm_end"],
            "operator": "matrix-free (two-block Kronecker form)"}
        tv = q["two_vector_gs"]
        if tv:
            result["hubbard4x4_quench"]["two_vector_gs"] = {
                "E0": tv["E0"], "solve_s": round(tv["solve_s"], 3), "coefficients": tv["coefficients"],
                "rel_diff_vs_dcgs2": abs(tv["E0"] - q["E0"]) / abs(q["E0"]),
                "rayleigh_rel_diff": abs(tv["rayleigh"] - tv["E0"]) / abs(tv["E0"])}
        log(f"quench: {q['steps']} steps in {q['rt_s']:.2f} s ({q['rt_step_ms']:.1f} ms/step), "
            f"energy drift {abs(q['energy_end'] - q['energy_start']):.2e}")
        # the same quench with the real-time basis kept semi-orthogonal by
        # partial reorthogonalisation (a labelled second row; the row above,
        # fully reorthogonalised, stays the figure)
        q2 = WL.hubbard_quench(ctx, 4, 4, 5, 5, U0=4.0, U1=2.0, k_gs=120, k_rt=30, dt=0.05,
                               steps=args.quench_steps, matfree=True, rt_method=2)
        dd = max(abs(a - b) for a, b in zip(q2["docc"], q["docc"]))
        dpsi = float(torch.linalg.vector_norm(q2["psi"] - q["psi"]) / torch.linalg.vector_norm(q["psi"]))
        result["hubbard4x4_quench"]["partial_reorth"] = {
            "rt_s": round(q2["rt_s"], 3), "rt_step_ms": round(q2["rt_step_ms"], 2),
            "full_passes": q2["rt_full_steps"], "max_abs_docc_diff": dd, "final_state_rel_diff": dpsi,
            "energy_drift": abs(q2["energy_end"] - q2["energy_start"])}
        log(f"quench, partial reorthogonalisation: {q2['rt_step_ms']:.1f} ms/step, {q2['rt_full_steps']} full "
            f"passes, max |d docc| {dd:.1e}, final state rel. diff {dpsi:.1e}")
        del q, q2

    if not args.no_traj:
        torch.cuda.empty_cache()
        result["hubbard_l8_trajectories"] = trajectories(ctx, args)

    if not args.no_extras:
        result["extras"] = extras(ctx)

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(csr, n, nnz, B, x, args.cpu_seconds)
        result["cpu_baseline_more"] = cpu_more(csr, n, args)

    print(json.dumps(result), flush=True)


def kernel_short_name(name):
    """`void (anonymous namespace)::k<4, 8>(args)` -> `k<4, 8>`."""
    n = name.replace("void ", "").replace("(anonymous namespace)::", "")
    depth = 0
    for i, c in enumerate(n):
        if c == "<":
            depth += 1
        elif c == ">":
            depth -= 1
        elif c == "(" and depth == 0:
            return n[:i]
    return n


def trace_summary(outdir, window=""):
    """Dispatch count, span and GPU busy time of a rocprofv3 kernel trace
    (`run_kernel_trace.csv`, or its gzip, under `outdir`): busy is the length
    of the union of the kernel intervals, so kernels that overlap on several
    streams count once.  window: only the dispatches from the first to the
    last one whose kernel name starts with this prefix (one engine's run
    inside a longer program), with the kernel time per kernel there;
    "FIRST..LAST": from the first dispatch whose name starts with FIRST to
    the last one whose name starts with LAST."""
    import csv
    import gzip
    path = os.path.join(outdir, "run_kernel_trace.csv")
    fh = open(path, newline="") if os.path.exists(path) else gzip.open(path + ".gz", "rt", newline="")
    with fh:
        rows = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), kernel_short_name(r["Kernel_Name"]))
                      for r in csv.DictReader(fh))
    if window:
        first, _, last = window.partition("..")
        last = last or first
        i0 = next((i for i, r in enumerate(rows) if r[2].startswith(first)), None)
        i1 = next((i for i in range(len(rows) - 1, -1, -1) if rows[i][2].startswith(last)), None)
        rows = rows[i0:i1 + 1] if i0 is not None and i1 is not None and i1 >= i0 else []
    if not rows:
        return {"dispatches": 0, "span_s": 0.0, "busy_s": 0.0, "busy_frac": 0.0}
    busy, cs, ce = 0, rows[0][0], rows[0][1]
    for s, e, _ in rows[1:]:
        if s > ce:
            busy += ce - cs
            cs, ce = s, e
        else:
            ce = max(ce, e)
    busy += ce - cs
    span = max(e for _, e, _ in rows) - rows[0][0]
    out = {"dispatches": len(rows), "span_s": span / 1e9, "busy_s": busy / 1e9,
           "busy_frac": busy / span if span else 0.0}
    if window:
        per = {}
        for s, e, k in rows:
            c = per.setdefault(k, [0, 0])
            c[0] += 1
            c[1] += e - s
        out["window"] = window
        out["kernels"] = [{"kernel": k, "calls": c, "ms": round(t / 1e6, 3)}
                          for k, (c, t) in sorted(per.items(), key=lambda kv: -kv[1][1])]
    return out


def kron_products(count):
    """Hubbard 4x4 (5,5) sector, H(U = 2) as the matrix-free two-block
    product (the quench's real-time operator): `count` timed products after
    two untimed ones (the first product of a block shape picks the kernel
    variants).  Prints one JSON line."""
    import torch

    from qudrip_amd import operators as Q
    from qudrip_amd.runtime import Basis, Context
    from qudrip_amd.workloads import make_operator
    ctx = Context(0)
    nsite = 16
    idx, h = Q.hubbard(nsite, Q.hubbard_bonds_square(4, 4), 1.0, 2.0)
    basis = Basis(ctx, Q.hubbard_sector(idx, nsite, 5, 5))
    op, keep = make_operator(ctx, basis, h, 2 * nsite, True)
    n = basis.size
    x = ctx.seeded_state(n, 3)
    y = ctx.zeros(n)
    for _ in range(2):
        op.apply(x, y)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(count):
        op.apply(x, y)
    e1.record(st)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / count
    print(json.dumps({"workload": "Hubbard 4x4 (5,5) matrix-free product", "n": n, "products": count,
                      "us_per_product": round(us, 1), "algorithmic_GBs_32n": round(32 * n / us / 1e3, 1)}),
          flush=True)
    for hnd in [op] + keep + [basis]:
        hnd.close()


def extras(ctx):
    """Rows beyond the Krylov core, measured: the 4x4 Hubbard (5,5) sector
    filtered to two doubly occupied sites (getNd), and a qubits-times-mode
    matrix (XY chain N=20 at half filling times 16 boson levels, 2,956,096
    states): its `>>` build and its product (algorithmic bytes as for any
    CSR, 20 nnz + 4 (n+1) + 32 n)."""
    from math import comb

    import torch

    from qudrip_amd import operators as Q
    from qudrip_amd.runtime import CSR, Basis, LinOp, TermProgram
    log = lambda m: print("[bench] " + m, file=sys.stderr, flush=True)
    out = {}
    L = 16
    idx = Q.getQbits(2 * L)
    sec = Q.Constrain(idx, Q.getNup(idx, L) == 5, Q.getNdo(idx, L) == 5, Q.getNd(idx, L) == 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b = Basis(ctx, sec)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    out["basis_hub4x4_55_nd2"] = {"states": b.size, "closed_form": comb(16, 2) * comb(14, 3) * comb(11, 3),
                                  "of_sector": 19079424, "build_ms": round(ms, 2)}
    log(f"Nd-filtered basis: {b.size} of 19079424 states in {ms:.1f} ms")
    b.close()
    q, idxm, ham = Q.holstein_chain(20, 16, 0.9, 0.45)
    sec = Q.Constrain(idxm, Q.getNset(q) == 10)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, 20)
    c = CSR(ctx, b, p)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    op = LinOp.from_csr(ctx, c)
    n, nnz = c.n, c.nnz
    x = ctx.seeded_state(n, 6)
    y = ctx.zeros(n)
    for _ in range(5):
        op.apply(x, y)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    reps = 50
    for _ in range(reps):
        op.apply(x, y)
    e1.record(st)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    Bm = 20 * nnz + 4 * (n + 1) + 32 * n
    out["holstein20_mode16"] = {"n": n, "nnz": nnz, "build_ms": round(build_ms, 2), "spmv_us": round(us, 2),
                                "spmv_GBs": round(Bm / (us * 1e-6) / 1e9, 1)}
    log(f"qubits x mode: n={n} nnz={nnz} build {build_ms:.1f} ms, product {us:.1f} us "
        f"= {Bm / (us * 1e-6) / 1e9:.0f} GB/s")
    for h in (op, c, p, b):
        h.close()
    return out


def trajectories(ctx, args):
    """Dissipative L=8 Hubbard chain, 4096 trajectories to T=10 (dt=0.05).
    Checks: E[N(T)] = 8 (1 - 0.05 dt)^200 within 5 sigma, and the jump
    sequences of the first trajectories equal the oracle's on the same random
    streams (timed as the CPU sample, single thread, precomputed operators)."""
    import numpy as np

    from qudrip_amd import workloads as WL
    log = lambda m: print("[bench] " + m, file=sys.stderr, flush=True)
    r = WL.hubbard_trajectories(ctx, ntraj=args.traj, want_log=True, log=log)
    ch, nsteps = r["channels"], r["nsteps"]
    is_loss = np.array([c[0] == 1 for c in ch] + [False])
    N = 8 - is_loss[r["log"]].sum(axis=1)
    expect = 8 * (1 - 0.05 * 0.05) ** nsteps
    sigma = N.std() / np.sqrt(len(N))
    out = {"L": 8, "n_start": r["n"], "ntraj": args.traj, "steps": nsteps, "dt": 0.05, "k_rt": 13,
           "run_s": round(r["run_s"], 3), "traj_steps_per_s": round(args.traj * nsteps / r["run_s"], 1),
           "jumps": r["jumps"], "losses": r["losses"], "sectors": r["sectors"],
           "krylov_form": "lanczos" if r["lanczos"] else "arnoldi",
           "lanczos_fallback_steps": r["lanczos_fallback_steps"],
           "N_T": round(float(N.mean()), 5), "N_T_closed_form": round(expect, 5), "N_T_sigma": round(float(sigma), 5),
           "closed_form_z": round(abs(N.mean() - expect) / sigma, 3),
           "density_site0_up_T": round(float(r["mean"][-1][0]), 6),
           "density_site0_up_T_sigma": round(float(np.sqrt(max(r["var"][-1][0], 0.0) / args.traj)), 6)}
    # the same run with the Krylov step fully reorthogonalised (the Arnoldi
    # form, delayed CGS2 over the whole basis): its time, and how far the
    # Lanczos form's results are from it
    if r["lanczos"]:
        old_form = os.environ.get("QD_TRAJ_KRYLOV")
        os.environ["QD_TRAJ_KRYLOV"] = "arnoldi"
        try:
            ra = WL.hubbard_trajectories(ctx, ntraj=args.traj, want_log=True)
        finally:
            if old_form is None:
                os.environ.pop("QD_TRAJ_KRYLOV", None)
            else:
                os.environ["QD_TRAJ_KRYLOV"] = old_form
        out["arnoldi_form"] = {"run_s": round(ra["run_s"], 3),
                               "traj_steps_per_s": round(args.traj * nsteps / ra["run_s"], 1),
                               "logs_equal": bool((ra["log"] == r["log"]).all()),
                               "max_abs_density_diff": float(np.abs(ra["mean"] - r["mean"]).max())}
        log(f"trajectories, Arnoldi form (full reorthogonalisation): {ra['run_s']:.3f} s, logs equal "
            f"{out['arnoldi_form']['logs_equal']}, max |d density| {out['arnoldi_form']['max_abs_density_diff']:.1e}")
    # plain: one trajectory at a time through the single-vector evolve (same
    # random streams, so the jump sequences must agree), on a sample
    ns_plain = 8
    _, plog, psecs = WL.plain_trajectories(ctx, r["ham"], 8, 4, 4, r["psi0"], ch, ns_plain, nsteps, 0.05, 13, 5)
    out["plain_gpu"] = {"traj_steps_per_s": round(ns_plain * nsteps / psecs, 1), "sample_trajectories": ns_plain,
                        "logs_equal_batched": bool((plog == r["log"][:ns_plain]).all())}
    out["batched_vs_plain"] = round(out["traj_steps_per_s"] / out["plain_gpu"]["traj_steps_per_s"], 1)
    log(f"trajectories plain: {out['plain_gpu']['traj_steps_per_s']} traj-steps/s; batched "
        f"{out['batched_vs_plain']}x")
    if args.traj_cpu > 0 and not args.no_cpu:
        import math

        import oracle
        from qudrip_amd import operators as Q
        psi0 = r["psi0"].cpu().numpy()
        t0 = time.perf_counter()
        _, olog = oracle.traj_run(r["ham"], 8, 4, 4, psi0, ch, args.traj_cpu, nsteps, 0.05, 13, 5)
        cpu_s = time.perf_counter() - t0
        out["oracle_prefix_logs_equal"] = bool((olog == r["log"][:args.traj_cpu]).all())
        # the loop as the reference runs it: jump operators built by `>>`,
        # multiplied and evaluated every step, H_eff assembled as a stored
        # matrix before each Arnoldi step (hubbard_diss.cpp:443,462-463,514-529)
        g = Q.getHopFermiGate(Q.getQbits(16))
        lsums = [math.sqrt(rate) * g(m, m) for (_, m, rate) in ch]
        nf = max(1, args.traj_cpu // 4)
        t0 = time.perf_counter()
        _, flog = oracle.traj_run_rebuild(r["ham"], lsums, 8, 4, 4, psi0, ch, nf, nsteps, 0.05, 13, 5)
        f_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(nf * nsteps / f_s, 2), "unit": "trajectory-steps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{nf} trajectories x {nsteps} steps in {f_s:.1f} s, oracle restatement "
                                         "of the reference loop (jump operators and H_eff rebuilt every step), "
                                         "1 thread",
                               "logs_equal_batched": bool((flog == r["log"][:nf]).all())}
        out["cpu_precomputed"] = {"value": round(args.traj_cpu * nsteps / cpu_s, 2), "unit": "trajectory-steps/s",
                                  "cores": 1, "kind": "port",
                                  "sample": f"{args.traj_cpu} trajectories x {nsteps} steps in {cpu_s:.1f} s, oracle "
                                            "restatement with operators precomputed per sector, 1 thread"}
        out["batched_vs_cpu_reference_loop"] = round(out["traj_steps_per_s"] / out["cpu_baseline"]["value"], 1)
        out["batched_vs_cpu_precomputed"] = round(out["traj_steps_per_s"] / out["cpu_precomputed"]["value"], 1)
    log(f"trajectories: N(T) = {out['N_T']} vs {out['N_T_closed_form']} (z={out['closed_form_z']})")
    return out


def serving(ctx, args):
    """Poisson stream (200 req/s, seed 4): 70 % XXZ ground states (N in
    16..26, Jz ~ U[0,2]), 30 % 3x4 Hubbard real-time solves (U ~ U[2,8], 100
    steps of 0.05).  Served: the solve server (structure cache, concurrent
    workers, ground states stopped when the Ritz-residual bound r^2/gap or the
    move of E0 over 4 Krylov vectors is <= 1e-12 |E0|, matrix-free real-time
    operator with the Krylov dimension set per step by Saad's estimate).  The
    Krylov bases (ground states and real-time steps) use partial
    reorthogonalisation by default.  Rows, to
    separate what serving buys from what the per-request algorithm buys:
      served             -- the stream, args.serve_workers workers;
      served_1_worker    -- the stream, one worker;
      one_at_a_time_same_rules -- the server's algorithms and cached
                            structure, requests one at a time (closed loop);
      served_full_reorth -- the stream, full reorthogonalisation (the drop-in
                            path's scheme);
      plain              -- each request alone through the drop-in path
                            (basis + stored matrix per request, Lanczos
                            k = 150 with full reorthogonalisation, evolve with
                            kry_dim_rt = 30).
    Latency = arrival -> result (queueing included); service = start on a
    worker -> result.  Accuracy: served vs plain energies per request."""
    import numpy as np

    from qudrip_amd import serving as SV
    log = lambda m: print("[bench] " + m, file=sys.stderr, flush=True)
    reqs = SV.request_stream(args.serve_requests, rate=200.0, seed=4)
    served = SV.run_served(reqs, workers=args.serve_workers, log=log)
    log(f"served: {len(reqs)} requests in {served['span_s']:.2f} s ({served['solves_per_s']:.2f} /s)")
    one = SV.run_served(reqs, workers=1)
    log(f"served, one worker: {one['span_s']:.2f} s ({one['solves_per_s']:.2f} /s)")
    seq = SV.run_served(reqs, workers=1, closed_loop=True)
    log(f"one at a time, same rules: {seq['span_s']:.2f} s ({seq['solves_per_s']:.2f} /s)")
    full = SV.run_served(reqs, workers=args.serve_workers, gs_method=1)
    log(f"served, full reorthogonalisation: {full['span_s']:.2f} s ({full['solves_per_s']:.2f} /s)")
    plain = SV.run_plain(reqs, ctx)
    log(f"plain: {plain['busy_s']:.2f} s busy ({plain['solves_per_s']:.2f} /s)")
    rel_gs, rel_gs_full, d_rt, drift = [], [], [], []
    for r in reqs:
        s, p = served["results"][r["tag"]], plain["results"][r["tag"]]
        if s["status"] != 0:
            continue
        if r["kind"] == "gs":
            rel_gs.append(abs(s["energy"] - p["energy"]) / abs(p["energy"]))
            f = full["results"][r["tag"]]
            if f["status"] == 0:
                rel_gs_full.append(abs(f["energy"] - p["energy"]) / abs(p["energy"]))
        else:
            d_rt.append(abs(s["observable"] - p["observable"]))
            drift.append(abs(s["energy"] - s["energy_start"]) / abs(s["energy_start"]))
    its = [served["results"][r["tag"]]["iterations"] for r in reqs if r["kind"] == "gs"]
    rt_m = [served["results"][r["tag"]]["krylov_mean"] for r in reqs if r["kind"] == "rt"]
    busy = {}
    for r in reqs:
        x = served["results"][r["tag"]]
        key = f"gs_N{r['N']}" if r["kind"] == "gs" else "rt_3x4"
        b = busy.setdefault(key, [0, 0.0, 0.0])
        b[0] += 1
        b[1] += x["t_end"] - x["t_start"]
        b[2] += plain["results"][r["tag"]]["time_s"]

    def row(out, workers):
        l50, l99 = SV.latency_percentiles(out["results"])
        s50, s99 = SV.service_percentiles(out["results"])
        return {"solves_per_s": round(out["solves_per_s"], 3), "span_s": round(out["span_s"], 3),
                "p50_latency_s": round(l50, 4), "p99_latency_s": round(l99, 4),
                "p50_service_s": round(s50, 4), "p99_service_s": round(s99, 4), "workers": workers,
                "failed": sum(1 for x in out["results"].values() if x["status"] != 0)}

    srow = row(served, args.serve_workers)
    srow.update({"max_running": served["stats"]["max_running"],
                 "peak_hbm_gb": round(served["stats"]["peak_bytes"] / 1e9, 2),
                 "model_setup_s": round(served["setup_s"], 3),
                 "gs_krylov_dim_mean": round(float(np.mean(its)), 1) if its else None,
                 "rt_krylov_dim_mean": round(float(np.mean(rt_m)), 2) if rt_m else None,
                 "per_kind_count_served_busy_s_plain_s": {k: [v[0], round(v[1], 3), round(v[2], 3)]
                                                         for k, v in sorted(busy.items())}})
    frow = row(full, args.serve_workers)
    frow.update({"speedup_vs_plain": round(full["solves_per_s"] / plain["solves_per_s"], 3),
                 "reaches_5x_vs_plain": bool(full["solves_per_s"] >= 5.0 * plain["solves_per_s"]),
                 "max_rel_dE_gs": max(rel_gs_full) if rel_gs_full else None})
    p50, p99 = SV.latency_percentiles(plain["results"])
    q50, q99 = SV.service_percentiles(plain["results"])
    return {"requests": len(reqs), "rate_per_s": 200.0, "gs_fraction": 0.7,
            "served": srow,
            "served_1_worker": row(one, 1),
            "one_at_a_time_same_rules": row(seq, 1),
            "served_full_reorth": frow,
            "plain": {"solves_per_s": round(plain["solves_per_s"], 3), "busy_s": round(plain["busy_s"], 3),
                      "p50_latency_s": round(p50, 4), "p99_latency_s": round(p99, 4),
                      "p50_service_s": round(q50, 4), "p99_service_s": round(q99, 4)},
            "speedup": round(served["solves_per_s"] / plain["solves_per_s"], 3),
            "gain_from_concurrency": round(served["solves_per_s"] / one["solves_per_s"], 3),
            "gain_from_server_rules_and_cache": round(seq["solves_per_s"] / plain["solves_per_s"], 3),
            "max_rel_dE_gs": max(rel_gs) if rel_gs else None,
            "max_abs_d_docc_rt": max(d_rt) if d_rt else None,
            "max_rel_energy_drift_rt": max(drift) if drift else None}


def cpu_baseline(csr, n, nnz, B, x, budget_s):
    """Oracle (single-thread C port of the reference's ColMajor scatter SpMV)
    on this host, over the same CSR arrays (bit-identical to the oracle's own
    build), repeated until ~budget_s seconds."""
    import oracle
    outer, inner, vals = csr.download()
    xv = x.cpu().numpy()
    reps, t_tot = 0, 0.0
    while t_tot < budget_s and reps < 200:
        t0 = time.perf_counter()
        oracle.spmv((outer, inner, vals), xv)
        t_tot += time.perf_counter() - t0
        reps += 1
    gbs = B * reps / t_tot / 1e9
    out = {"value": round(gbs, 3), "unit": "GB/s", "cores": 1, "kind": "port",
           "sample": f"{reps} SpMVs of the same XXZ N=24 CSR (n={n}, nnz={nnz}) in {t_tot:.1f} s, "
                     f"oracle ColMajor scatter (sparse_algorithm.hpp:378), 1 thread",
           "spmv_s": round(t_tot / reps, 4)}
    out.update(host_cpu())
    return out


def cpu_more(csr, n, args):
    """The other CPU rows, single thread:
    - XXZ N=24 ground state: the oracle's getLanczosSolver (two-pass 3-term
      recurrence, sparse_algorithm.hpp:395-473,564-585) at k = 150 on the same
      CSR (about 15 s on a current server core);
    - Hubbard matrix-free product (OperatorSum::map_acc, operator.hpp:215-232)
      on the 3x4 (6,6) sector (n = 853,776), projected to the 4x4 (5,5) sector
      by states x terms x log2(states) (the rank lookup is a binary search);
    - with --cpu-full, the 4x4 product itself (basis scan and one product,
      a few minutes)."""
    import math

    import numpy as np

    import oracle
    from qudrip_amd import operators as Q
    from qudrip_amd import serving as SV
    log = lambda m: print("[bench] " + m, file=sys.stderr, flush=True)
    out = {}
    outer, inner, vals = csr.download()
    csc = (outer, inner, vals)
    psi = oracle.seeded_state(n, 2)
    k = 150
    t0 = time.perf_counter()
    r = oracle.lanczos_gs(csc, k, psi)
    gs_s = time.perf_counter() - t0
    out["xxz24_ground_state"] = {
        "k": k, "solve_s": round(gs_s, 2), "E0": r["E0"],
        "rel_err": abs(r["E0"] - E_BETHE_XXX24) / abs(E_BETHE_XXX24) if n == 2704156 else None,
        "method": "oracle getLanczosSolver (two passes of the 3-term recurrence), 1 thread"}
    log(f"cpu: XXZ24 Lanczos k={k} in {gs_s:.1f} s")
    nsite, bonds, nup, ndn = SV.hubbard_lattice()
    idx, H = Q.hubbard(nsite, bonds, 1.0, 4.0)
    b = oracle.basis_scan(Q.hubbard_sector(idx, nsite, nup, ndn))
    xs = oracle.seeded_state(len(b), 3)
    ys = np.zeros(len(b), dtype=np.complex128)
    t0 = time.perf_counter()
    oracle.map_acc(b, H, xs, ys)
    s34 = time.perf_counter() - t0
    n44, nt34, nt44 = 19079424, len(H.terms), len(Q.hubbard(16, Q.hubbard_bonds_square(4, 4), 1.0, 4.0)[1].terms)
    proj = s34 * (n44 / len(b)) * (nt44 / nt34) * (math.log2(n44) / math.log2(len(b)))
    out["hubbard_matfree_product"] = {"hub3x4_n": len(b), "hub3x4_s": round(s34, 2), "hub4x4_n": n44,
                                      "hub4x4_s": round(proj, 1), "hub4x4_is": "projected",
                                      "method": "oracle map_acc (operator.hpp:215-232), 1 thread"}
    log(f"cpu: Hub3x4 map_acc in {s34:.1f} s (4x4 projected {proj:.0f} s)")
    if args.cpu_full:
        idx4, H4 = Q.hubbard(16, Q.hubbard_bonds_square(4, 4), 1.0, 2.0)
        t0 = time.perf_counter()
        b4 = oracle.basis_scan(Q.hubbard_sector(idx4, 16, 5, 5))
        scan_s = time.perf_counter() - t0
        x4 = oracle.seeded_state(len(b4), 3)
        y4 = np.zeros(len(b4), dtype=np.complex128)
        t0 = time.perf_counter()
        oracle.map_acc(b4, H4, x4, y4)
        s44 = time.perf_counter() - t0
        out["hubbard_matfree_product"].update({"hub4x4_s": round(s44, 1), "hub4x4_is": "measured",
                                               "hub4x4_basis_scan_s": round(scan_s, 1)})
        log(f"cpu: Hub4x4 basis scan {scan_s:.1f} s, map_acc {s44:.1f} s")
    # one real-time step of the quench (k_rt = 30) makes 29 products; the
    # projection counts only those (the MGS vector work comes on top)
    p44 = out["hubbard_matfree_product"]["hub4x4_s"]
    out["hub4x4_quench_step_s_projected"] = round(29 * p44, 1)
    out["hub4x4_quench_200_steps_h_projected"] = round(200 * 29 * p44 / 3600, 1)
    out.update(host_cpu())
    return out


if __name__ == "__main__":
    main()
