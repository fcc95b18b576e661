#!/usr/bin/env python3
"""Benchmark of the recursive Gram-Schmidt QR (later_rgsqrf) on one MI355X.

Workload (the 16384^2 square QR the headline metric is quoted on): A is
m = n = 16384 FP32 column-major, entries N(0,1) from the Philox generator
(seed 0), resident in HBM before timing.  One step is one complete QR
factorisation A = QR (A overwritten by Q) in the mixed-precision form: FP16
hi/lo-split MFMA products, FP64-summed panels, second Gram-Schmidt pass.  The
plain-FP32 form (FP32 MFMA, one pass, the reference's algorithm in FP32) is
timed on the same input beside it.

value = 2 n^2 (m - n/3) / t, the reference's own GFLOPS formula
(test/test_qr.cu:85-86) in TFLOP/s, for the mixed form.  Accuracy of each form
is reported with FP64 accumulation: ||I - Q^T Q||_F and ||A - QR||_F/||A||_F.

roofline: the dominant kernel (node GEMMs) timed live with HIP events on the
launch stream; achieved = algorithmic flops 2 m n1 n2 per projection / update
launch summed over one step, divided by the summed kernel time.
cpu_baseline: host LAPACK sgeqrf + sorgqr (scipy's OpenBLAS, threads = the
process's cores) on a bounded square sample and on the 8192 x 1024 workload,
beside the oracle's C restatement of the FP32 recursion (OpenMP), TFLOP/s by
the same formula.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 16384]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MFMA_PEAK_TFLOPS = {"f16x3": 2500.0, "f16": 2500.0, "bf16": 2500.0, "f32": 157.3, "i8": 5000.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_form(later, torch, A0, form, steps, warmup, graph=True):
    """Times `steps` factorisations, each of a fresh copy of A0 (the copy is
    outside the timed events).  With graph=True the factorisation is one
    replay of a graph-captured plan; the first eager call is timed too."""
    m, n = A0.shape
    o = later.opts(form) if isinstance(form, str) else form
    work = later.workspace(m, n, o)
    hwork = later.hworkspace(m, n, o)  # 16-bit operand staging (None for the FP32 form)
    A = later.colmajor(m, n)
    R = later.colmajor(n, n)
    s = torch.cuda.current_stream()

    def one(run):
        A.copy_(A0)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record(s)
        run()
        ev1.record(s)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1)

    eager_ms = one(lambda: later.rgsqrf(A, R=R, o=o, work=work, hwork=hwork))
    run = lambda: later.rgsqrf(A, R=R, o=o, work=work, hwork=hwork)
    plan = None
    if graph:
        plan = later.Plan(A, R, o, work, hwork)
        run = plan.execute
    times = []
    for it in range(warmup + steps):
        t = one(run)
        if it >= warmup:
            times.append(t)
    if plan is not None:
        plan.close()
    return times, eager_ms, A, R, (work, hwork), o


def profile_step(later, torch, A0, form, work, o):
    """One more step with kernel-class timing; returns the per-class totals."""
    m, n = A0.shape
    A = later.colmajor(m, n)
    R = later.colmajor(n, n)
    A.copy_(A0)
    torch.cuda.synchronize()
    later.prof_begin()
    later.rgsqrf(A, R=R, o=o, work=work[0], hwork=work[1])
    return later.prof_end()


def vendor_sgeqrf(torch, A0, reps=2):
    """rocSOLVER sgeqrf on the same matrix, timed like the reference times
    cuSOLVER SGEQRF beside rgsqrf (test/test_qr.cu:184-206): R and the
    Householder vectors only, no explicit Q.  Returns ms or None."""
    import ctypes
    try:
        rb = ctypes.CDLL("/opt/rocm/lib/librocblas.so")
        rs = ctypes.CDLL("/opt/rocm/lib/librocsolver.so")
    except OSError as e:
        log(f"[bench] rocSOLVER unavailable: {e}")
        return None
    m, n = A0.shape
    h = ctypes.c_void_p()
    if rb.rocblas_create_handle(ctypes.byref(h)) != 0:
        return None
    s = torch.cuda.current_stream()
    rb.rocblas_set_stream(h, ctypes.c_void_p(s.cuda_stream))
    A = A0.clone().t().contiguous().t()
    tau = torch.empty(min(m, n), device="cuda", dtype=torch.float32)
    rs.rocsolver_sgeqrf.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_void_p]
    best = None
    for _ in range(reps):
        A.copy_(A0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(s)
        rc = rs.rocsolver_sgeqrf(h, m, n, A.data_ptr(), A.stride(1), tau.data_ptr())
        e1.record(s)
        torch.cuda.synchronize()
        if rc != 0:
            log(f"[bench] rocsolver_sgeqrf status {rc}")
            return None
        t = e0.elapsed_time(e1)
        best = t if best is None else min(best, t)
    rb.rocblas_destroy_handle(h)
    return best


def hbm_rate(m: int, n: int, ms: float) -> float:
    """Algorithmic GB/s of a tall-skinny factorisation: read A, write Q (8 m n
    bytes) and write R (4 n^2)."""
    return (8.0 * m * n + 4.0 * n * n) / (ms * 1e-3) / 1e9


def single_config(later, torch, name, m, n, kind, forms, steps=3):
    """One factorisation configuration, timed by graph replay, per form."""
    A0 = (later.generate_normal if kind == "normal" else later.generate_uniform)(m, n, seed=0, mat=0)
    out = {"m": m, "n": n, "input": kind + " seed 0"}
    for form in forms:
        times, eager_ms, Q, R, _, _ = time_form(later, torch, A0, form, steps, 1)
        ms = sum(times) / len(times)
        out[form] = {"ms": ms, "tflops_ref": later.flops_ref(m, n) / ms / 1e9,
                     "gbps": hbm_rate(m, n, ms), "orth": later.check_orth(Q),
                     "backward": later.check_backward(A0, Q, R)}
        log(f"[bench] {name} {m}x{n} {form}: {ms:.3f} ms  {out[form]['tflops_ref']:.2f} TFLOP/s(ref)  "
            f"{out[form]['gbps']:.0f} GB/s  orth={out[form]['orth']:.2e} bwd={out[form]['backward']:.2e}")
        del Q, R
    del A0
    torch.cuda.empty_cache()
    return out


def blas_ceiling(torch, M=16384, N=8192, K=8192, reps=5):
    """The vendor library's dense GEMM rate on this GPU (torch.matmul, i.e.
    hipBLASLt) at the root update's shape, FP16 and FP32: the practical
    ceilings the node GEMMs are compared with."""
    out = {}
    for name, dt in (("fp16", torch.float16), ("fp32", torch.float32)):
        a = torch.randn(M, K, device="cuda", dtype=dt)
        b = torch.randn(K, N, device="cuda", dtype=dt)
        for _ in range(2):
            c = a @ b
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            c = a @ b
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"ms": ms, "tflops": 2.0 * M * N * K / ms / 1e9, "shape": [M, N, K]}
        del a, b, c
    torch.cuda.empty_cache()
    log(f"[bench] vendor GEMM ceiling {M}x{N}x{K}: fp16 {out['fp16']['tflops']:.0f} TFLOP/s, "
        f"fp32 {out['fp32']['tflops']:.0f} TFLOP/s")
    return out


def square_variants(later, torch, A0, steps=2):
    """The 16384^2 square workload's other forms: FP16 and BF16 node GEMMs
    with and without the second pass (Cholesky QR leaves when repeated), FP32
    products with the second pass, on
    the same N(0,1) input; and the graded input A = U diag(s) V^T (U, V the
    Q factors of two Gaussian matrices, s geometric from 1 to 1e-4) in the
    mixed and plain-FP32 forms."""
    m, n = A0.shape
    out = {}
    variants = {
        "fp16": later.opts("reference"),
        "fp16_reorth": later.opts("reference", reorth=later.REORTH_PASS, panel=later.PANEL_CHOLQR1, chunk=256,
                                  chunk2=32),
        "bf16": later.opts("reference", gemm=later.GEMM_BF16, gemm_pass2=later.GEMM_BF16),
        "bf16_reorth": later.opts("reference", gemm=later.GEMM_BF16, gemm_pass2=later.GEMM_BF16,
                                  reorth=later.REORTH_PASS, panel=later.PANEL_CHOLQR1),
        # The plain-FP32 form with the mixed form's second pass and leaves:
        # FP32 products everywhere (does FP32 reach the gate at all?).
        "fp32_reorth": later.opts("fp32", reorth=later.REORTH_PASS, panel=later.PANEL_CHOLQR1),
    }

    def run(name, A, o):
        times, eager_ms, Q, R, _, _ = time_form(later, torch, A, o, steps, 1)
        ms = sum(times) / len(times)
        e = {"ms": ms, "tflops_ref": later.flops_ref(m, n) / ms / 1e9, "orth": later.check_orth(Q),
             "backward": later.check_backward(A, Q, R)}
        e["gates_met"] = e["orth"] <= 1e-5 and e["backward"] <= 1e-6
        log(f"[bench] C2 {name}: {ms:.2f} ms  {e['tflops_ref']:.1f} TFLOP/s(ref)  orth={e['orth']:.2e} "
            f"bwd={e['backward']:.2e}")
        del Q, R
        torch.cuda.empty_cache()
        return e

    for name, o in variants.items():
        out[name] = run(name, A0, o)
    # Graded input.
    U = later.generate_normal(m, n, 1, 0)
    later.rgsqrf(U, o=later.opts("mixed"))
    V = later.generate_normal(n, n, 2, 0)
    later.rgsqrf(V, o=later.opts("mixed"))
    s = torch.pow(torch.tensor(1e-4, dtype=torch.float64, device="cuda"),
                  torch.arange(n, device="cuda", dtype=torch.float64) / (n - 1)).float()
    G = (U * s[None, :]) @ V.t()
    Ag = later.colmajor(m, n)
    Ag.copy_(G)
    del U, V, G
    torch.cuda.empty_cache()
    for form in ("mixed", "fp32"):
        out["graded_" + form] = run("graded (cond 1e4) " + form, Ag, later.opts(form))
    del Ag
    torch.cuda.empty_cache()
    return out


def batched_config(later, torch, count=4096, steps=3, single=False):
    """4096 independent matrices, m log-uniform in [64, 4096], n uniform in
    [16, min(m, 256)], entries 2u - 1 (Philox, seed 0, matrix index i).  The
    headline engine is the persistent kernel (one launch; a workgroup per
    matrix from a device work queue; per-bucket autotuned variants), in the
    FP32-product form with reorthogonalisation ("mixed" options) and the
    BF16 one; the grouped-launch graph of the shared recursion tree is timed
    beside it."""
    import numpy as np
    rng = np.random.default_rng(0)
    m = np.rint(np.exp(rng.uniform(np.log(64), np.log(4096), count))).astype(int)
    n = np.array([rng.integers(16, min(mi, 256) + 1) for mi in m])
    As = [later.generate_uniform(int(mi), int(ni), seed=0, mat=i) for i, (mi, ni) in enumerate(zip(m, n))]
    A0 = [a.clone() for a in As]
    Rs = [later.colmajor(int(ni), int(ni)) for ni in n]
    flops = float(np.sum(2.0 * m * n * n))
    nbytes = float(np.sum(8.0 * m * n))
    s = torch.cuda.current_stream()
    if single:  # one persistent launch of the FP32-product form, for profilers
        plan = later.BatchWg(As, Rs, later.opts("mixed"))
        plan.execute()
        torch.cuda.synchronize()
        plan.close()
        log(f"[bench] one persistent batched launch over {count} matrices done")
        return None

    def timed(plan):
        times = []
        for it in range(steps + 1):
            for a, a0 in zip(As, A0):
                a.copy_(a0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record(s)
            plan.execute()
            e1.record(s)
            torch.cuda.synchronize()
            if it:
                times.append(e0.elapsed_time(e1))
        worst = 0.0
        for i in range(0, count, 256):  # spot checks
            worst = max(worst, later.check_orth(As[i]))
        return sum(times) / len(times), worst

    out = {"count": count, "shapes": "m log-uniform [64, 4096], n uniform [16, min(m, 256)], seed 0",
           "flops_2mn2": flops, "bytes_8mn": nbytes}
    forms = {"persistent_fp32_reorth": later.opts("mixed"),
             "persistent_bf16_reorth": later.opts("reference", gemm=later.GEMM_BF16, reorth=later.REORTH_PASS)}
    for name, o in forms.items():
        t0 = time.time()
        plan = later.BatchWg(As, Rs, o)
        build_s = time.time() - t0
        ms, worst = timed(plan)
        bad = int(plan.status.abs().sum())
        out[name] = {"ms": ms, "tflops_mgs": flops / ms / 1e9, "gbps": nbytes / ms / 1e6, "launches": 1,
                     "plan_build_s": build_s, "orth_max_sampled": worst, "breakdowns": bad}
        log(f"[bench] C4 batched {count} matrices ({name}): {ms:.2f} ms  {flops / ms / 1e9:.2f} TFLOP/s(2mn^2)  "
            f"{nbytes / ms / 1e6:.0f} GB/s  1 launch  orth<={worst:.1e}")
        plan.close()
    t0 = time.time()
    plan = later.BatchPlan(As, Rs, later.opts("mixed"))
    build_s = time.time() - t0
    ms, worst = timed(plan)
    out["grouped_mixed"] = {"ms": ms, "tflops_mgs": flops / ms / 1e9, "gbps": nbytes / ms / 1e6,
                            "launches": plan.launches, "plan_build_s": build_s, "orth_max_sampled": worst}
    log(f"[bench] C4 batched {count} matrices (grouped launches, mixed): {ms:.2f} ms  "
        f"{flops / ms / 1e9:.2f} TFLOP/s(2mn^2)  launches={plan.launches}  orth<={worst:.1e}")
    plan.close()
    del As, A0, Rs
    torch.cuda.empty_cache()
    return out


def out_of_core_config(later, torch, m=65536, n=8192, budget=1 << 30, forms=("reference", "mixed")):
    """later_oc_qr_rec on a 65536 x 8192 N(0,1) matrix (2 GiB) in pinned host
    memory with a 1 GiB device budget: wall time of the synchronous call,
    host-device copies included (the rate is PCIe-inclusive), and the FP64
    metrics of the result."""
    out = {"m": m, "n": n, "device_budget_bytes": budget, "input": "normal seed 0",
           "timing": "wall clock of the synchronous call, host<->device copies included"}
    A0 = later.generate_normal(m, n, seed=0, mat=0)
    host = torch.empty((n, m), dtype=torch.float32, pin_memory=True).t()
    Rh = torch.empty((n, n), dtype=torch.float32, pin_memory=True).t()
    for form in forms:
        o = later.opts(form)
        host.copy_(A0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        later.oc_qr_rec(host, Rh, budget, o)
        ms = (time.perf_counter() - t0) * 1e3
        Q = later.colmajor(m, n)
        Q.copy_(host)
        R = later.colmajor(n, n)
        R.copy_(Rh)
        torch.cuda.synchronize()
        e = {"ms": ms, "tflops_ref": later.flops_ref(m, n) / ms / 1e9, "orth": later.check_orth(Q),
             "backward": later.check_backward(A0, Q, R)}
        out[form] = e
        log(f"[bench] out-of-core {m}x{n} ({budget >> 20} MiB device) {form}: {ms:.1f} ms  "
            f"{e['tflops_ref']:.1f} TFLOP/s(ref, PCIe-inclusive)  orth={e['orth']:.2e} bwd={e['backward']:.2e}")
        del Q, R
        torch.cuda.empty_cache()
    return out


def serving_config(later, torch, count=64, rates=(100.0, 1e9), slots=8, budget=16 << 30):
    """Least-squares serving: requests with m log-uniform in [128, 32768],
    n log-uniform in [16, 1024] (n <= m), condition number log-uniform in
    [1, 1e6] (A = U diag(s) V^T, s_i = cond^(-i/(n-1)), U and V from FP64 QR
    of Gaussians), k uniform in [1, 16] right-hand sides, b = A x + 1e-3 noise,
    arriving as a Poisson process of each of `rates` requests/s (1e9: all at
    once, the saturation throughput).  Latency is the server's own
    submit-to-done time; every x is checked against FP64 least squares and
    against the first-order LS perturbation bound for a backward error of
    1e-6, kappa eps + kappa^2 eps ||r|| / (||A|| ||x||)."""
    import numpy as np
    rng = np.random.default_rng(1)
    g = torch.Generator(device="cuda").manual_seed(1)
    probs = []
    for i in range(count):
        m = int(round(np.exp(rng.uniform(np.log(128), np.log(32768)))))
        n = min(m, int(round(np.exp(rng.uniform(np.log(16), np.log(1024))))))
        k = int(rng.integers(1, 17))
        cond = float(10 ** rng.uniform(0, 6))
        U = torch.linalg.qr(torch.randn(m, n, device="cuda", dtype=torch.float64, generator=g))[0]
        V = torch.linalg.qr(torch.randn(n, n, device="cuda", dtype=torch.float64, generator=g))[0]
        sv = cond ** (-torch.arange(n, device="cuda", dtype=torch.float64) / max(n - 1, 1))
        A64 = (U * sv) @ V.T
        x = torch.randn(n, k, device="cuda", dtype=torch.float64, generator=g)
        b64 = A64 @ x + 1e-3 * torch.randn(m, k, device="cuda", dtype=torch.float64, generator=g)
        A = A64.float().t().contiguous().t()
        B = b64.float().t().contiguous().t()
        Ad, Bd = A.double(), B.double()
        x64 = torch.linalg.lstsq(Ad, Bd).solution
        svals = torch.linalg.svdvals(Ad)
        kap = float(svals[0] / svals[-1])
        eps = 1e-6
        rn = float(torch.linalg.norm(Bd - Ad @ x64))
        bound = kap * eps + kap * kap * eps * rn / (float(svals[0]) * float(torch.linalg.norm(x64)))
        probs.append((A, B, x64, m, n, k, cond, bound))
        del U, V, A64, b64, Ad, Bd
    torch.cuda.synchronize()
    runs = {}
    for rate in rates:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
        srv = later.Server(slots=slots, hbm_budget_bytes=budget)
        ids, lat, errs, ratio, done = [None] * count, [], [], [], 0
        t0 = time.perf_counter()
        nxt = 0
        pending = set()
        while done < count:
            now = time.perf_counter() - t0
            while nxt < count and arrivals[nxt] <= now:
                A, B = probs[nxt][0], probs[nxt][1]
                ids[nxt] = srv.submit(A, B)
                pending.add(nxt)
                nxt += 1
            srv.poll()
            for i in sorted(pending):
                r = srv.result(ids[i])
                if r["state"] in (3, 4):
                    A, B, x64, m, n, k, cond, bound = probs[i]
                    if r["state"] == 3:
                        X, _ = srv.fetch(ids[i], m, n, k)
                        e = float(torch.linalg.norm(X.double() - x64) / torch.linalg.norm(x64))
                        errs.append(e)
                        ratio.append(e / (bound + 1e-7))
                        lat.append(r["latency_s"])
                    srv.release(ids[i])
                    pending.discard(i)
                    done += 1
            if time.perf_counter() - t0 > 120:
                break
        wall = time.perf_counter() - t0
        stats = srv.stats()
        srv.close()
        lat_ms = np.array(lat) * 1e3
        key = "saturation" if rate >= 1e8 else f"rate_{rate:g}"
        runs[key] = {"requests": count, "completed": len(lat), "rate_per_s": None if rate >= 1e8 else rate,
                     "slots": slots,
                     "p50_ms": float(np.percentile(lat_ms, 50)) if len(lat) else None,
                     "p99_ms": float(np.percentile(lat_ms, 99)) if len(lat) else None,
                     "solves_per_s": len(lat) / wall, "wall_s": wall,
                     "x_rel_err_median": float(np.median(errs)) if errs else None,
                     "x_rel_err_max": float(np.max(errs)) if errs else None,
                     "err_over_ls_bound_max": float(np.max(ratio)) if ratio else None,
                     "within_10x_ls_bound": int(np.sum(np.array(ratio) < 10.0)),
                     "peak_pool_bytes": stats["peak_bytes"]}
        o = runs[key]
        log(f"[bench] C5 serving {count} requests, {'all at once' if rate >= 1e8 else f'{rate:g}/s'}: "
            f"p50 {o['p50_ms']:.1f} ms  p99 {o['p99_ms']:.1f} ms  {o['solves_per_s']:.1f} solves/s  "
            f"|x-x64|/|x64| median {o['x_rel_err_median']:.1e} max {o['x_rel_err_max']:.1e}  "
            f"{o['within_10x_ls_bound']}/{len(lat)} within 10x the LS bound")
    out = dict(runs["rate_100"]) if "rate_100" in runs else {}
    out["runs"] = runs
    out["workload"] = ("m log-uniform [128, 32768], n log-uniform [16, 1024], cond log-uniform [1, 1e6], "
                       "k in [1, 16], Poisson arrivals")
    del probs
    torch.cuda.empty_cache()
    return out


def read_traffic(k):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, calibrated on a device
    copy of known size); None when the summary is absent."""
    import re
    path = os.path.join(ROOT, "profiles", "gemm_update_x3_16384_traffic.txt")
    try:
        txt = open(path).read()
    except OSError:
        return None
    mt = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt, re.M)
    return float(mt.group(1)) if mt else None


def _lapack_qr_seconds(A):
    """Host LAPACK sgeqrf + sorgqr (explicit Q, as rgsqrf returns) on a copy of A."""
    from scipy.linalg import lapack
    import numpy as np
    X = np.array(A, dtype=np.float32, order="F", copy=True)
    t = time.perf_counter()
    qr, tau, _, info = lapack.sgeqrf(X, overwrite_a=1)
    q, _, info2 = lapack.sorgqr(qr, tau, overwrite_a=1)
    dt = time.perf_counter() - t
    if info or info2:
        raise RuntimeError(f"LAPACK status {info} {info2}")
    return dt, q


def cpu_baseline(n_sample: int, n_port: int):
    """Host-core baselines (reported only): LAPACK sgeqrf + sorgqr through
    scipy's OpenBLAS with threads = the process's cores, on (1) the 8192 x 1024
    uniform workload at full size (one warm-up, median of 5) and (2) a square
    N(0,1) sample of the headline workload (one run); and the oracle's C
    restatement of the FP32 recursion (OpenMP) on a smaller square sample."""
    from oracle import oracle as O
    import scipy.linalg.lapack  # noqa: F401  (loads its BLAS before the thread limit is set)
    from threadpoolctl import threadpool_info, threadpool_limits
    cores = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        cores = min(cores, int(omp))
    ref = lambda m, n: 2.0 * n * n * (m - n / 3.0)
    with threadpool_limits(limits=cores, user_api="blas"):
        blas = [f"{i.get('internal_api')} {i.get('version')} threads={i.get('num_threads')}"
                for i in threadpool_info() if i.get("user_api") == "blas"]
        A1 = O.gen_uniform(8192, 1024, 0, 0)
        _lapack_qr_seconds(A1)
        c1 = sorted(_lapack_qr_seconds(A1)[0] for _ in range(5))[2]
        As = O.gen_normal(n_sample, n_sample, 0, 0)
        dts, Q = _lapack_qr_seconds(As)
        orth = O.orth(Q)
    Ap = O.gen_normal(n_port, n_port, 0, 0)
    t = time.perf_counter()
    O.rgsqrf(Ap, "f32")
    dtp = time.perf_counter() - t
    return {"value": ref(n_sample, n_sample) / dts / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "what": "host LAPACK sgeqrf + sorgqr (scipy's OpenBLAS), explicit Q; rate by 2 n^2 (m - n/3)",
            "sample": f"square m=n={n_sample} N(0,1) seed 0, one run after a warm-up: {dts:.2f} s, "
                      f"||I-Q^TQ||_F={orth:.2e}; BLAS: {'; '.join(blas)}",
            "c1_8192x1024": {"seconds_median_of_5": c1, "tflops_ref": ref(8192, 1024) / c1 / 1e12},
            "restatement": {"tflops_ref": ref(n_port, n_port) / dtp / 1e12, "seconds": dtp,
                            "sample": f"oracle C restatement of the FP32 recursion, m=n={n_port} N(0,1), "
                                      f"OpenMP threads={omp or os.cpu_count()}"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--m", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=6144, help="square sample of the LAPACK baseline")
    ap.add_argument("--cpu-port-sample", type=int, default=4096, help="square sample of the restatement")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="time eager launches instead of graph replays")
    ap.add_argument("--no-vendor", action="store_true", help="skip the rocSOLVER sgeqrf baseline")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the secondary workloads (8192x1024, tall-skinny, batched, serving)")
    ap.add_argument("--single", choices=("mixed", "fp32", "reference"),
                    help="one eager factorisation of the headline workload and nothing else (for profilers)")
    ap.add_argument("--single-batched", action="store_true",
                    help="one persistent launch of the 4096-matrix batched workload and nothing else (for profilers)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="three 1 GiB device-to-device copies (the HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this QR path runs on one GPU; multi-GPU row sharding is not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    import later_amd as later

    n = args.n
    m = args.m or n
    if args.copy_calibration:
        x = torch.empty(1 << 28, device="cuda", dtype=torch.float32)
        y = torch.empty_like(x)
        x.fill_(1.0)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("[bench] copied 3 x 1 GiB")
        return
    if args.single_batched:
        batched_config(later, torch, single=True)
        return
    if args.single:
        o = later.opts(args.single)
        A = later.generate_normal(m, n, seed=0, mat=0)
        R = later.colmajor(n, n)
        later.rgsqrf(A, R=R, o=o, work=later.workspace(m, n, o), hwork=later.hworkspace(m, n, o))
        torch.cuda.synchronize()
        log(f"[bench] one {args.single} factorisation of {m}x{n} done")
        return
    log(f"[bench] device {torch.cuda.get_device_name(0)}  m={m} n={n}  steps={args.steps} warmup={args.warmup}")
    A0 = later.generate_normal(m, n, seed=0, mat=0)
    torch.cuda.synchronize()

    res = {}
    for form in ("fp32", "mixed"):
        t0 = time.time()
        times, eager_ms, Q, R, work, o = time_form(later, torch, A0, form, args.steps, args.warmup,
                                                   graph=not args.no_graph)
        ms = sum(times) / len(times)
        entry = {"ms": ms, "ms_all": times, "eager_ms": eager_ms, "tflops_ref": later.flops_ref(m, n) / ms / 1e9,
                 "tflops_mgs": later.flops_mgs(m, n) / ms / 1e9}
        if not args.no_check:
            entry["orth"] = later.check_orth(Q)
            entry["orth_over_n"] = entry["orth"] / n
            entry["backward"] = later.check_backward(A0, Q, R)
        entry["prof"] = profile_step(later, torch, A0, form, work, o)
        res[form] = entry
        log(f"[bench] {form}: {ms:.2f} ms (eager {eager_ms:.2f})  {entry['tflops_ref']:.1f} TFLOP/s(ref)  "
            f"orth={entry.get('orth', float('nan')):.3e} bwd={entry.get('backward', float('nan')):.3e} "
            f"({time.time() - t0:.1f} s)")

    mixed, fp32 = res["mixed"], res["fp32"]
    # Dominant kernel of the mixed form: the streaming FP16 hi/lo kernel on
    # pre-split operands and 256-row tiles, gemm_lds_kernel<3, false, 256, 4>
    # (the node updates, the R2 R1 composition and the first pass's large
    # projections), timed per launch with HIP events on its stream.  achieved
    # = its algorithmic flops (2 m n1 n2 per update or projection, n^3 / 3 for
    # the triangular composition) summed, over its summed duration; each such
    # flop costs three MFMA flops (hi*hi, hi*lo, lo*hi).
    pr = mixed["prof"]
    k = pr["gemm_update_x3"]
    achieved = k["flops"] / (k["ms"] * 1e-3) / 1e12 if k["ms"] > 0 else 0.0
    traffic = read_traffic(k)
    roofline = {"bound": "mfma", "achieved": achieved, "peak": MFMA_PEAK_TFLOPS["f16x3"], "unit": "TFLOP/s",
                "frac": achieved / MFMA_PEAK_TFLOPS["f16x3"], "traffic": traffic,
                "kernel": "gemm_lds_kernel<3,false,256,4> (node updates A2 -= Q1 R12, R2 R1, first-pass Q1^T A2; "
                          "FP16 hi/lo pieces)",
                "launches": k["launches"], "kernel_ms": k["ms"], "avg_launch_us": 1e3 * k["ms"] / max(k["launches"], 1),
                "algorithmic_bytes_per_launch": k["bytes"] / max(k["launches"], 1),
                "mfma_flops_issued_tflops": 3.0 * achieved}
    # The second-largest: the int8 four-slice second-pass projections
    # (gemm_lds_kernel<6, false, 128, 4>); each algorithmic flop costs
    # thirteen int8 MFMA operations (the slice pairs of weight 2^-7g, g <= 4).
    k8 = pr.get("proj_i8", {"ms": 0.0, "launches": 0, "flops": 0.0, "bytes": 0.0})
    a8 = k8["flops"] / (k8["ms"] * 1e-3) / 1e12 if k8["ms"] > 0 else 0.0
    roofline_i8 = {"bound": "mfma", "achieved": a8, "peak": MFMA_PEAK_TFLOPS["i8"], "unit": "TOP/s",
                   "frac": 13.0 * a8 / MFMA_PEAK_TFLOPS["i8"],
                   "kernel": "gemm_lds_kernel<6,false,128,4> (second-pass Q1^T A2, int8 slices)",
                   "launches": k8["launches"], "kernel_ms": k8["ms"],
                   "avg_launch_us": 1e3 * k8["ms"] / max(k8["launches"], 1),
                   "int8_ops_issued_tops": 13.0 * a8,
                   "note": "achieved = algorithmic flops / s; frac = issued int8 ops / int8 peak"}

    vend = None if args.no_vendor else vendor_sgeqrf(torch, A0)
    if vend:
        log(f"[bench] rocSOLVER sgeqrf: {vend:.2f} ms  {later.flops_ref(m, n) / vend / 1e9:.1f} TFLOP/s(ref)")

    extra = None
    if not args.no_extra:
        extra = {
            "vendor_gemm_ceiling": blas_ceiling(torch),
            "C2": square_variants(later, torch, A0),
            "C1": single_config(later, torch, "C1", 8192, 1024, "uniform", ("reference", "mixed")),
            "C3a": single_config(later, torch, "C3a", 1048576, 128, "normal", ("reference", "mixed", "fp8", "bf16")),
            "C3b": single_config(later, torch, "C3b", 262144, 256, "normal", ("reference", "mixed", "fp8", "bf16")),
            "C4": batched_config(later, torch),
            "C5": serving_config(later, torch),
            "out_of_core": out_of_core_config(later, torch),
        }

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline sample n={args.cpu_sample}")
        cpu = cpu_baseline(args.cpu_sample, args.cpu_port_sample)

    out = {
        "metric": "QR TFLOP/s 16384^2 mixed vs plain FP32 at ||I-Q^TQ||<=1e-5; LS-solve p50/p99",
        "value": mixed["tflops_ref"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16x3+i8x4-mfma/f32+i32-acc/f64-panel",
        "data": "synthetic: Philox N(0,1) seed 0, generated on device",
        "config": {"workload": f"square QR m={m} n={n} FP32 column-major N(0,1), mixed-precision rgsqrf "
                               f"with second Gram-Schmidt pass", "m": m, "n": n},
        "timing": "graph replay of the captured factorisation" if not args.no_graph else "eager launches",
        "fp32_form": {"tflops_ref": fp32["tflops_ref"], "ms": fp32["ms"], "eager_ms": fp32["eager_ms"],
                      "orth": fp32.get("orth"),
                      "backward": fp32.get("backward")},
        "mixed_form": {"tflops_ref": mixed["tflops_ref"], "tflops_mgs": mixed["tflops_mgs"],
                       "eager_ms": mixed["eager_ms"],
                       "orth": mixed.get("orth"), "orth_over_n": mixed.get("orth_over_n"),
                       "backward": mixed.get("backward")},
        "speedup_vs_fp32": fp32["ms"] / mixed["ms"],
        "vendor_sgeqrf": ({"ms": vend, "tflops_ref": later.flops_ref(m, n) / vend / 1e9,
                           "what": "rocSOLVER sgeqrf (R + reflectors, no explicit Q), same input"}
                          if vend else None),
        "gates_met": (mixed.get("orth", 1.0) <= 1e-5 and mixed.get("backward", 1.0) <= 1e-6),
        "roofline": roofline,
        "roofline_second": roofline_i8,
        "cpu_baseline": cpu,
        "kernel_classes": {k: v for k, v in pr.items()},
        "other_workloads": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
