"""Batched QR (grouped launches per step of the shared recursion tree) on the
GPU, against the single-matrix path and FP64 LAPACK on every matrix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _shapes(count, seed, mmax=1500):
    rng = np.random.default_rng(seed)
    m = np.exp(rng.uniform(np.log(64), np.log(mmax), count)).round().astype(int)
    n = np.array([rng.integers(16, min(mi, 256) + 1) for mi in m])
    return m, n


@pytest.mark.parametrize("form", ["reference", "mixed"])
def test_batch_matches_single(later, oracle, form):
    import torch
    m, n = _shapes(40, 1)
    m[:3] = [64, 300, 1000]
    n[:3] = [64, 33, 256]
    As = [later.generate_uniform(int(mi), int(ni), seed=0, mat=i) for i, (mi, ni) in enumerate(zip(m, n))]
    A0 = [a.clone().t().contiguous().t() for a in As]
    Rs = [later.colmajor(int(ni), int(ni)) for ni in n]
    o = later.opts(form)
    plan = later.BatchPlan(As, Rs, o)
    assert plan.launches > 0
    for rep in range(2):
        for a, a0 in zip(As, A0):
            a.copy_(a0)
        plan.execute()
        torch.cuda.synchronize()
        for i in range(len(As)):
            A = A0[i].cpu().numpy().copy(order="F")
            Q = As[i].cpu().numpy().copy(order="F")
            R = Rs[i].cpu().numpy().copy(order="F")
            Q64, R64 = oracle.lapack_qr64(A)
            As1 = A0[i].clone().t().contiguous().t()
            R1 = later.rgsqrf(As1, o=o)
            torch.cuda.synchronize()
            Rsgl = R1.cpu().numpy()
            e_sgl = oracle.rel_err(Rsgl, R64)
            assert np.array_equal(R, np.triu(R)), i
            assert oracle.rel_err(R, R64) < 2 * e_sgl + 2e-6, (i, m[i], n[i])
            bwd = oracle.backward(A, Q, R)
            assert bwd < (5e-4 if form == "reference" else 1e-6), (i, bwd)
            if form == "mixed":
                assert oracle.orth(Q) < 1e-5, i
    plan.close()


# ---------------------------------------------------------------- one persistent launch
def _batch_inputs(later, m, n, seed=0):
    As = [later.generate_uniform(int(mi), int(ni), seed=seed, mat=i) for i, (mi, ni) in enumerate(zip(m, n))]
    A0 = [a.clone().t().contiguous().t() for a in As]
    Rs = [later.colmajor(int(ni), int(ni)) for ni in n]
    return As, A0, Rs


@pytest.mark.parametrize("form,orth_max,bwd_max", [("mixed", 3e-5, 1e-6), ("fp32", 3e-5, 1e-6),
                                                   ("bf16", 0.3, 5e-2), ("reference", 0.3, 1e-2)])
def test_batch_wg_vs_lapack(later, oracle, form, orth_max, bwd_max):
    """The persistent one-launch kernel (a workgroup per matrix from the
    device queue): R against sign-normalised FP64 LAPACK, Q's orthogonality
    and the backward error, on uneven shapes including n = 1, n = m, n not a
    multiple of 16 and m not a multiple of 64; two executions give the same
    bits (the queue resets itself).  With FP32 products the orthogonality is
    a small multiple of the FP32 rounding level (1.1e-5 for the 77 x 77
    square case, 2.6e-6 at most over the benchmark's distribution); BF16 and
    the single-pass FP16 form are reported, not gated, at their precision."""
    import torch
    m, n = _shapes(48, 7, mmax=2500)
    m[:6] = [64, 300, 1000, 77, 513, 2048]
    n[:6] = [64, 1, 256, 77, 17, 200]
    As, A0, Rs = _batch_inputs(later, m, n)
    o = later.opts("mixed" if form == "mixed" else "reference")
    if form == "fp32":
        o = later.opts("fp32", reorth=later.REORTH_PASS)
    elif form == "bf16":
        o = later.opts("reference", gemm=later.GEMM_BF16, reorth=later.REORTH_PASS)
    plan = later.BatchWg(As, Rs, o)
    outs = []
    for rep in range(2):
        for a, a0 in zip(As, A0):
            a.copy_(a0)
        plan.execute()
        torch.cuda.synchronize()
        outs.append([(a.clone(), r.clone()) for a, r in zip(As, Rs)])
    assert int(plan.status.abs().sum()) == 0
    for (q0, r0), (q1, r1) in zip(*outs):
        assert torch.equal(q0, q1) and torch.equal(r0, r1)
    for i in range(len(As)):
        A = A0[i].cpu().numpy().copy(order="F")
        Q = As[i].cpu().numpy().copy(order="F")
        R = Rs[i].cpu().numpy().copy(order="F")
        assert np.array_equal(R, np.triu(R)), i
        Q64, R64 = oracle.lapack_qr64(A)
        assert oracle.orth(Q) < orth_max, (i, m[i], n[i], oracle.orth(Q))
        assert oracle.backward(A, Q, R) < bwd_max, (i, m[i], n[i])
        kappa = np.linalg.cond(R64)
        if form in ("mixed", "fp32"):
            assert oracle.rel_err(R, R64) < 1e-6 * kappa + 1e-6, (i, kappa)
        else:
            assert oracle.rel_err(R, R64) < 5e-2 * max(1.0, np.sqrt(n[i] / 16))
    plan.close()


def test_batch_wg_full_distribution_and_tuning(later, oracle):
    """The benchmark's 4096-matrix distribution (m log-uniform in [64, 4096],
    n uniform in [16, min(m, 256)]) in one launch, every variant of the
    tuning table forced in turn: gates with FP64 metrics on a sample of 64
    matrices, and the variants agree with each other to the FP32 level."""
    import torch
    rng = np.random.default_rng(0)
    cnt = 4096
    m = np.rint(np.exp(rng.uniform(np.log(64), np.log(4096), cnt))).astype(int)
    n = np.array([rng.integers(16, min(mi, 256) + 1) for mi in m])
    As, A0, Rs = _batch_inputs(later, m, n)
    sample = rng.choice(cnt, 64, replace=False)
    results = {}
    for v in (None, 0, 1, 2, 3):
        plan = later.BatchWg(As, Rs, later.opts("mixed"), tuning=None if v is None else [v] * later.BATCH_BUCKETS)
        for a, a0 in zip(As, A0):
            a.copy_(a0)
        plan.execute()
        torch.cuda.synchronize()
        assert int(plan.status.abs().sum()) == 0
        results[v] = [Rs[i].cpu().numpy().copy(order="F") for i in sample]
        for j, i in enumerate(sample):
            A = A0[i].cpu().numpy().copy(order="F")
            Q = As[i].cpu().numpy().copy(order="F")
            assert later.check_orth(As[i]) < 1e-5
            assert oracle.backward(A, Q, results[v][j]) < 1e-6
        plan.close()
    for v in (0, 1, 2, 3):
        for a, b in zip(results[None], results[v]):
            assert oracle.rel_err(a, b) < 1e-5
