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
