"""Least-squares solve and the continuous-batching server on the GPU.

Oracle: FP64 LAPACK least squares (numpy lstsq) on the same inputs.  The
server's results must be bitwise identical whether a request runs alone or
concurrently with others (its kernels depend only on its own shape).
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(later, A):
    import torch
    m, n = A.shape
    T = later.colmajor(m, n)
    T.copy_(torch.from_numpy(np.asarray(A, dtype=np.float32)))
    return T


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _graded(m, n, cond, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = cond ** (-np.arange(n) / max(n - 1, 1))
    return np.asfortranarray((U * s) @ V.T, dtype=np.float32)


def _problem(m, n, k, cond, seed):
    A = _graded(m, n, cond, seed)
    rng = np.random.default_rng(seed + 1)
    x = rng.standard_normal((n, k))
    b = A.astype(np.float64) @ x + 1e-3 * rng.standard_normal((m, k)) * np.linalg.norm(A) / np.sqrt(m * n)
    B = np.asfortranarray(b, dtype=np.float32)
    x64 = np.linalg.lstsq(A.astype(np.float64), B.astype(np.float64), rcond=None)[0]
    return A, B, x64


def _ls_bound(A, B, x64, eps=1e-6):
    """First-order LS perturbation bound for a backward error eps of the QR:
    kappa eps + kappa^2 eps ||r|| / (||A|| ||x||)."""
    A64 = A.astype(np.float64)
    s = np.linalg.svd(A64, compute_uv=False)
    kappa = s[0] / s[-1]
    r = B.astype(np.float64) - A64 @ x64
    return kappa * eps + kappa ** 2 * eps * np.linalg.norm(r) / (s[0] * np.linalg.norm(x64))


@pytest.mark.parametrize("n,k", [(64, 1), (300, 5), (1000, 16), (8192, 2)])
def test_trsm_upper(later, oracle, n, k):
    rng = np.random.default_rng(0)
    T = np.triu(rng.standard_normal((n, n))) / np.sqrt(n) + np.diag(1.0 + rng.random(n))
    T = np.asfortranarray(T, dtype=np.float32)
    B = np.asfortranarray(rng.standard_normal((n, k)), dtype=np.float32)
    Bd = _dev(later, B)
    later.trsm_upper(_dev(later, T), Bd)
    X = _np(Bd)
    ref = np.linalg.solve(T.astype(np.float64), B.astype(np.float64))
    assert np.linalg.norm(X - ref) / np.linalg.norm(ref) < 1e-6 * np.linalg.cond(T.astype(np.float64))


@pytest.mark.parametrize("m,n,k,cond", [(2000, 300, 4, 1e2), (4096, 1024, 16, 1e4), (777, 100, 1, 1e1)])
def test_gels(later, m, n, k, cond):
    A, B, x64 = _problem(m, n, k, cond, 0)
    X, R = later.gels(_dev(later, A), _dev(later, B), o=later.opts("mixed"))
    err = np.linalg.norm(_np(X) - x64) / np.linalg.norm(x64)
    assert err < 50 * 6e-8 * cond


def _run_server(later, problems, slots, budget):
    """Submits every problem at once and drains the server, fetching and
    releasing each result as soon as it is done (a finished request holds its
    pool memory until released, so waiting requests need the releases)."""
    import time
    import torch
    srv = later.Server(slots=slots, hbm_budget_bytes=budget)
    try:
        dev = [(_dev(later, A), _dev(later, B)) for A, B, _ in problems]
        torch.cuda.synchronize()
        ids = [srv.submit(Ad, Bd) for Ad, Bd in dev]
        out = [None] * len(problems)
        pending = set(range(len(problems)))
        t0 = time.time()
        while pending:
            srv.poll()
            for i in sorted(pending):
                r = srv.result(ids[i])
                if r["state"] in (3, 4):
                    assert r["state"] == 3, r
                    A, B, _ = problems[i]
                    X, R = srv.fetch(ids[i], A.shape[0], A.shape[1], B.shape[1])
                    out[i] = (_np(X), _np(R), r)
                    srv.release(ids[i])
                    pending.discard(i)
            assert time.time() - t0 < 120, "server did not drain"
        stats = srv.stats()
        assert stats["used_bytes"] == 0
        return out, stats
    finally:
        srv.close()


def test_server_mix_equals_alone(later):
    rng = np.random.default_rng(5)
    probs = []
    for i in range(12):
        m = int(rng.integers(128, 3000))
        n = int(rng.integers(16, min(m, 512)))
        k = int(rng.integers(1, 17))
        probs.append(_problem(m, n, k, 10 ** rng.uniform(0, 6), 100 + i))
    mixed, stats = _run_server(later, probs, slots=6, budget=2 << 30)
    for i, p in enumerate(probs):
        alone, _ = _run_server(later, [p], slots=1, budget=2 << 30)
        assert np.array_equal(alone[0][0], mixed[i][0]), i
        assert np.array_equal(alone[0][1], mixed[i][1]), i
    for (X, R, r), (A, B, x64) in zip(mixed, probs):
        err = np.linalg.norm(X - x64) / np.linalg.norm(x64)
        assert err < 10 * _ls_bound(A, B, x64) + 1e-6, (r, err)


def test_server_budget_queues_and_rejects(later):
    import torch
    probs = [_problem(2048, 256, 4, 10.0, i) for i in range(6)]
    # A budget that fits about two requests at a time: the rest must queue.
    need = 4 * (2 * 2048 * 256 + 256 * 256 + 2 * 2048 * 4 + 2 * 256 * 4) + 4 * later.workspace_size(2048, 256,
                                                                                                    later.opts("mixed"))
    out, stats = _run_server(later, probs, slots=6, budget=int(2.5 * need) + (4 << 20))
    assert stats["peak_bytes"] <= stats["pool_bytes"]
    assert stats["completed"] == 6
    srv = later.Server(slots=2, hbm_budget_bytes=8 << 20)
    try:
        A, B, _ = _problem(4096, 1024, 4, 10.0, 9)
        rid = srv.submit(_dev(later, A), _dev(later, B))
        r = srv.result(rid)
        assert r["state"] == 4 and r["status"] == later.ERR_BUDGET
        assert srv.stats()["rejected"] == 1
        srv.release(rid)
    finally:
        srv.close()


def test_server_cancel(later):
    """Cancelling a waiting request removes it from the queue at once; a
    running one finishes its enqueued stage and is retired at the next poll;
    a finished one cannot be cancelled.  The others' results are those they
    give alone, and the pool drains to zero."""
    import torch
    probs = [_problem(1500, 200, 3, 100.0, 40 + i) for i in range(4)]
    srv = later.Server(slots=1, hbm_budget_bytes=2 << 30)
    try:
        devs = [(_dev(later, A), _dev(later, B)) for A, B, _ in probs]
        ids = [srv.submit(A, B) for A, B in devs]
        # one slot: request 0 runs, 1..3 wait
        assert srv.result(ids[0])["state"] in (1, 2, 5)
        assert srv.result(ids[2])["state"] == 0
        srv.cancel(ids[2])
        assert srv.result(ids[2])["state"] == 6 and srv.result(ids[2])["status"] == later.ERR_CANCELLED
        srv.cancel(ids[0])  # running: retired after its current stage
        t0 = time.time()
        while any(srv.result(i)["state"] not in (3, 4, 6) for i in ids):
            srv.poll()
            assert time.time() - t0 < 60, "server did not drain"
        torch.cuda.synchronize()
        assert srv.result(ids[0])["state"] == 6
        assert [srv.result(i)["state"] for i in (ids[1], ids[3])] == [3, 3]
        with pytest.raises(later.LaterError):
            srv.cancel(ids[1])
        got = {}
        for j in (1, 3):
            A, B, _ = probs[j]
            X, R = srv.fetch(ids[j], A.shape[0], A.shape[1], B.shape[1])
            got[j] = (_np(X), _np(R))
        for i in ids:
            srv.release(i)
        assert srv.stats()["used_bytes"] == 0
    finally:
        srv.close()
    for j in (1, 3):
        alone, _ = _run_server(later, [probs[j]], slots=1, budget=2 << 30)
        assert np.array_equal(alone[0][0], got[j][0]) and np.array_equal(alone[0][1], got[j][1])
