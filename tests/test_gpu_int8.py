"""GPU parity of the int8 four-slice projections (LATER_GEMM_I8X4).

Each operand column is scaled by 2^e (max |x| 2^e in [32, 64)) and written as
s0 + s1 2^-7 + s2 2^-14 + s3 2^-21 with integer slices |s| <= 64; the
projection sums the thirteen slice products of weight 2^-7g, g = i + j <= 4,
in int32 (exact), combines the five sums in FP64 and scales by 2^-(e_i + e_j).
Every step but the last rounding is exact, so a NumPy model of the same
definition gives the kernel's bits (single K split: m < 512).  Longer
products are checked against the FP64 product with the representation bound
(each value carries at most 2^-27 of its column's largest magnitude, and the
dropped slice pairs weigh 2^-35), and the factorisation with int8
second-pass projections against the gates.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy().astype(np.float32, copy=False).copy(order="F")


def _dev(later, A):
    import torch
    m, n = A.shape
    T = later.colmajor(m, n)
    T.copy_(torch.from_numpy(np.asarray(A, dtype=np.float32)))
    return T


def _sync():
    import torch
    torch.cuda.synchronize()


def _slices(X):
    """(exponents, four int64 slice arrays) of the columns of X (float32)."""
    X = np.asarray(X, dtype=np.float32)
    mx = np.abs(X).max(axis=0)
    e = np.zeros(X.shape[1], dtype=np.int64)
    ok = (mx > 0) & np.isfinite(mx)
    e[ok] = np.clip(6 - np.frexp(mx[ok])[1], -126, 126)
    r = (X * np.ldexp(np.float32(1), e).astype(np.float32)).astype(np.float32)
    S = []
    for _ in range(4):
        s = np.rint(r).astype(np.float32)
        S.append(s.astype(np.int64))
        r = ((r - s) * np.float32(128)).astype(np.float32)
    return e, S


def _model(Q, B):
    """The int8 projection Q^T B, rounded to FP32 as the kernel does."""
    eq, SQ = _slices(Q)
    eb, SB = _slices(B)
    G = [sum(SQ[a].T @ SB[g - a] for a in range(max(0, g - 3), min(g, 3) + 1)) for g in range(5)]
    for g in G:
        assert np.abs(g).max() < 2 ** 31
    v = (((G[0].astype(np.float64) + G[1] * 2.0 ** -7) + G[2] * 2.0 ** -14) + G[3] * 2.0 ** -21) + G[4] * 2.0 ** -28
    return (v * np.ldexp(1.0, -(eq[:, None] + eb[None, :]))).astype(np.float32)


@pytest.mark.parametrize("m,n1,n2", [(300, 100, 36), (511, 129, 257), (64, 64, 64), (200, 7, 5), (448, 256, 128)])
def test_int8_proj_bits(later, oracle, m, n1, n2):
    Q = oracle.gen_normal(m, n1, 1, 0)
    B = oracle.gen_uniform(m, n2, 2, 0)
    # A zero column, a column spanning many binades, a huge and a tiny one.
    B[:, 0] = 0.0
    if n2 > 2:
        B[:, 1] *= (2.0 ** ((np.arange(m) % 40) - 20)).astype(np.float32)
        B[:, 2] *= np.float32(1e30)
    Q[:, -1] *= np.float32(1e-30)
    C = later.colmajor(n1, n2)
    later.gemm_proj(_dev(later, Q), _dev(later, B), C, later.GEMM_I8X4, 0, pieces=True)
    _sync()
    assert np.array_equal(_np(C), _model(Q, B))


@pytest.mark.parametrize("m,n1,n2", [(16384, 256, 256), (4096, 640, 128), (131008, 64, 64)])
def test_int8_proj_envelope(later, oracle, m, n1, n2):
    Q = oracle.gen_normal(m, n1, 3, 0)
    B = oracle.gen_uniform(m, n2, 4, 0)
    C = later.colmajor(n1, n2)
    later.gemm_proj(_dev(later, Q), _dev(later, B), C, later.GEMM_I8X4, 0, pieces=True)
    _sync()
    Qd, Bd = Q.astype(np.float64), B.astype(np.float64)
    ref = Qd.T @ Bd
    # Representation: |x - x~| <= 2^-27 max|column| per value; dropped
    # pairs below 2^-33 max max; split-K partials rounded to FP32.
    mq, mb = np.abs(Qd).max(axis=0), np.abs(Bd).max(axis=0)
    bound = 2.0 ** -25 * m * mq[:, None] * mb[None, :] + 2.0 ** -22 * np.abs(Qd).T @ np.abs(Bd)
    err = np.abs(_np(C).astype(np.float64) - ref)
    assert (err <= bound).all(), float((err / bound).max())
    # and far tighter than the bound in the mean: the slices carry ~27 bits
    assert np.linalg.norm(err) / np.linalg.norm(np.abs(Qd).T @ np.abs(Bd)) < 2.0 ** -24


def test_int8_proj_needs_staging_and_short_k(later):
    """The slices are staged in hwork, and exact int32 sums need K <= 131008
    rows."""
    import torch
    m = 131008 + 64
    Q = torch.zeros((16, m), device="cuda").t()
    B = torch.zeros((16, m), device="cuda").t()
    C = later.colmajor(16, 16)
    with pytest.raises(later.LaterError) as e:
        later.gemm_proj(Q, B, C, later.GEMM_I8X4, 0, pieces=True)
    assert e.value.code == later.ERR_SHAPE
    with pytest.raises(later.LaterError) as e:
        later.gemm_proj(Q[:1000], B[:1000], C, later.GEMM_I8X4, 0, pieces=False)
    assert e.value.code == later.ERR_WORKSPACE


def test_int8_second_pass_gates(later):
    """The mixed form with int8 second-pass projections: the gates, the same
    bits on a second run, and different bits from the three-piece form (so
    the int8 kernels ran)."""
    import torch
    n = 4096
    A0 = later.generate_normal(n, n, 0, 0)
    o = later.opts("mixed", gemm_pass2=later.GEMM_I8X4)
    out = []
    for _ in range(2):
        A = A0.clone()
        R = later.rgsqrf(A, o=o)
        _sync()
        out.append((A, R))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
    Q, R = out[0]
    assert later.check_orth(Q) < 1e-5
    assert later.check_backward(A0, Q, R) < 1e-6
    A6 = A0.clone()
    R6 = later.rgsqrf(A6, o=later.opts("mixed", gemm_pass2=later.GEMM_F16X6))
    _sync()
    assert not torch.equal(R6, R)


def test_int8_tall_uses_three_pieces(later):
    """Above 131008 rows a node's int8 projection runs the three-piece FP16
    split instead: the same bits as asking for it."""
    import torch
    m, n = 140000, 256
    A0 = later.generate_normal(m, n, 5, 0)
    res = []
    for p in (later.GEMM_I8X4, later.GEMM_F16X6):
        A = A0.clone()
        R = later.rgsqrf(A, o=later.opts("mixed", gemm_pass2=p))
        _sync()
        res.append((A, R))
    assert torch.equal(res[0][0], res[1][0]) and torch.equal(res[0][1], res[1][1])
    assert later.check_orth(res[0][0]) < 1e-5


def test_int8_unaligned_columns(later):
    """lda = m + 3 and m not a multiple of 64: every column starts off a
    16-byte boundary, so the splits take their scalar paths; the mixed form
    still meets the gates and repeats its bits."""
    import torch
    m, n = 3001, 1100
    A0 = later.generate_normal(m, n, 7, 0)
    out = []
    for _ in range(2):
        Ap = later.colmajor(m, n, ld=m + 3)
        Ap.copy_(A0)
        R = later.rgsqrf(Ap, o=later.opts("mixed"))
        _sync()
        out.append((Ap, R))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
    Q, R = out[0]
    assert later.check_orth(Q) < 1e-5
    assert later.check_backward(A0, Q, R) < 1e-6
