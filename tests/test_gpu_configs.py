"""GPU parity at the benchmark configurations and for the kernels that only
the large nodes select.

- Tall-skinny QR at full size (1048576 x 128, 262144 x 256, N(0,1) seed 0):
  FP64 metrics on the device against the gates (mixed form) and against the
  oracle's restatement of the reference recursion (reference form); one tree
  level smaller (131072 x 128, 65536 x 256): R against FP64 LAPACK and Q
  against the restatement.
- The streaming GEMM kernels on pre-split 16-bit pieces, element-wise against
  the FP64 product envelope and bitwise against the in-kernel conversion.
- The FP32 -> FP16 / BF16 operand cast (util/util.cu:24-32), bit-exact over
  all 2^32 FP32 patterns.
- The BF16 node-GEMM form against the restatement; the 4096^2 mixed form's R
  against LAPACK.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24


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


# ---------------------------------------------------------------- tall-skinny
@pytest.mark.parametrize("m,n", [(1048576, 128), (262144, 256)])
@pytest.mark.parametrize("form", ["reference", "fp32", "mixed"])
def test_tall_skinny_full_size(later, oracle, m, n, form):
    """Gates with FP64 accumulation on the device: mixed ||I - Q^T Q||_F <= 1e-5
    and backward error <= 1e-6; fp32 backward error <= 1e-6; the reference
    form within 1.5x of the restatement's own metrics on the same input."""
    print(f"tall-skinny {m}x{n} {form}", flush=True)
    A0 = later.generate_normal(m, n, 0, 0)
    A = A0.clone()
    R = later.rgsqrf(A, o=later.opts(form))
    _sync()
    orth, bwd = later.check_orth(A), later.check_backward(A0, A, R)
    if form == "mixed":
        assert orth <= 1e-5 and bwd <= 1e-6, (orth, bwd)
    elif form == "fp32":
        assert bwd <= 1e-6, bwd
    else:
        An = _np(A0)
        Qo, Ro = oracle.rgsqrf(An, "f16")
        assert bwd < 1.5 * oracle.backward(An, Qo, Ro) + 1e-7
        assert orth < 1.5 * oracle.orth(Qo) + 1e-6
    Rn = _np(R)
    assert np.array_equal(Rn, np.triu(Rn)) and np.all(np.diag(Rn) > 0)


@pytest.mark.parametrize("m,n", [(131072, 128), (65536, 256)])
@pytest.mark.parametrize("form,prec", [("reference", "f16"), ("fp32", "f32"), ("mixed", None)])
def test_tall_skinny_vs_lapack(later, oracle, m, n, form, prec):
    """One tree level below the full size: R against sign-normalised FP64
    LAPACK; Q and R against the restatement for the reference forms."""
    A = oracle.gen_normal(m, n, 0, 0)
    Ad = _dev(later, A)
    R = later.rgsqrf(Ad, o=later.opts(form))
    _sync()
    Q, Rg = _np(Ad), _np(R)
    Q64, R64 = oracle.lapack_qr64(A)
    if prec is None:
        assert oracle.rel_err(Rg, R64) < 1e-6
        assert oracle.rel_err(Q, Q64) < 1e-6
        assert oracle.orth(Q) < 1e-6
        return
    Qo, Ro = oracle.rgsqrf(A, prec)
    e_or = oracle.rel_err(Ro, R64)
    if prec == "f32":
        assert oracle.rel_err(Rg, Ro) < 2e-6
        assert oracle.rel_err(Q, Qo) < 1e-5
        assert oracle.rel_err(Rg, R64) < 2e-6
    else:
        assert oracle.rel_err(Rg, Ro) < 0.6 * e_or + 1e-6
        assert oracle.rel_err(Rg, R64) < 1.5 * e_or + 1e-6
        assert oracle.rel_err(Q, Qo) < 2 * oracle.rel_err(Qo, Q64) + 1e-6


# ---------------------------------------------------------------- streaming GEMMs
def _envelope(X, prec):
    import torch
    t = torch.from_numpy(np.asarray(X, dtype=np.float32))
    if prec == 1:
        t = t.half().float()
    elif prec == 2:
        t = t.bfloat16().float()
    return t.numpy().astype(np.float64)


@pytest.mark.parametrize("m,n1,n2,prec,chunk", [
    (16384, 1024, 1024, 3, 256),   # folded hi/lo projection, 128-row streaming tiles
    (16384, 1024, 1024, 4, 32),    # three-piece projection
    (16384, 1024, 1024, 3, 32),
    (8192, 1024, 512, 1, 0),       # FP16 and BF16 pieces
    (8192, 1024, 512, 2, 0),
    (4096, 40, 1024, 3, 256),      # thin: register-staged kernel on pieces
])
def test_streaming_projection(later, oracle, m, n1, n2, prec, chunk):
    Q = oracle.gen_normal(m, n1, 1, 0) * np.float32(1.0 / np.sqrt(m))
    B = oracle.gen_normal(m, n2, 2, 0)
    Qd, Bd = _dev(later, Q), _dev(later, B)
    Cp = later.colmajor(n1, n2)
    Cf = later.colmajor(n1, n2)
    later.gemm_proj(Qd, Bd, Cp, prec, chunk, pieces=True)
    later.gemm_proj(Qd, Bd, Cf, prec, chunk, pieces=False)
    _sync()
    Cg = _np(Cp).astype(np.float64)
    assert np.array_equal(_np(Cp), _np(Cf))  # staged and fused operands: same bits
    scale = np.abs(Q.astype(np.float64)).T @ np.abs(B.astype(np.float64))
    if prec in (3, 4):
        ref = Q.astype(np.float64).T @ B.astype(np.float64)
        tol = 4e-6 if prec == 3 else 4 * U32 * np.sqrt(chunk)
    else:
        ref = _envelope(Q, prec).T @ _envelope(B, prec)
        tol = 4 * U32 * np.sqrt(m)
    err = np.abs(Cg - ref) / scale
    assert err.max() < 8 * tol, err.max()
    assert np.linalg.norm(Cg - ref) / np.linalg.norm(scale) < tol


@pytest.mark.parametrize("m,n1,n2,prec", [
    (4096, 2048, 2048, 3),   # hi/lo update, 256 x 128 streaming tiles (eight waves)
    (4096, 1024, 1024, 3),   # fewer tiles than CUs: 128-row streaming tiles
    (4096, 1024, 2048, 1),
    (4096, 1024, 2048, 2),
    (4096, 512, 1024, 4),
])
def test_streaming_update(later, oracle, m, n1, n2, prec):
    Q = oracle.gen_normal(m, n1, 1, 0) * np.float32(1.0 / np.sqrt(m))
    C = oracle.gen_normal(n1, n2, 3, 0)
    B = oracle.gen_normal(m, n2, 2, 0)
    Qd, Cd = _dev(later, Q), _dev(later, C)
    Bp, Bf = _dev(later, B), _dev(later, B)
    later.gemm_update(Qd, Cd, Bp, prec, pieces=True)
    later.gemm_update(Qd, Cd, Bf, prec, pieces=False)
    _sync()
    assert np.array_equal(_np(Bp), _np(Bf))
    Bg = _np(Bp).astype(np.float64)
    if prec in (3, 4):
        ref = B.astype(np.float64) - Q.astype(np.float64) @ C.astype(np.float64)
        tol = 4e-6 if prec == 3 else 4 * U32 * np.sqrt(n1)
    else:
        ref = B.astype(np.float64) - _envelope(Q, prec) @ _envelope(C, prec)
        tol = 4 * U32 * np.sqrt(n1)
    scale = np.abs(B.astype(np.float64)) + np.abs(Q.astype(np.float64)) @ np.abs(C.astype(np.float64))
    err = np.abs(Bg - ref) / scale
    assert err.max() < 8 * tol, err.max()
    assert np.linalg.norm(Bg - ref) / np.linalg.norm(scale) < tol


# ---------------------------------------------------------------- the cast
@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("pieces", [False, True])
def test_cast_all_fp32_patterns(later, oracle, prec, pieces):
    """C = Q^T B with m = 1, Q = [1]: every output is the 16-bit cast of one
    B entry, returned in FP32.  All 2^32 patterns, compared on the device with
    torch's device cast; that cast itself is compared with the oracle's
    (util/util.cu:24-32, RN-even; overflow to Inf, FP16 subnormals) on every
    61st pattern, which covers every exponent.  Zeros compare by value (the
    MFMA adds the product to +0, so -0 comes back as +0); NaN stays NaN."""
    import torch
    N = 65535 * 64
    Q = torch.ones((1, 1), device="cuda", dtype=torch.float32)
    B = later.colmajor(1, N)
    C = later.colmajor(1, N)
    bad = 0
    total = 1 << 32
    for start in range(0, total, N):
        cnt = min(N, total - start)
        bits = torch.arange(start, start + cnt, device="cuda", dtype=torch.int64)
        bits = torch.where(bits >= (1 << 31), bits - (1 << 32), bits).to(torch.int32)
        x = bits.view(torch.float32)
        Bv = B[:, :cnt]
        Bv.copy_(x.view(1, cnt))
        Cv = C[:, :cnt]
        later.gemm_proj(Q, Bv, Cv, prec, 0, pieces=pieces)
        y = Cv.reshape(cnt)
        e = (x.half() if prec == 1 else x.bfloat16()).float()
        same = (y == e) | (torch.isnan(y) & torch.isnan(e))
        bad += int((~same).sum())
    assert bad == 0
    # torch's device cast against the oracle
    sample = np.arange(0, total, 61, dtype=np.uint64).astype(np.uint32)
    xs = sample.view(np.float32)
    dev = torch.from_numpy(xs).cuda()
    if prec == 1:
        got = (dev.half().view(torch.int16).cpu().numpy().view(np.uint16))
        want = oracle.f32_to_f16_bits(xs)
        nan = np.isnan(xs)
        assert np.array_equal(got[~nan], want[~nan])
        assert np.all((want[nan] & 0x7C00) == 0x7C00) and np.all((got[nan] & 0x7C00) == 0x7C00)
    else:
        got = dev.bfloat16().view(torch.int16).cpu().numpy().view(np.uint16)
        L = oracle.lib()
        want = np.array([L.oracle_f32_to_bf16(float(v)) for v in xs[:: 4099]], dtype=np.uint16)
        g = got[:: 4099]
        nan = np.isnan(xs[:: 4099])
        assert np.array_equal(g[~nan], want[~nan])


# ---------------------------------------------------------------- forms
@pytest.mark.parametrize("m,n", [(2048, 512), (4096, 1024)])
def test_rgsqrf_bf16_form(later, oracle, m, n):
    """BF16 node GEMMs (the reference form with BF16 operands) against the
    restatement with the same BF16 rounding."""
    A = oracle.gen_uniform(m, n, 0, 0)
    Ad = _dev(later, A)
    R = later.rgsqrf(Ad, o=later.opts("reference", gemm=later.GEMM_BF16, gemm_pass2=later.GEMM_BF16))
    _sync()
    Q, Rg = _np(Ad), _np(R)
    Qo, Ro = oracle.rgsqrf(A, "bf16")
    Q64, R64 = oracle.lapack_qr64(A)
    e_or = oracle.rel_err(Ro, R64)
    assert oracle.rel_err(Rg, Ro) < 0.6 * e_or + 1e-6
    assert oracle.rel_err(Rg, R64) < 1.5 * e_or + 1e-6
    assert oracle.backward(A, Q, Rg) < 1.5 * oracle.backward(A, Qo, Ro)


def test_mixed_square_4096_vs_lapack(later, oracle):
    """The 4096^2 mixed form (256-row streaming update tiles, XCD-balanced
    triangular composition) against sign-normalised FP64 LAPACK."""
    n = 4096
    A = oracle.gen_normal(n, n, 0, 0)
    Ad = _dev(later, A)
    R = later.rgsqrf(Ad, o=later.opts("mixed"))
    _sync()
    Q, Rg = _np(Ad), _np(R)
    Q64, R64 = oracle.lapack_qr64(A)
    kappa = np.linalg.cond(A.astype(np.float64))
    assert oracle.orth(Q) < 1e-5
    assert oracle.backward(A, Q, Rg) < 1e-6
    assert oracle.rel_err(Rg, R64) < 2e-6 * max(1.0, kappa / 1000)
