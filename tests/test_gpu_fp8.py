"""FP8 e4m3 projections (LATER_GEMM_FP8) and the tall-skinny forms built on
them, on the GPU.

- The product C = Q^T B on staged e4m3 operands, element-wise against the
  FP64 product of the oracle's block quantisation of the same operands
  (e4m3 round-to-nearest-even after a power-of-two scale per 128 rows of
  each column, pinned in tests/test_oracle.py against torch's float8_e4m3fn
  cast): only the FP32 accumulation separates the two.
- rgsqrf in the 'fp8' and 'bf16' forms (first-pass projections in e4m3 or
  BF16, hi/lo updates, Cholesky QR leaves; the second pass as refinement)
  against the gates and FP64 LAPACK, at the tall-skinny configurations
  (1048576 x 128, 262144 x 256) and one tree level smaller.
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


@pytest.mark.parametrize("m,n1,n2,kind", [
    (4096, 128, 128, "normal"),
    (16384, 64, 64, "normal"),        # thin, split-K
    (1000, 40, 72, "normal"),         # ragged K and sides
    (8192, 512, 1024, "graded"),      # 256-row tiles; columns spanning 2^-20 .. 1 inside one block
    (131072, 128, 128, "normal"),
])
def test_fp8_projection(later, oracle, m, n1, n2, kind):
    import torch
    print(f"fp8 projection {m} {n1} {n2} {kind}", flush=True)
    Q = oracle.gen_normal(m, n1, 1, 0)
    B = oracle.gen_normal(m, n2, 2, 0)
    if kind == "graded":
        Q = Q * np.exp2(-20.0 * np.random.default_rng(0).random((m, 1))).astype(np.float32)
        B = B * np.float32(1e30)
    Qd, Bd = _dev(later, Q), _dev(later, B)
    C = later.colmajor(n1, n2)
    later.gemm_proj(Qd, Bd, C, later.GEMM_FP8, 0, pieces=True)
    torch.cuda.synchronize()
    Cg = _np(C).astype(np.float64)
    Qq, Bq = oracle.fp8_quantize(Q), oracle.fp8_quantize(B)
    ref = Qq.T @ Bq
    scale = np.abs(Qq).T @ np.abs(Bq)
    err = np.abs(Cg - ref) / scale
    tol = 4 * U32 * np.sqrt(m)
    assert err.max() < 8 * tol, err.max()
    assert np.linalg.norm(Cg - ref) / np.linalg.norm(scale) < tol
    # the quantisation is far coarser than the accumulation: a product of the
    # unquantised operands is not what the kernel computed
    exact = Q.astype(np.float64).T @ B.astype(np.float64)
    assert np.linalg.norm(Cg - exact) > 100 * np.linalg.norm(Cg - ref)


def test_fp8_projection_needs_staging(later):
    import torch
    Q = later.generate_normal(1024, 64, 1, 0)
    C = later.colmajor(64, 64)
    with pytest.raises(later.LaterError) as e:
        later.gemm_proj(Q, Q, C, later.GEMM_FP8, 0, pieces=False)
    assert e.value.code == later.ERR_WORKSPACE
    with pytest.raises(later.LaterError) as e:
        later.gemm_update(Q, C, Q.clone(), later.GEMM_FP8)
    assert e.value.code == later.ERR_OPTION
    with pytest.raises(later.LaterError) as e:
        later.rgsqrf(later.generate_normal(1024, 256, 1, 0), o=later.opts("fp8"), hwork=None)
    assert e.value.code == later.ERR_WORKSPACE
    torch.cuda.synchronize()


@pytest.mark.parametrize("m,n", [(1048576, 128), (262144, 256)])
@pytest.mark.parametrize("form", ["fp8", "bf16"])
def test_tall_skinny_refined_full_size(later, m, n, form):
    """Gates at the full tall-skinny sizes with FP64 accumulation on the
    device: ||I - Q^T Q||_F <= 1e-5 and ||A - QR||_F / ||A||_F <= 1e-6."""
    import torch
    print(f"tall-skinny {form} {m}x{n}", flush=True)
    A0 = later.generate_normal(m, n, 0, 0)
    A = A0.clone()
    R = later.rgsqrf(A, o=later.opts(form))
    torch.cuda.synchronize()
    orth, bwd = later.check_orth(A), later.check_backward(A0, A, R)
    assert orth <= 1e-5 and bwd <= 1e-6, (orth, bwd)
    Rn = _np(R)
    assert np.array_equal(Rn, np.triu(Rn)) and np.all(np.diag(Rn) > 0)


@pytest.mark.parametrize("m,n", [(131072, 128), (65536, 256), (20000, 300)])
@pytest.mark.parametrize("form", ["fp8", "bf16"])
def test_tall_skinny_refined_vs_lapack(later, oracle, m, n, form):
    import torch
    A = oracle.gen_normal(m, n, 0, 0)
    Ad = _dev(later, A)
    R = later.rgsqrf(Ad, o=later.opts(form))
    torch.cuda.synchronize()
    Q, Rg = _np(Ad), _np(R)
    Q64, R64 = oracle.lapack_qr64(A)
    assert oracle.rel_err(Rg, R64) < 1e-6
    assert oracle.rel_err(Q, Q64) < 1e-6
    assert oracle.orth(Q) < 1e-6
