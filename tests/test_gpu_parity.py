"""GPU parity: the HIP library (through the C-ABI) against the oracle.

Tolerances are stated per test.  The oracle's GEMMs accumulate in FP64 while
the MFMA kernels accumulate in FP32, so floating-point results agree to a few
units of FP32 rounding scaled by the problem (accumulation length, condition
number), never bit for bit; integer work (the generator) is bit-exact.
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


# ---------------------------------------------------------------- generator
@pytest.mark.parametrize("m,n,seed,mat", [(1000, 37, 0, 0), (4096, 3, 7, 11), (1, 1, 0, 5)])
def test_generate_uniform_bit_exact(later, oracle, m, n, seed, mat):
    G = later.generate_uniform(m, n, seed, mat)
    _sync()
    assert np.array_equal(_np(G), oracle.gen_uniform(m, n, seed, mat))


def test_generate_normal_matches(later, oracle):
    m, n = 4096, 9
    G = _np(later.generate_normal(m, n, 0, 0))
    ref = oracle.gen_normal(m, n, 0, 0)
    # Device and host libm may differ in the last FP64 bit: at most 1 FP32 ulp.
    ulp = np.abs(G.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1 and (ulp > 0).mean() < 1e-3


# ---------------------------------------------------------------- GEMMs
def _round(X, prec):
    import torch
    t = torch.from_numpy(np.asarray(X, dtype=np.float32))
    if prec == 1:
        t = t.half().float()
    elif prec == 2:
        t = t.bfloat16().float()
    return t.numpy().astype(np.float64)


PRECS = [(0, "f32"), (1, "f16"), (2, "bf16"), (3, "f16x3"), (4, "f16x6")]


@pytest.mark.parametrize("m,n1,n2", [(4096, 128, 128), (1000, 100, 36), (16384, 32, 32), (300, 7, 5),
                                     (2048, 256, 640)])
@pytest.mark.parametrize("prec,name", PRECS)
@pytest.mark.parametrize("chunk", [0, 256])
def test_gemm_proj(later, oracle, m, n1, n2, prec, name, chunk):
    Q = oracle.gen_uniform(m, n1, 1, 0) * np.float32(0.05)
    B = oracle.gen_uniform(m, n2, 2, 0)
    C = later.colmajor(n1, n2)
    later.gemm_proj(_dev(later, Q), _dev(later, B), C, prec, chunk)
    _sync()
    Cg = _np(C).astype(np.float64)
    if prec in (3, 4):
        # Split products of the FP32 operands: hi/lo (2^-22) or three pieces
        # (2^-33), FP32 accumulation (folded into FP64 every chunk rows).
        ref = Q.astype(np.float64).T @ B.astype(np.float64)
        tol = 4e-6 if prec == 3 else 4 * U32 * np.sqrt(chunk if chunk else m)
    else:
        ref = _round(Q, prec).T @ _round(B, prec)
        tol = 4 * U32 * np.sqrt(m)
    scale = np.abs(Q.astype(np.float64)).T @ np.abs(B.astype(np.float64))
    err = np.linalg.norm(Cg - ref) / np.linalg.norm(scale)
    assert err < tol, (name, err)


@pytest.mark.parametrize("m,n1,n2", [(4096, 128, 128), (1000, 100, 36), (16384, 32, 32), (300, 7, 5),
                                     (2048, 640, 256)])
@pytest.mark.parametrize("prec,name", PRECS)
def test_gemm_update(later, oracle, m, n1, n2, prec, name):
    Q = oracle.gen_uniform(m, n1, 1, 0) * np.float32(0.05)
    C = oracle.gen_uniform(n1, n2, 3, 0)
    B = oracle.gen_uniform(m, n2, 2, 0)
    Bd = _dev(later, B)
    later.gemm_update(_dev(later, Q), _dev(later, C), Bd, prec)
    _sync()
    Bg = _np(Bd).astype(np.float64)
    if prec in (3, 4):
        ref = B.astype(np.float64) - Q.astype(np.float64) @ C.astype(np.float64)
        tol = 4e-6 if prec == 3 else 4 * U32 * np.sqrt(n1)
    else:
        ref = B.astype(np.float64) - _round(Q, prec) @ _round(C, prec)
        tol = 4 * U32 * np.sqrt(n1)
    scale = np.abs(B.astype(np.float64)) + np.abs(Q.astype(np.float64)) @ np.abs(C.astype(np.float64))
    err = np.linalg.norm(Bg - ref) / np.linalg.norm(scale)
    assert err < tol, (name, err)


# ---------------------------------------------------------------- panel
@pytest.mark.parametrize("m,w", [(256, 32), (4096, 32), (16384, 32), (1000, 32), (287, 32), (300, 7),
                                 (64, 1), (70000, 32)])
@pytest.mark.parametrize("fp64", [0, 1, 2])
def test_caqr_panel(later, oracle, m, w, fp64):
    A = oracle.gen_uniform(m, w, 0, 0)
    Ad = _dev(later, A)
    R = later.caqr_panel(Ad, fp64=fp64)
    _sync()
    Qg, Rg = _np(Ad), _np(R)
    Qo, Ro = oracle.caqr32(A)
    Q64, R64 = oracle.lapack_qr64(A)
    assert oracle.rel_err(Rg, Ro) < 2e-6
    assert oracle.rel_err(Qg, Qo) < 4e-6
    assert oracle.rel_err(Rg, R64) < 2e-6
    assert np.array_equal(Rg, np.triu(Rg)) and np.all(np.diag(Rg) > 0)
    assert oracle.orth(Qg) < (2e-6 if fp64 else 4e-6) * np.sqrt(w)


# ---------------------------------------------------------------- rgsqrf
def _run(later, A, form, **over):
    Ad = _dev(later, A)
    o = later.opts(form, **over)
    R, info = later.rgsqrf(Ad, o=o, info=True)
    _sync()
    assert info.status == 0
    return _np(Ad), _np(R), info


@pytest.mark.parametrize("m,n", [(1024, 128), (2048, 256), (4096, 512), (1000, 100), (300, 33), (256, 256)])
def test_rgsqrf_fp32_form(later, oracle, m, n):
    A = oracle.gen_uniform(m, n, 0, 0)
    Q, R, _ = _run(later, A, "fp32")
    Qo, Ro = oracle.rgsqrf(A, "f32")
    Q64, R64 = oracle.lapack_qr64(A)
    kappa = np.linalg.cond(A.astype(np.float64))
    assert oracle.rel_err(R, Ro) < 2e-6 * max(1.0, kappa / 100)
    assert oracle.rel_err(R, R64) < 2e-6 * max(1.0, kappa / 100)
    assert oracle.rel_err(Q, Q64) < 2e-6 * max(1.0, kappa / 10)
    assert oracle.backward(A, Q, R) < 1e-6
    assert np.array_equal(R, np.triu(R))


@pytest.mark.parametrize("m,n", [(2048, 256), (4096, 512), (1000, 200)])
def test_rgsqrf_reference_form_f16(later, oracle, m, n):
    """FP16 node GEMMs (QR/later_rgsqrf.cu:43-56): GPU vs the restatement."""
    A = oracle.gen_uniform(m, n, 0, 0)
    Q, R, _ = _run(later, A, "reference")
    Qo, Ro = oracle.rgsqrf(A, "f16")
    Q64, R64 = oracle.lapack_qr64(A)
    e_gpu, e_or = oracle.rel_err(R, R64), oracle.rel_err(Ro, R64)
    # Both carry the FP16 operand error (~u16 = 4.9e-4 scaled); they agree
    # with each other far below it.
    assert oracle.rel_err(R, Ro) < 0.2 * e_or + 1e-5
    assert e_gpu < 2 * e_or + 1e-5
    assert abs(oracle.backward(A, Q, R) - oracle.backward(A, Qo, Ro)) < 0.5 * oracle.backward(A, Qo, Ro)


def test_rgsqrf_c1_reference_config(later, oracle):
    """Config C1: m=8192 n=1024 U[-1,1] seed 0, cutoff 128, FP16 GEMMs."""
    m, n = 8192, 1024
    Ad = later.generate_uniform(m, n, 0, 0)
    A = _np(Ad)
    R = later.rgsqrf(Ad, o=later.opts("reference"))
    Q, Rg = _np(Ad), _np(R)
    Qo, Ro = oracle.rgsqrf(A, "f16")
    Q64, R64 = oracle.lapack_qr64(A)
    e_or = oracle.rel_err(Ro, R64)
    # Two independent FP16 rounding realisations (GPU and restatement).
    assert oracle.rel_err(Rg, Ro) < 0.6 * e_or
    assert oracle.rel_err(Rg, R64) < 1.5 * e_or
    bg, bo = oracle.backward(A, Q, Rg), oracle.backward(A, Qo, Ro)
    assert bg < 1.5 * bo
    assert oracle.orth(Q) < 1.5 * oracle.orth(Qo)


@pytest.mark.parametrize("m,n,kind", [(2048, 2048, "normal"), (4096, 1024, "uniform"), (1000, 300, "uniform")])
def test_rgsqrf_mixed_form(later, oracle, m, n, kind):
    A = oracle.gen_normal(m, n, 0, 0) if kind == "normal" else oracle.gen_uniform(m, n, 0, 0)
    Q, R, _ = _run(later, A, "mixed")
    Q64, R64 = oracle.lapack_qr64(A)
    Qh = np.linalg.qr(A)[0]
    orth_h = oracle.orth(Qh)  # FP32 Householder's own orthogonality
    # The second pass's FP16 hi/lo products with FP32 chunk sums leave a
    # floor a few times above Householder's (measured on the CPU model: ~3x).
    assert oracle.orth(Q) < 6 * orth_h
    assert oracle.backward(A, Q, R) < 1e-6
    kappa = np.linalg.cond(A.astype(np.float64))
    assert oracle.rel_err(R, R64) < 2e-6 * max(1.0, kappa / 100)


def test_rgsqrf_graded_mixed_meets_gates(later, oracle):
    """A = U diag(s) V^T, s geometric 1..1e-4 (config C2 (ii) at n=1024)."""
    n = 1024
    rng = np.random.default_rng(0)
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = 10.0 ** (-4 * np.arange(n) / (n - 1))
    A = np.asfortranarray((U * s) @ V.T, dtype=np.float32)
    Q, R, _ = _run(later, A, "mixed")
    assert oracle.orth(Q) < 1e-5
    assert oracle.backward(A, Q, R) < 1e-6
    Qf, Rf, _ = _run(later, A, "fp32")
    assert oracle.orth(Qf) > 1e-4  # one pass of block CGS loses orthogonality here


def test_rgsqrf_mixed_square_4096(later, oracle):
    """A square mixed-form factorisation whose R2 R1 composition has 16 tile
    rows of 256, so it runs the XCD-balanced triangular tile order: the gates
    (FP64 metrics on the device, themselves checked against the oracle) and
    the same bits on a second run."""
    import torch
    n = 4096
    A0 = later.generate_normal(n, n, 0, 0)
    o = later.opts("mixed")
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


def test_rgsqrf_deterministic(later, oracle):
    A = oracle.gen_uniform(4096, 512, 0, 0)
    for form in ("reference", "fp32", "mixed"):
        Q1, R1, _ = _run(later, A, form)
        Q2, R2, _ = _run(later, A, form)
        assert np.array_equal(Q1, Q2) and np.array_equal(R1, R2), form


@pytest.mark.parametrize("form,m,n", [("reference", 4096, 1024), ("mixed6", 3000, 1100), ("fp32", 2048, 1024),
                                     ("mixed6", 2048, 2048)])
def test_presplit_pieces_give_fused_bits(later, oracle, form, m, n):
    """Nodes of 512+ columns stage their 16-bit operands as pre-split pieces
    in the 16-bit arena `hwork` when it is as large as the query asks; one
    element short of it (or without it), the GEMMs convert in the kernel.
    Both give the same bits.  (mixed6: the mixed form with three-piece
    second-pass projections; its int8 slices exist only staged, see
    test_int8_needs_the_arena.)"""
    A = oracle.gen_uniform(m, n, 0, 0)
    o = later.opts("mixed", gemm_pass2=later.GEMM_F16X6) if form == "mixed6" else later.opts(form)
    lw, lh = later.workspace_sizes(m, n, o)
    if form == "fp32":  # FP32 products stage nothing
        assert lh == 0
        return
    assert lh > 0
    hw = later.hworkspace(m, n, o)
    out = []
    for h in (hw, hw[:-1], None):
        Ad = _dev(later, A)
        R, info = later.rgsqrf(Ad, o=o, info=True, hwork=h)
        _sync()
        assert info.status == 0
        out.append((_np(Ad), _np(R), info.kernels))
    for q in out[1:]:
        assert np.array_equal(out[0][0], q[0]) and np.array_equal(out[0][1], q[1])
    assert out[0][2] > out[1][2] == out[2][2]  # the split launches ran with the full arena only


def test_rgsqrf_honours_lda(later, oracle):
    import torch
    m, n, ld = 1000, 200, 1037
    A = oracle.gen_uniform(m, n, 0, 0)
    Q1, R1, _ = _run(later, A, "reference")
    big = later.colmajor(m, n, ld=ld)
    big.copy_(torch.from_numpy(A))
    R = later.rgsqrf(big, o=later.opts("reference"))
    _sync()
    assert big.stride(1) == ld
    assert np.array_equal(_np(big), Q1) and np.array_equal(_np(R), R1)


def test_legacy_entry_point_matches_ex(later, oracle):
    import torch
    m, n = 2048, 256
    A = oracle.gen_uniform(m, n, 0, 0)
    Q1, R1, _ = _run(later, A, "reference")
    Ad = _dev(later, A)
    R = later.colmajor(n, n)
    R.fill_(7.0)
    work = torch.empty(m // 256 * 32 * n, device="cuda")     # test/test_qr.cu:73
    hwork = torch.empty(m * n, device="cuda", dtype=torch.int16)  # test/test_qr.cu:76
    later.rgsqrf_legacy(Ad, R, work, hwork)
    _sync()
    assert np.array_equal(_np(Ad), Q1) and np.array_equal(_np(R), R1)


def test_breakdown_sets_status(later, oracle):
    import torch
    A = oracle.gen_uniform(512, 64, 0, 0)
    A[:, 5] = 0.0
    Ad = _dev(later, A)
    o = later.opts("reference")
    work = later.workspace(512, 64, o)
    R = later.rgsqrf(Ad, o=o, work=work)
    assert later.read_status(work) & 1
    assert torch.isfinite(Ad).all() and torch.isfinite(R).all()
    ok = later.workspace(512, 64, o)
    later.rgsqrf(_dev(later, oracle.gen_uniform(512, 64, 0, 0)), o=o, work=ok)
    assert later.read_status(ok) == 0


def test_workspace_too_small_is_an_error(later):
    import torch
    A = later.generate_uniform(4096, 512)
    A0 = A.clone()
    for hw in (None, torch.empty(4096, device="cuda", dtype=torch.int16)):
        with pytest.raises(later.LaterError) as e:
            later.rgsqrf(A, o=later.opts("mixed"), work=torch.empty(10, device="cuda"), hwork=hw)
        assert e.value.code == later.ERR_WORKSPACE
    torch.cuda.synchronize()
    assert torch.equal(A, A0)  # nothing was enqueued
    # a 16-bit arena large enough for the whole FP32 layout takes it instead
    lw = later.workspace(4096, 512, later.opts("mixed")).numel()
    hw = torch.empty(2 * lw + 256, device="cuda", dtype=torch.int16)
    R = later.rgsqrf(A, o=later.opts("mixed"), work=torch.empty(10, device="cuda"), hwork=hw)
    Rf = later.rgsqrf(A0, o=later.opts("mixed"), hwork=None)
    torch.cuda.synchronize()
    assert torch.equal(R, Rf) and torch.equal(A, A0)


def test_int8_needs_the_arena(later, oracle):
    """The mixed form's int8 second-pass projections need their slices
    staged in hwork: one element short of the queried arena (or without
    it) those nodes run the three-piece FP16 split, whose bits the run then
    has, and both meet the gates."""
    import torch
    m, n = 2048, 2048
    A = oracle.gen_uniform(m, n, 0, 0)
    o8, o6 = later.opts("mixed"), later.opts("mixed", gemm_pass2=later.GEMM_F16X6)
    hw = later.hworkspace(m, n, o8)
    out = []
    for o, h in ((o8, hw), (o8, hw[:-1]), (o8, None), (o6, None)):
        Ad = _dev(later, A)
        R = later.rgsqrf(Ad, o=o, hwork=h)
        _sync()
        out.append((Ad, R))
    assert not torch.equal(out[0][1], out[1][1])
    for q in out[2:]:
        assert torch.equal(out[1][0], q[0]) and torch.equal(out[1][1], q[1])
    A0 = _dev(later, A)
    for Q, R in out[:2]:
        assert later.check_orth(Q) < 1e-5 and later.check_backward(A0, Q, R) < 1e-6


# ---------------------------------------------------------------- -check
@pytest.mark.parametrize("m,n", [(1000, 100), (4096, 256)])
def test_check_metrics_match_oracle(later, oracle, m, n):
    A = oracle.gen_uniform(m, n, 0, 0)
    Q, R = oracle.rgsqrf(A, "f16")
    Qd, Rd, Ad = _dev(later, Q), _dev(later, R), _dev(later, A)
    o_g, o_o = later.check_orth(Qd), oracle.orth(Q)
    b_g, b_o = later.check_backward(Ad, Qd, Rd), oracle.backward(A, Q, R)
    assert abs(o_g - o_o) < 1e-9 * max(1.0, o_o) + 1e-12
    assert abs(b_g - b_o) < 1e-6 * b_o + 1e-15


@pytest.mark.parametrize("m,n,kind", [(2048, 2048, "normal"), (4096, 1024, "uniform")])
def test_rgsqrf_mixed_cholqr2_panel(later, oracle, m, n, kind):
    A = oracle.gen_normal(m, n, 0, 0) if kind == "normal" else oracle.gen_uniform(m, n, 0, 0)
    Q, R, _ = _run(later, A, "mixed", panel=later.PANEL_CHOLQR2)
    Qh = np.linalg.qr(A)[0]
    assert oracle.orth(Q) < 6 * oracle.orth(Qh)
    assert oracle.backward(A, Q, R) < 1e-6


@pytest.mark.parametrize("form", ["reference", "mixed"])
def test_plan_replays_match_eager(later, oracle, form):
    """A graph-captured plan replayed several times, with eager work and
    synchronisations in between, gives the eager call's bits every time."""
    import torch
    m, n = 4096, 1024
    A = oracle.gen_uniform(m, n, 0, 0)
    Qe, Re, _ = _run(later, A, form)
    o = later.opts(form)
    Ad = _dev(later, A)
    R = later.colmajor(n, n)
    plan = later.Plan(Ad, R, o)
    assert plan.kernels > 10
    src = _dev(later, A)
    for rep in range(3):
        Ad.copy_(src)
        plan.execute()
        _sync()
        assert np.array_equal(_np(Ad), Qe) and np.array_equal(_np(R), Re), rep
        _run(later, oracle.gen_uniform(512, 64, rep, 1), form)  # eager work in between
    plan.close()


@pytest.mark.parametrize("m,w", [(4096, 32), (300, 5), (1000000, 32)])
def test_cholqr2_panel_ill_conditioned(later, oracle, m, w):
    """Cholesky QR twice on a panel with kappa = 1e6 (graded columns)."""
    rng = np.random.default_rng(0)
    U, _ = np.linalg.qr(rng.standard_normal((min(m, 20000), w)))
    if m > 20000:
        U = np.vstack([U] * (m // 20000) + [U[: m % 20000]])
        U, _ = np.linalg.qr(U)
    V, _ = np.linalg.qr(rng.standard_normal((w, w)))
    s = 10.0 ** (-6 * np.arange(w) / max(w - 1, 1))
    A = np.asfortranarray((U * s) @ V.T, dtype=np.float32)
    Ad = _dev(later, A)
    R = later.caqr_panel(Ad, fp64=2)
    _sync()
    Qg, Rg = _np(Ad), _np(R)
    assert oracle.orth(Qg) < 2e-6 * np.sqrt(w)
    assert oracle.backward(A, Qg, Rg) < 1e-6
    assert np.array_equal(Rg, np.triu(Rg))


def _graded(m, n, decades, seed=0):
    """A = U diag(s) V^T with s geometric from 1 to 10^-decades."""
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = 10.0 ** (-decades * np.arange(n) / max(n - 1, 1))
    return np.asfortranarray((U * s) @ V.T, dtype=np.float32)


@pytest.mark.parametrize("m,n", [(4096, 64), (300, 50), (100000, 64), (1000, 33), (64, 64), (500, 65),
                                 (2000, 97)])
def test_cholqr2_wide_leaf(later, oracle, m, n):
    """Cholesky QR leaves of 33..64 columns (the blocked 64 x 64 factor and
    inverse) in the mixed form, alone or beside a narrow leaf (65 = 64 + 1,
    97 = 64 + 33), kappa = 1e6: Householder-level orthogonality, backward
    error and R against LAPACK in FP64."""
    A = _graded(m, n, 6)
    Q, R, info = _run(later, A, "mixed")
    Q64, R64 = oracle.lapack_qr64(A)
    assert oracle.orth(Q) < 2e-6 * np.sqrt(n)
    assert oracle.backward(A, Q, R) < 1e-6
    assert np.array_equal(R, np.triu(R)) and np.all(np.diag(R) > 0)
    kappa = np.linalg.cond(A.astype(np.float64))
    assert oracle.rel_err(R, R64) < 2e-6 * max(1.0, kappa / 100)


@pytest.mark.parametrize("dup", [(3, 4), (40, 41), (10, 45)])
def test_cholqr2_wide_leaf_breakdown(later, oracle, dup):
    """A repeated column in the first block, in the second block, and across
    the two: the status bit is set, the dependent column of Q is zero, the
    others stay orthonormal, and A = QR still holds (R keeps the column's
    projections on the earlier ones)."""
    m, n = 2000, 64
    A = oracle.gen_uniform(m, n, 0, 0)
    A[:, dup[1]] = A[:, dup[0]]
    Ad = _dev(later, A)
    o = later.opts("mixed")
    work = later.workspace(m, n, o)
    R, info = later.rgsqrf(Ad, o=o, work=work, info=True)
    _sync()
    assert info.status == 0  # the call itself succeeded
    assert later.read_status(work) & 1
    Q, Rg = _np(Ad), _np(R)
    keep = [j for j in range(n) if j != dup[1]]
    assert np.all(Q[:, dup[1]] == 0)
    assert np.isfinite(Q).all() and np.isfinite(Rg).all()
    assert oracle.orth(Q[:, keep]) < 2e-6 * np.sqrt(n)
    assert oracle.backward(A, Q, Rg) < 1e-6


def test_cholqr2_panel_breakdown(later, oracle):
    import torch
    A = oracle.gen_uniform(2000, 16, 0, 0)
    A[:, 4] = A[:, 3]  # rank deficient
    Ad = _dev(later, A)
    work = torch.empty(later.caqr_panel_workspace(2000, 16, 2), device="cuda")
    later.caqr_panel(Ad, fp64=2, work=work)
    assert later.read_status(work) & 1
