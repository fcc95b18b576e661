"""The drop-in boundary on the GPU: the reference callers' workspace sizes and
aliasing, the polar-decomposition caller's stacked shape, non-finite input,
least squares with default options, and the compiled C driver.

Oracles: the C restatement of the reference recursion and FP64 LAPACK
(oracle/), and the library's own status-form results on the same buffers.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _np(t):
    return t.detach().cpu().numpy().astype(np.float32, copy=False).copy(order="F")


def _dev(later, A, ld=None):
    import torch
    m, n = A.shape
    T = later.colmajor(m, n, ld=ld)
    T.copy_(torch.from_numpy(np.asarray(A, dtype=np.float32)))
    return T


def _sync():
    import torch
    torch.cuda.synchronize()


def _ref_sizes(m, n):
    """test/test_qr.cu:73-76: lwork = m/256*32*n floats, lhwork = m*n halves."""
    return (m // 256) * 32 * n, m * n


@pytest.mark.parametrize("m,n", [(100, 7), (200, 64), (255, 128), (300, 33), (1000, 100), (1024, 128),
                                 (2048, 256), (512, 512), (8192, 1024)])
def test_reference_workspace_sizes_run(later, m, n):
    """The status form with exactly the reference caller's sizes succeeds and
    gives the bits of a call with the full queried workspace; so does the
    legacy entry point.  The status word is work[0] when lwork >= 64, else
    hwork's first word."""
    import torch
    A = later.generate_uniform(m, n, 0, 0)
    A0 = A.clone()
    o = later.opts("reference")
    Rq = later.rgsqrf(A, o=o)
    _sync()
    Qq, Rq = _np(A), _np(Rq)
    lw, lh = _ref_sizes(m, n)
    work = torch.zeros(max(lw, 64), device="cuda")
    hwork = torch.zeros(lh, device="cuda", dtype=torch.int16)
    Ad = A0.clone()
    R = later.colmajor(n, n)
    L = later.lib()
    info = later.later_info()
    rc = L.later_rgsqrf_ex(later.ctxt(), m, n, Ad.data_ptr(), m, R.data_ptr(), n, work.data_ptr(), lw,
                           hwork.data_ptr(), lh, ctypes.byref(o), ctypes.byref(info))
    assert rc == later.OK
    st = later.read_status(work if lw >= 64 else hwork)
    assert st == 0
    assert np.array_equal(_np(Ad), Qq) and np.array_equal(_np(R), Rq)
    Ad = A0.clone()
    R.fill_(7.0)
    L.later_rgsqrf(later.ctxt(), m, n, Ad.data_ptr(), m, R.data_ptr(), n, work.data_ptr(), lw, hwork.data_ptr(), lh)
    _sync()
    assert np.array_equal(_np(Ad), Qq) and np.array_equal(_np(R), Rq)


@pytest.mark.parametrize("n", [256, 512])
def test_polar_caller_stacked_shape(later, oracle, n):
    """The QR that the reference's polar decomposition issues
    (EVD/later_qdwh_polar.cu:72-79): X = [sqrt(c) A; I] of 2n x n with lda = 2n.
    With a separate R it matches the oracle's restatement of the reference
    recursion (FP16 node GEMMs); the status form rejects R aliasing `work`
    (LATER_ERR_ALIAS), and the legacy entry, called exactly as that caller
    calls it (R = work, ldr = n, lwork = 2n/256*32*n, lhwork = 2n*n), still
    writes the same Q and R."""
    import torch
    m = 2 * n
    A = oracle.gen_uniform(n, n, 3, 0)
    A = A / np.float32(np.linalg.norm(A.astype(np.float64)))
    c = np.float32(2.7)
    X = np.asfortranarray(np.vstack([np.sqrt(c) * A, np.eye(n, dtype=np.float32)]), dtype=np.float32)
    Xd = _dev(later, X, ld=m)
    o = later.opts("reference")
    R = later.rgsqrf(Xd, o=o)
    _sync()
    Q, Rg = _np(Xd), _np(R)
    Qo, Ro = oracle.rgsqrf(X, "f16")
    Q64, R64 = oracle.lapack_qr64(X)
    if n >= 512:
        # Rows 768.. of the identity block are zero in the first 32 columns,
        # so one 256-row block of the first TSQR panel is all zero and the
        # reference's block MGS divides by a zero norm (QR/panel.cu:281-285):
        # its result is NaN.  The same QR of a row permutation of X (same R,
        # no zero block) gives the restatement's error level instead.
        assert np.isnan(Ro).any()
        perm = np.random.default_rng(n).permutation(m)
        Qo, Ro = oracle.rgsqrf(np.asfortranarray(X[perm]), "f16")
        Qo = np.asfortranarray(Qo[np.argsort(perm)])
        assert not np.isnan(Ro).any()
    e_or = oracle.rel_err(Ro, R64)
    if n < 512:
        assert oracle.rel_err(Rg, Ro) < 0.6 * e_or + 1e-6
    assert oracle.rel_err(Rg, R64) < 1.5 * e_or + 1e-6
    assert oracle.backward(X, Q, Rg) < 1.5 * oracle.backward(X, Qo, Ro) + 1e-7
    # R aliasing work: rejected by the status form
    lw, lh = 2 * n // 256 * 32 * n, 2 * n * n
    work = torch.zeros(n * n, device="cuda")  # the caller's work also holds R (n x n)
    hwork = torch.zeros(lh, device="cuda", dtype=torch.int16)
    L = later.lib()
    Xa = _dev(later, X, ld=m)
    rc = L.later_rgsqrf_ex(later.ctxt(), m, n, Xa.data_ptr(), m, work.data_ptr(), n, work.data_ptr(), lw,
                           hwork.data_ptr(), lh, ctypes.byref(o), None)
    assert rc == later.ERR_ALIAS
    _sync()
    assert np.array_equal(_np(Xa), X)  # untouched
    # ... and handled by the legacy entry point as the caller uses it
    L.later_rgsqrf(later.ctxt(), m, n, Xa.data_ptr(), m, work.data_ptr(), n, work.data_ptr(), lw,
                   hwork.data_ptr(), lh)
    _sync()
    Rw = work[: n * n].view(n, n).t().cpu().numpy()
    assert np.array_equal(_np(Xa), Q)
    assert np.array_equal(np.asfortranarray(Rw), Rg)


@pytest.mark.parametrize("form", ["reference", "fp32", "mixed"])
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_nonfinite_input_sets_status(later, oracle, form, bad):
    """A NaN or Inf entry: the call returns, enqueued work finishes, and the
    device status word has bit 0 set (Q and R are then unspecified)."""
    m, n = 2048, 512
    A = oracle.gen_uniform(m, n, 0, 0)
    A[17, 300] = bad
    Ad = _dev(later, A)
    o = later.opts(form)
    work = later.workspace(m, n, o)
    R, info = later.rgsqrf(Ad, o=o, work=work, info=True)
    _sync()
    assert info.status == later.OK
    assert later.read_status(work) & 1
    ok = later.workspace(m, n, o)
    later.rgsqrf(_dev(later, oracle.gen_uniform(m, n, 0, 0)), o=o, work=ok)
    assert later.read_status(ok) == 0


@pytest.mark.parametrize("m,n,k", [(128, 16, 256), (1000, 16, 128), (4096, 32, 64), (65536, 32, 32)])
def test_gels_default_options(later, m, n, k):
    """Least squares with the default options at shapes whose Q^T B
    partials need the full queried workspace (64-float header, column
    exponents, partials), against FP64 least squares."""
    import torch
    rng = np.random.default_rng(m + n + k)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, n)), dtype=np.float32)
    B = np.asfortranarray(rng.uniform(-1, 1, (m, k)), dtype=np.float32)
    x64 = np.linalg.lstsq(A.astype(np.float64), B.astype(np.float64), rcond=None)[0]
    lw = ctypes.c_int64()
    assert later.lib().later_ls_workspace(m, n, k, None, ctypes.byref(lw)) == later.OK
    proj = later.gemm_workspace(m, n, k, later.GEMM_F16X3, 256, False)[0]
    assert lw.value >= 64 + proj
    # a guard region past the queried workspace must stay untouched
    work = torch.full((lw.value + 4096,), 3.0, device="cuda")
    Ad, Bd = _dev(later, A), _dev(later, B)
    R = later.colmajor(n, n)
    X = later.colmajor(n, k)
    rc = later.lib().later_gels(later.ctxt(), m, n, k, Ad.data_ptr(), m, R.data_ptr(), n, Bd.data_ptr(), m,
                                X.data_ptr(), n, work.data_ptr(), lw.value, None, None)
    assert rc == later.OK
    _sync()
    assert torch.all(work[lw.value:] == 3.0)
    Xg = _np(X).astype(np.float64)
    assert np.linalg.norm(Xg - x64) / np.linalg.norm(x64) < 1e-4


def _driver(*args, timeout=120):
    exe = os.path.join(ROOT, "later_amd", "test_qr")
    p = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    line = [ln for ln in p.stdout.splitlines() if ln.startswith("result ")]
    res = dict(kv.split("=", 1) for kv in line[-1].split()[1:]) if line else {}
    return p.returncode, res, p.stdout + p.stderr


@pytest.mark.parametrize("m,n,form", [(1024, 256, None), (100, 7, None), (4096, 512, "fp32"),
                                      (4096, 1024, "mixed")])
def test_c_driver(later, oracle, m, n, form):
    """The plain-C driver test_qr (test/test_qr.cu:22-101) through
    include/later.h and the shared library: the reference-form run equals
    the oracle's restatement in its metrics, the mixed form meets the gates."""
    args = [1, m, n, "-check"] + (["-form", form] if form else [])
    rc, res, out = _driver(*args)
    assert rc == 0, out
    assert res["status"] == "0"
    orth, bwd = float(res["orth"]), float(res["backward"])
    assert abs(float(res["orth_over_n"]) - orth / n) <= 1e-6 * orth / n
    A = oracle.gen_uniform(m, n, 0, 0)
    if form is None:
        Qo, Ro = oracle.rgsqrf(A, "f16")
        assert bwd < 1.5 * oracle.backward(A, Qo, Ro) + 1e-7
        assert orth < 1.5 * oracle.orth(Qo) + 1e-6
    elif form == "fp32":
        assert bwd < 1e-6
    else:
        assert orth < 1e-5 and bwd < 1e-6


def test_c_driver_usage(later):
    rc, _, out = _driver(2, 1024, 256)
    assert rc == 2 and "only 1" in out
    rc, _, _ = _driver(1, 10, 20)
    assert rc == 2


@pytest.mark.parametrize("form", [None, "mixed"])
def test_c_driver_out_of_core(later, form):
    """The C driver's out-of-core mode (the role of test/test_oc_qr.cu:69-75):
    8192 x 2048 (64 MiB) with a 40 MiB device budget through
    later_oc_qr_rec on pinned host memory; the reference form's metrics
    match the in-core run's within 2x, the mixed form meets the gates."""
    m, n = 8192, 2048
    args = [1, m, n, "-check", "-oc", 40] + (["-form", form] if form else [])
    rc, res, out = _driver(*args)
    assert rc == 0, out
    assert res["status"] == "0" and res["oc_mib"] == "40"
    orth, bwd = float(res["orth"]), float(res["backward"])
    if form is None:
        rc2, res2, out2 = _driver(1, m, n, "-check", "-form", "reference")
        assert rc2 == 0, out2
        assert orth <= 2 * float(res2["orth"]) + 1e-6 and bwd <= 2 * float(res2["backward"]) + 1e-7
    else:
        assert orth < 1e-5 and bwd < 1e-6
