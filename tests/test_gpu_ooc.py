"""Out-of-core QR (later_oc_qr_rec; QR/later_oc_qr.cu:29-94) on host
matrices with a device budget below the matrix size, so that the recursion
runs out-of-core nodes over streamed row blocks.

The reference's own out-of-core update does not form A2 - Q1 R12 (its
OC_gemm never reads C from the host), so parity is pinned against the
in-core factorisation of the same input and options and against FP64
LAPACK: the same FP64 metrics on the device, R against LAPACK's R, and
reproducibility (bit-identical results for repeated calls and for pinned
and pageable host memory).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _host(m, n, pinned, ld=None):
    import torch
    ld = m if ld is None else ld
    buf = torch.empty((n, ld), dtype=torch.float32, pin_memory=pinned)
    return buf.t()[:m, :]


def _metrics(later, A0_dev, Q_host, R_host):
    import torch
    m, n = Q_host.shape
    Q = later.colmajor(m, n)
    Q.copy_(Q_host)
    R = later.colmajor(n, n)
    R.copy_(R_host)
    torch.cuda.synchronize()
    return later.check_orth(Q), later.check_backward(A0_dev, Q, R)


def _r_err(R, A0):
    """max_j ||R(:,j) - R64(:,j)|| / ||A0(:,j)|| with LAPACK's signs matched."""
    R64 = np.linalg.qr(A0.astype(np.float64), mode="r")
    s = np.sign(np.diag(R64)) * np.sign(np.diag(R))
    return float(np.max(np.linalg.norm(R - s[:, None] * R64, axis=0) / np.linalg.norm(A0, axis=0)))


def _case(later, m, n, form, budget, pinned=True, ld=None, seed=3):
    import torch
    A0 = later.generate_uniform(m, n, seed, 0)
    torch.cuda.synchronize()
    Ah = _host(m, n, pinned, ld)
    Ah.copy_(A0.cpu())
    Rh = _host(n, n, pinned)
    Rh.fill_(7.0)
    info = later.oc_qr_rec(Ah, Rh, budget, later.opts(form), info=True)
    return A0, Ah, Rh, info


@pytest.mark.parametrize("form", ["fp32", "reference", "mixed"])
def test_oc_matches_in_core(later, form):
    """8192 x 1024 with half the matrix's bytes as budget (out-of-core nodes
    at the top two levels): the FP64 metrics within 2x (+ a small absolute
    term) of the in-core call's, the mixed form at its gates, R against
    LAPACK as close as the in-core R, R's strictly-lower triangle zero."""
    m, n = 8192, 1024
    print(f"oc {m}x{n} {form}", flush=True)
    A0, Ah, Rh, info = _case(later, m, n, form, m * n * 4 // 2)
    assert info.status == 0
    orth, bwd = _metrics(later, A0, Ah, Rh)
    A = A0.clone()
    R = later.rgsqrf(A, o=later.opts(form))
    import torch
    torch.cuda.synchronize()
    orth_ic, bwd_ic = later.check_orth(A), later.check_backward(A0, A, R)
    assert orth <= 2 * orth_ic + 1e-6 and bwd <= 2 * bwd_ic + 1e-7, (orth, orth_ic, bwd, bwd_ic)
    if form == "mixed":
        assert orth <= 1e-5 and bwd <= 1e-6, (orth, bwd)
    Rn, A0n = Rh.numpy(), A0.cpu().numpy()
    assert np.all(np.tril(Rn, -1) == 0)
    e_oc, e_ic = _r_err(Rn, A0n), _r_err(R.cpu().numpy(), A0n)
    assert e_oc <= 2 * e_ic + 1e-6, (e_oc, e_ic)


def test_oc_reproducible_pinned_and_pageable(later):
    """Ragged shape (5000 x 1000, lda 5003): repeated calls give the same
    bits, pinned and pageable host memory give the same bits."""
    m, n = 5000, 1000
    budget = 12 << 20  # the matrix is 19 MiB; the smallest budget that fits is 9.6 MiB
    _, A1, R1, _ = _case(later, m, n, "mixed", budget, pinned=True, ld=m + 3)
    _, A2, R2, _ = _case(later, m, n, "mixed", budget, pinned=True, ld=m + 3)
    A0, A3, R3, _ = _case(later, m, n, "mixed", budget, pinned=False, ld=m + 3)
    for X, Y in ((A1, A2), (R1, R2), (A1, A3), (R1, R3)):
        assert np.array_equal(X.numpy(), Y.numpy())
    orth, bwd = _metrics(later, A0, A1, R1)
    assert orth <= 1e-5 and bwd <= 1e-6, (orth, bwd)


def test_oc_in_core_when_it_fits(later):
    """A budget that holds the whole factorisation runs it in core: the
    same bits as later_rgsqrf_ex with the same options and operand arena."""
    import torch
    m, n = 4096, 512
    A0, Ah, Rh, info = _case(later, m, n, "reference", 1 << 30)
    A = A0.clone()
    R = later.rgsqrf(A, o=later.opts("reference"))
    torch.cuda.synchronize()
    assert np.array_equal(Ah.numpy(), A.cpu().numpy())
    assert np.array_equal(Rh.numpy(), R.cpu().numpy())


def test_oc_budget_too_small_touches_nothing(later):
    """A budget that cannot hold a 32-column leaf is rejected before any
    copy: A and R keep their contents."""
    m, n = 65536, 256
    import torch
    A0 = later.generate_uniform(m, n, 1, 0)
    Ah = _host(m, n, True)
    Ah.copy_(A0.cpu())
    Rh = _host(n, n, True)
    Rh.fill_(7.0)
    before = Ah.numpy().copy()
    with pytest.raises(RuntimeError):
        later.oc_qr_rec(Ah, Rh, m * 32 * 4 // 2, later.opts("fp32"))
    assert np.array_equal(Ah.numpy(), before)
    assert np.all(Rh.numpy() == 7.0)
    torch.cuda.synchronize()
