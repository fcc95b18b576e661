"""later_rgsqrf_noalloc: the QR with its workspace in the library's own device
array, through the C-ABI, against later_rgsqrf_ex on caller buffers (same
kernels, bit-identical Q and R) -- two calls in a row, the second smaller,
so the second runs on the first call's leftover contents -- and the size
limit rejected before anything is enqueued."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("reference", (2048, 1024), (700, 300)),
    ("mixed", (4096, 1024), (1500, 200)),
    ("fp32", (262144, 256), (8192, 128)),
]


def _ex(later, A0, o):
    import torch
    A = A0.clone()
    m, n = A.shape
    R = later.colmajor(n, n)
    work = later.workspace(m, n, o)
    rc = later.lib().later_rgsqrf_ex(later.ctxt(), m, n, A.data_ptr(), m, R.data_ptr(), n, work.data_ptr(),
                                     work.numel(), None, 0, ctypes.byref(o), None)
    assert rc == later.OK
    torch.cuda.synchronize()
    return A, R


def _noalloc(later, A0, o):
    import torch
    A = A0.clone()
    m, n = A.shape
    R = later.colmajor(n, n)
    R.fill_(5.0)
    info = later.later_info()
    rc = later.lib().later_rgsqrf_noalloc(later.ctxt(), m, n, A.data_ptr(), m, R.data_ptr(), n, ctypes.byref(o),
                                          ctypes.byref(info))
    assert rc == later.OK and info.status == later.OK
    torch.cuda.synchronize()
    return A, R


@pytest.mark.parametrize("form,big,small", CASES)
def test_noalloc_matches_status_form(later, form, big, small):
    import torch
    o = later.opts(form)
    for i, (m, n) in enumerate((big, small)):
        print(f"noalloc {form} {m}x{n}", flush=True)
        A0 = later.generate_normal(m, n, 11, i)
        Qe, Re = _ex(later, A0, o)
        Qn, Rn = _noalloc(later, A0, o)
        assert torch.equal(Qn, Qe) and torch.equal(Rn, Re)
        assert later.check_backward(A0, Qn, Rn) < (1e-3 if form == "reference" else 1e-6)


def test_noalloc_rejects_oversize(later):
    import torch
    m = n = 8192  # the mixed form needs far more than the array holds
    o = later.opts("mixed")
    lw = ctypes.c_int64()
    assert later.lib().later_rgsqrf_workspace(m, n, ctypes.byref(o), ctypes.byref(lw), None) == later.OK
    assert lw.value > 40 * 1024 * 1024
    A = later.colmajor(m, n)
    A.fill_(1.0)
    R = later.colmajor(n, n)
    R.fill_(2.0)
    rc = later.lib().later_rgsqrf_noalloc(later.ctxt(), m, n, A.data_ptr(), m, R.data_ptr(), n, ctypes.byref(o),
                                          None)
    assert rc == later.ERR_WORKSPACE
    torch.cuda.synchronize()
    assert bool((A == 1.0).all()) and bool((R == 2.0).all())
    rc = later.lib().later_rgsqrf_noalloc(later.ctxt(), 10, 20, A.data_ptr(), 10, R.data_ptr(), 20, None, None)
    assert rc == later.ERR_SHAPE
