"""The C-ABI library loads and exports what include/later.h declares; host-side
argument checking and the workspace query run without a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "later.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(later_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol(later):
    names = _declared()
    assert "later_rgsqrf" in names and "later_rgsqrf_ex" in names
    L = ctypes.CDLL(later.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version(later):
    assert "gfx950" in later.version()


def test_options(later):
    o = later.opts("reference")
    assert (o.gemm, o.panel, o.reorth, o.nmin) == (later.GEMM_F16, later.PANEL_MGS, later.REORTH_NONE, 128)
    o = later.opts("fp32")
    assert o.gemm == later.GEMM_F32
    o = later.opts("mixed")
    assert o.gemm == later.GEMM_F16X3 and o.reorth == later.REORTH_PASS
    assert o.gemm_pass2 == later.GEMM_I8X4 and o.chunk == 0 and o.chunk2 == 32
    assert o.nmin == 256 and o.nmin2 == 128
    assert later.opts("reference").nmin2 == 0


@pytest.mark.parametrize("m,n", [(16384, 16384), (8192, 1024), (1048576, 128), (262144, 256), (100, 7)])
def test_workspace_fits_reference_sizes(later, m, n):
    """Default options fit the reference's caller-side sizes
    (test/test_qr.cu:73-76: lwork = m/256*32*n floats, lhwork = m*n halves)."""
    need = later.workspace_size(m, n)
    ref_work = max(m // 256 * 32 * n, 0)
    ref_h = m * n
    assert need <= ref_work + ref_h // 2


def test_workspace_query_rejects_bad_shape(later):
    with pytest.raises(later.LaterError) as e:
        later.workspace_size(10, 20)
    assert e.value.code == later.ERR_SHAPE


def test_ex_rejects_without_touching_device(later):
    L = later.lib()
    c = later.laterCtxt(None, None)
    info = later.later_info()
    # n > m
    assert L.later_rgsqrf_ex(c, 10, 20, None, 10, None, 20, None, 0, None, 0, None, ctypes.byref(info)) \
        == later.ERR_SHAPE
    # lda < m
    assert L.later_rgsqrf_ex(c, 100, 20, 1 << 20, 50, 2 << 20, 20, 3 << 20, 1 << 20, None, 0, None, None) \
        == later.ERR_SHAPE
    # bad option
    o = later.opts("reference", gemm=9)
    assert L.later_rgsqrf_ex(c, 100, 20, 1 << 20, 100, 2 << 20, 20, 3 << 20, 1 << 20, None, 0,
                             ctypes.byref(o), None) == later.ERR_OPTION
    # workspace too small
    assert L.later_rgsqrf_ex(c, 100, 20, 1 << 20, 100, 2 << 20, 20, 3 << 20, 1, None, 0, None, None) \
        == later.ERR_WORKSPACE
    # R overlapping work
    need = later.workspace_size(100, 20)
    assert L.later_rgsqrf_ex(c, 100, 20, 1 << 24, 100, 2 << 20, 20, (2 << 20) + 64, need, None, 0, None,
                             None) == later.ERR_ALIAS


@pytest.mark.parametrize("form", ["fp32", "reference", "mixed"])
def test_oc_budget_checked_before_any_copy(later, form):
    """The out-of-core driver's dry run rejects a budget below what a 32-column
    leaf or a 256-row node block needs (LATER_ERR_WORKSPACE), before touching
    the device or the host matrices; bad shapes and options come first."""
    import numpy as np
    L = later.lib()
    c = later.laterCtxt(None, None)
    m, n = 65536, 256
    A = np.full((n, m), 3.0, np.float32)
    R = np.full((n, n), 7.0, np.float32)
    o = later.opts(form)
    info = later.later_info()
    rc = L.later_oc_qr_rec(c, m, n, A.ctypes.data, m, R.ctypes.data, n, 4 << 20, ctypes.byref(o),
                           ctypes.byref(info))
    assert rc == later.ERR_WORKSPACE and info.status == rc
    assert np.all(A == 3.0) and np.all(R == 7.0)
    assert L.later_oc_qr_rec(c, 10, 20, A.ctypes.data, 10, R.ctypes.data, 20, 1 << 30, None, None) \
        == later.ERR_SHAPE
    for kv in ({"chunk": 17}, {"gemm": 9}, {"gemm_update": 4}, {"panel": 5}):
        bad = later.opts(form, **kv)
        assert L.later_oc_qr_rec(c, m, n, A.ctypes.data, m, R.ctypes.data, n, 1 << 30, ctypes.byref(bad),
                                 None) == later.ERR_OPTION
