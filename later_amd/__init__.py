"""MI355X-native recursive Gram-Schmidt QR (LATER's ``later_rgsqrf``).

Thin Python face of the C-ABI in ``include/later.h``; the compute path is the
hand-written HIP library ``liblater_amd.so`` built from ``later_amd/csrc``.
Device memory and streams come from PyTorch-ROCm tensors (plumbing only): every
matrix is FP32 column-major, i.e. a tensor ``A`` of shape (m, n) with
``A.stride(0) == 1`` and leading dimension ``A.stride(1)``; ``colmajor(m, n)``
allocates one.

There is no fallback: if the HIP library is missing or fails to load, every
entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblater_amd.so")

# Status codes (include/later.h)
OK, ERR_SHAPE, ERR_WORKSPACE, ERR_OPTION, ERR_DEVICE, ERR_BREAKDOWN, ERR_ALIAS, ERR_BUDGET, ERR_CANCELLED = range(9)
GEMM_F32, GEMM_F16, GEMM_BF16, GEMM_F16X3, GEMM_F16X6, GEMM_FP8, GEMM_I8X4 = range(7)
PANEL_MGS, PANEL_MGS64, PANEL_CHOLQR2, PANEL_CHOLQR1 = range(4)
REORTH_NONE, REORTH_NODE, REORTH_PASS = range(3)

_STATUS_NAMES = {
    ERR_SHAPE: "bad shape", ERR_WORKSPACE: "workspace too small", ERR_OPTION: "bad option",
    ERR_DEVICE: "HIP runtime error", ERR_BREAKDOWN: "breakdown (zero or non-finite column norm)",
    ERR_ALIAS: "aliased buffers", ERR_BUDGET: "exceeds HBM budget", ERR_CANCELLED: "cancelled",
}


class LaterError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        super().__init__(f"{what}: status {code} ({_STATUS_NAMES.get(code, 'unknown')})")
        self.code = code


class laterCtxt(ctypes.Structure):
    _fields_ = [("stream", ctypes.c_void_p), ("handle", ctypes.c_void_p)]


class later_opts(ctypes.Structure):
    _fields_ = [("gemm", ctypes.c_int), ("gemm_pass2", ctypes.c_int), ("panel", ctypes.c_int),
                ("reorth", ctypes.c_int), ("nmin", ctypes.c_int), ("chunk", ctypes.c_int),
                ("chunk2", ctypes.c_int), ("gemm_update", ctypes.c_int), ("nmin2", ctypes.c_int)]

    def __repr__(self):
        return ("later_opts(gemm=%d, gemm_pass2=%d, panel=%d, reorth=%d, nmin=%d, chunk=%d, chunk2=%d, "
                "gemm_update=%d, nmin2=%d)" % (self.gemm, self.gemm_pass2, self.panel, self.reorth, self.nmin,
                                               self.chunk, self.chunk2, self.gemm_update, self.nmin2))


class later_info(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("kernels", ctypes.c_int), ("tree_levels", ctypes.c_int),
                ("flops_ref", ctypes.c_double), ("flops_mgs", ctypes.c_double)]


_lib = None
_vp, _i, _i64, _u32, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint32, ctypes.c_double

_SIGNATURES = {
    "later_opts_default": (None, [ctypes.POINTER(later_opts)]),
    "later_opts_mixed": (None, [ctypes.POINTER(later_opts)]),
    "later_opts_fp32": (None, [ctypes.POINTER(later_opts)]),
    "later_rgsqrf_workspace": (_i, [_i, _i, ctypes.POINTER(later_opts), ctypes.POINTER(_i64),
                                    ctypes.POINTER(_i64)]),
    "later_rgsqrf": (None, [laterCtxt, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i]),
    "later_rgsqrf_ex": (_i, [laterCtxt, _i, _i, _vp, _i, _vp, _i, _vp, _i64, _vp, _i64,
                             ctypes.POINTER(later_opts), ctypes.POINTER(later_info)]),
    "later_rgsqrf_noalloc": (_i, [laterCtxt, _i, _i, _vp, _i, _vp, _i, ctypes.POINTER(later_opts),
                                  ctypes.POINTER(later_info)]),
    "later_oc_qr_rec": (_i, [laterCtxt, _i, _i, _vp, _i, _vp, _i, _i64, ctypes.POINTER(later_opts),
                             ctypes.POINTER(later_info)]),
    "later_read_status": (_i, [laterCtxt, _vp, ctypes.POINTER(_i)]),
    "later_caqr_panel_workspace": (_i, [_i, _i, _i, ctypes.POINTER(_i64)]),
    "later_caqr_panel": (_i, [laterCtxt, _i, _i, _vp, _i, _vp, _i, _vp, _i64, _i]),
    "later_gemm_workspace": (_i, [_i, _i, _i, _i, _i, _i, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "later_gemm_proj": (_i, [laterCtxt, _i, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _i, _vp, _i64, _vp, _i64]),
    "later_gemm_update": (_i, [laterCtxt, _i, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp, _i64, _vp, _i64]),
    "later_generate_uniform": (_i, [laterCtxt, _i64, _i64, _vp, _i64, _u32, _u32]),
    "later_generate_normal": (_i, [laterCtxt, _i64, _i64, _vp, _i64, _u32, _u32]),
    "later_check_orth": (_i, [laterCtxt, _i, _i, _vp, _i, ctypes.POINTER(_d)]),
    "later_check_backward": (_i, [laterCtxt, _i, _i, _vp, _i, _vp, _i, _vp, _i, ctypes.POINTER(_d)]),
    "later_version": (ctypes.c_char_p, []),
    "later_trsm_upper": (_i, [laterCtxt, _i, _i, _vp, _i, _vp, _i]),
    "later_ls_workspace": (_i, [_i, _i, _i, ctypes.POINTER(later_opts), ctypes.POINTER(_i64)]),
    "later_gels": (_i, [laterCtxt, _i, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i64,
                        ctypes.POINTER(later_opts), ctypes.POINTER(later_info)]),
    "later_server_create": (_i, [_i, _i64, ctypes.POINTER(later_opts), ctypes.POINTER(_vp)]),
    "later_server_submit": (_i, [_vp, _i, _i, _i, _vp, _i, _vp, _i, ctypes.POINTER(ctypes.c_uint64)]),
    "later_server_poll": (_i, [_vp, ctypes.POINTER(_i)]),
    "later_server_result": (_i, [_vp, ctypes.c_uint64, ctypes.POINTER(_i), ctypes.POINTER(_i),
                                 ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                 ctypes.POINTER(_d), ctypes.POINTER(_d), ctypes.POINTER(_d), ctypes.POINTER(_i)]),
    "later_server_release": (_i, [_vp, ctypes.c_uint64]),
    "later_server_cancel": (_i, [_vp, ctypes.c_uint64]),
    "later_server_fetch": (_i, [_vp, ctypes.c_uint64, _vp, _i, _vp, _i, _vp, _i]),
    "later_server_stats": (_i, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "later_server_destroy": (_i, [_vp]),
    "later_plan_create": (_i, [_i, _i, _vp, _i, _vp, _i, _vp, _i64, _vp, _i64, ctypes.POINTER(later_opts),
                               ctypes.POINTER(_vp)]),
    "later_plan_execute": (_i, [_vp, _vp]),
    "later_plan_kernels": (_i, [_vp]),
    "later_plan_destroy": (_i, [_vp]),
    "later_batch_workspace": (_i, [_i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(later_opts),
                                   ctypes.POINTER(_i64)]),
    "later_batch_plan_create": (_i, [_i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_vp),
                                     ctypes.POINTER(_i), ctypes.POINTER(_vp), ctypes.POINTER(_i), _vp, _i64,
                                     ctypes.POINTER(later_opts), ctypes.POINTER(_vp)]),
    "later_batch_plan_execute": (_i, [_vp, _vp]),
    "later_batch_plan_launches": (_i, [_vp]),
    "later_batch_plan_destroy": (_i, [_vp]),
    "later_batch_wg_create": (_i, [_i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_vp),
                                   ctypes.POINTER(_i), ctypes.POINTER(_vp), ctypes.POINTER(_i), _vp,
                                   ctypes.POINTER(later_opts), ctypes.POINTER(_i), ctypes.POINTER(_vp)]),
    "later_batch_wg_execute": (_i, [_vp, _vp]),
    "later_batch_wg_destroy": (_i, [_vp]),
    "later_batch_wg_bucket": (_i, [_i, _i]),
    "later_prof_begin": (None, []),
    "later_prof_end": (_i, [ctypes.POINTER(_i), ctypes.POINTER(_d), ctypes.POINTER(_d), ctypes.POINTER(_d)]),
}


def lib() -> ctypes.CDLL:
    """The loaded HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python later_amd/build.py`")
        # torch ships its own libamdhip64.so.7; loading it first makes the
        # library bind to that same HIP runtime (one runtime per process).
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# Options
# --------------------------------------------------------------------------
def opts(form: str = "reference", **over) -> later_opts:
    """'reference' (FP16 node GEMMs, FP32 MGS panel, one pass), 'fp32'
    (FP32 MFMA everywhere), 'mixed' (FP16 hi/lo products, FP64 panel sums,
    second pass), or 'fp8' / 'bf16': the mixed form whose first-pass
    projections run e4m3 (per-128-row block scales) or BF16 operands, with
    hi/lo updates and nodes above 32 columns (for well-conditioned tall
    panels; the second pass is the refinement).  Keyword arguments override
    single fields."""
    o = later_opts()
    base = {"fp8": "mixed", "bf16": "mixed"}.get(form, form)
    fn = {"reference": "later_opts_default", "fp32": "later_opts_fp32", "mixed": "later_opts_mixed"}[base]
    getattr(lib(), fn)(ctypes.byref(o))
    if form in ("fp8", "bf16"):
        o.gemm = GEMM_FP8 if form == "fp8" else GEMM_BF16
        o.gemm_update = GEMM_F16X3
        o.nmin = 32
    for k, v in over.items():
        setattr(o, k, int(v))
    return o


# --------------------------------------------------------------------------
# Tensors (plumbing)
# --------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def colmajor(m: int, n: int, device="cuda", dtype=None, ld: int | None = None):
    """Uninitialised column-major (m x n) FP32 tensor with leading dimension ld."""
    torch = _torch()
    ld = m if ld is None else ld
    buf = torch.empty((n, ld), device=device, dtype=dtype or torch.float32)
    return buf.t()[:m, :]


def _ld(t) -> int:
    if t.dim() != 2 or t.stride(0) != 1:
        raise ValueError("expected a column-major 2-D tensor (stride(0) == 1)")
    return max(t.stride(1), t.shape[0], 1)


def ctxt(stream=None) -> laterCtxt:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return laterCtxt(ctypes.c_void_p(s.cuda_stream), None)


def _check(rc: int, what: str):
    if rc != OK:
        raise LaterError(rc, what)


# --------------------------------------------------------------------------
# QR
# --------------------------------------------------------------------------
def workspace_sizes(m: int, n: int, o: later_opts | None = None) -> tuple[int, int]:
    """(lwork, lhwork): FP32 elements of `work` (required) and 16-bit
    elements of `hwork` (optional operand staging)."""
    lw, lh = _i64(), _i64()
    _check(lib().later_rgsqrf_workspace(m, n, ctypes.byref(o) if o is not None else None,
                                        ctypes.byref(lw), ctypes.byref(lh)), "later_rgsqrf_workspace")
    return int(lw.value), int(lh.value)


def workspace_size(m: int, n: int, o: later_opts | None = None) -> int:
    return workspace_sizes(m, n, o)[0]


def workspace(m: int, n: int, o: later_opts | None = None, device="cuda"):
    torch = _torch()
    return torch.empty(max(workspace_size(m, n, o), 1), device=device, dtype=torch.float32)


def hworkspace(m: int, n: int, o: later_opts | None = None, device="cuda"):
    """The 16-bit staging arena of the query (None when nothing is staged)."""
    torch = _torch()
    lh = workspace_sizes(m, n, o)[1]
    return torch.empty(lh, device=device, dtype=torch.int16) if lh > 0 else None


def _hw(hwork):
    return (hwork.data_ptr(), int(hwork.numel())) if hwork is not None else (None, 0)


def rgsqrf(A, R=None, o: later_opts | None = None, work=None, stream=None, info: bool = False, hwork="auto"):
    """QR of the column-major FP32 tensor A in place: A <- Q.  Returns R (and
    the later_info when info=True).  Asynchronous on the current stream.
    hwork: the 16-bit staging arena ("auto": allocated by the query; None:
    no staging, the GEMMs convert in the kernel with the same bits)."""
    m, n = A.shape
    if R is None:
        R = colmajor(n, n, device=A.device)
    if work is None:
        work = workspace(m, n, o, device=A.device)
    if isinstance(hwork, str):
        hwork = hworkspace(m, n, o, device=A.device)
    hp, hn = _hw(hwork)
    inf = later_info()
    rc = lib().later_rgsqrf_ex(ctxt(stream), m, n, A.data_ptr(), _ld(A), R.data_ptr(), _ld(R),
                               work.data_ptr(), work.numel(), hp, hn,
                               ctypes.byref(o) if o is not None else None, ctypes.byref(inf))
    _check(rc, "later_rgsqrf_ex")
    return (R, inf) if info else R


def oc_qr_rec(A, R, device_bytes: int, o: later_opts | None = None, stream=None, info: bool = False):
    """Out-of-core QR of the HOST column-major FP32 tensor A in place (A <- Q,
    R <- R, both CPU tensors; pinned memory lets copies overlap products)
    using at most device_bytes of device memory.  Synchronous."""
    if A.device.type != "cpu" or R.device.type != "cpu":
        raise ValueError("oc_qr_rec takes host tensors")
    m, n = A.shape
    inf = later_info()
    rc = lib().later_oc_qr_rec(ctxt(stream), m, n, A.data_ptr(), _ld(A), R.data_ptr(), _ld(R), int(device_bytes),
                               ctypes.byref(o) if o is not None else None, ctypes.byref(inf))
    _check(rc, "later_oc_qr_rec")
    return inf if info else None


class Plan:
    """A graph-captured factorisation of A (in place) into R for fixed buffers."""

    def __init__(self, A, R, o: later_opts | None = None, work=None, hwork="auto"):
        m, n = A.shape
        self.A, self.R = A, R
        self.work = work if work is not None else workspace(m, n, o, device=A.device)
        self.hwork = hworkspace(m, n, o, device=A.device) if isinstance(hwork, str) else hwork
        hp, hn = _hw(self.hwork)
        self.o = o
        h = _vp()
        _check(lib().later_plan_create(m, n, A.data_ptr(), _ld(A), R.data_ptr(), _ld(R), self.work.data_ptr(),
                                       self.work.numel(), hp, hn, ctypes.byref(o) if o is not None else None,
                                       ctypes.byref(h)), "later_plan_create")
        self.h = h

    @property
    def kernels(self) -> int:
        return int(lib().later_plan_kernels(self.h))

    def execute(self, stream=None):
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        _check(lib().later_plan_execute(self.h, ctypes.c_void_p(s.cuda_stream)), "later_plan_execute")

    def close(self):
        if self.h:
            lib().later_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_BUCKETS = 56


class BatchWg:
    """Batched QR of the column-major tensors As (in place) into Rs in one
    persistent kernel launch over a device work queue (later_batch_wg_*).
    tuning: None (built-in table) or BATCH_BUCKETS variant ids."""

    def __init__(self, As, Rs, o: later_opts | None = None, tuning=None, status: bool = True):
        torch = _torch()
        cnt = len(As)
        I = _i * cnt
        P = _vp * cnt
        self.m = I(*[a.shape[0] for a in As])
        self.n = I(*[a.shape[1] for a in As])
        self.lda = I(*[_ld(a) for a in As])
        self.ldr = I(*[_ld(r) for r in Rs])
        self.A = P(*[a.data_ptr() for a in As])
        self.R = P(*[r.data_ptr() for r in Rs])
        self.status = torch.zeros(cnt, device=As[0].device, dtype=torch.int32) if status else None
        self.keep = (As, Rs)
        tun = (_i * BATCH_BUCKETS)(*[int(v) for v in tuning]) if tuning is not None else None
        h = _vp()
        _check(lib().later_batch_wg_create(cnt, self.m, self.n, self.A, self.lda, self.R, self.ldr,
                                           self.status.data_ptr() if status else None,
                                           ctypes.byref(o) if o is not None else None, tun, ctypes.byref(h)),
               "later_batch_wg_create")
        self.h = h

    def execute(self, stream=None):
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        _check(lib().later_batch_wg_execute(self.h, ctypes.c_void_p(s.cuda_stream)), "later_batch_wg_execute")

    def close(self):
        if self.h:
            lib().later_batch_wg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_bucket(m: int, n: int) -> int:
    return int(lib().later_batch_wg_bucket(m, n))


class BatchPlan:
    """Graph-captured batched QR of the column-major tensors As (in place) into Rs."""

    def __init__(self, As, Rs, o: later_opts | None = None):
        torch = _torch()
        cnt = len(As)
        I = _i * cnt
        P = _vp * cnt
        self.m = I(*[a.shape[0] for a in As])
        self.n = I(*[a.shape[1] for a in As])
        self.lda = I(*[_ld(a) for a in As])
        self.ldr = I(*[_ld(r) for r in Rs])
        self.A = P(*[a.data_ptr() for a in As])
        self.R = P(*[r.data_ptr() for r in Rs])
        lw = _i64()
        _check(lib().later_batch_workspace(cnt, self.m, self.n, ctypes.byref(o) if o is not None else None,
                                           ctypes.byref(lw)), "later_batch_workspace")
        self.work = torch.empty(max(lw.value, 1), device=As[0].device, dtype=torch.float32)
        self.keep = (As, Rs)
        h = _vp()
        _check(lib().later_batch_plan_create(cnt, self.m, self.n, self.A, self.lda, self.R, self.ldr,
                                             self.work.data_ptr(), self.work.numel(),
                                             ctypes.byref(o) if o is not None else None, ctypes.byref(h)),
               "later_batch_plan_create")
        self.h = h

    @property
    def launches(self) -> int:
        return int(lib().later_batch_plan_launches(self.h))

    def execute(self, stream=None):
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        _check(lib().later_batch_plan_execute(self.h, ctypes.c_void_p(s.cuda_stream)), "later_batch_plan_execute")

    def close(self):
        if self.h:
            lib().later_batch_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rgsqrf_legacy(A, R, work, hwork, stream=None):
    """The reference-signature entry point (include/LATER.h:39)."""
    m, n = A.shape
    lib().later_rgsqrf(ctxt(stream), m, n, A.data_ptr(), _ld(A), R.data_ptr(), _ld(R), work.data_ptr(),
                       int(work.numel()), hwork.data_ptr() if hwork is not None else None,
                       int(hwork.numel()) if hwork is not None else 0)


def read_status(work, stream=None) -> int:
    st = _i()
    _check(lib().later_read_status(ctxt(stream), work.data_ptr(), ctypes.byref(st)), "later_read_status")
    return int(st.value)


def caqr_panel_workspace(m: int, w: int, fp64=False) -> int:
    """FP32 elements of `work` that caqr_panel needs."""
    lw = _i64()
    _check(lib().later_caqr_panel_workspace(m, w, int(fp64), ctypes.byref(lw)), "later_caqr_panel_workspace")
    return int(lw.value)


def caqr_panel(A, R=None, fp64=False, work=None, stream=None):
    """mgs_caqr_panel_256x32 (QR/panel.cu:65-134) for an m x w panel, w <= 32.
    fp64: False/0 FP32 block MGS, True/1 FP64 sums, 2 Cholesky QR twice."""
    torch = _torch()
    m, w = A.shape
    if R is None:
        R = colmajor(w, w, device=A.device)
    if work is None:
        work = torch.empty(caqr_panel_workspace(m, w, fp64), device=A.device, dtype=torch.float32)
    _check(lib().later_caqr_panel(ctxt(stream), m, w, A.data_ptr(), _ld(A), R.data_ptr(), _ld(R),
                                  work.data_ptr(), work.numel(), int(fp64)), "later_caqr_panel")
    return R


def gemm_workspace(m: int, n1: int, n2: int, prec: int, chunk: int = 0, update: bool = False) -> tuple[int, int]:
    """(lwork, lhwork) of one component GEMM: FP32 work and 16-bit pieces."""
    lw, lh = _i64(), _i64()
    _check(lib().later_gemm_workspace(m, n1, n2, prec, chunk, int(update), ctypes.byref(lw), ctypes.byref(lh)),
           "later_gemm_workspace")
    return int(lw.value), int(lh.value)


def _gemm_bufs(device, lw, lh, pieces):
    torch = _torch()
    work = torch.empty(max(lw, 1), device=device, dtype=torch.float32)
    hwork = torch.empty(lh, device=device, dtype=torch.int16) if (pieces and lh > 0) else None
    return work, hwork


def gemm_proj(Q, B, C, prec: int, chunk: int = 0, work=None, stream=None, pieces: bool = False):
    """C = Q^T B.  pieces=True stages the operands as pre-split 16-bit pieces
    (the streaming kernels of the recursion's wide nodes)."""
    m, n1 = Q.shape
    n2 = B.shape[1]
    lw, lh = gemm_workspace(m, n1, n2, prec, chunk, False)
    w, hwork = _gemm_bufs(Q.device, lw, lh, pieces)
    work = w if work is None else work
    hp, hn = _hw(hwork)
    _check(lib().later_gemm_proj(ctxt(stream), m, n1, n2, Q.data_ptr(), _ld(Q), B.data_ptr(), _ld(B),
                                 C.data_ptr(), _ld(C), prec, chunk, work.data_ptr(), work.numel(), hp, hn),
           "later_gemm_proj")
    return C


def gemm_update(Q, C, B, prec: int, work=None, stream=None, pieces: bool = False):
    """B -= Q C (pieces as for gemm_proj)."""
    m, n1 = Q.shape
    n2 = C.shape[1]
    lw, lh = gemm_workspace(m, n1, n2, prec, 0, True)
    w, hwork = _gemm_bufs(Q.device, lw, lh, pieces)
    work = w if work is None else work
    hp, hn = _hw(hwork)
    _check(lib().later_gemm_update(ctxt(stream), m, n1, n2, Q.data_ptr(), _ld(Q), C.data_ptr(), _ld(C),
                                   B.data_ptr(), _ld(B), prec, work.data_ptr(), work.numel(), hp, hn),
           "later_gemm_update")
    return B


# --------------------------------------------------------------------------
# Inputs and -check metrics
# --------------------------------------------------------------------------
def generate_uniform(m: int, n: int, seed: int = 0, mat: int = 0, device="cuda", out=None, stream=None):
    A = colmajor(m, n, device=device) if out is None else out
    _check(lib().later_generate_uniform(ctxt(stream), m, n, A.data_ptr(), _ld(A), seed, mat),
           "later_generate_uniform")
    return A


def generate_normal(m: int, n: int, seed: int = 0, mat: int = 0, device="cuda", out=None, stream=None):
    A = colmajor(m, n, device=device) if out is None else out
    _check(lib().later_generate_normal(ctxt(stream), m, n, A.data_ptr(), _ld(A), seed, mat),
           "later_generate_normal")
    return A


def check_orth(Q, stream=None) -> float:
    """||I - Q^T Q||_F with FP64 accumulation (test/test_qr.cu:245-268 without the /n)."""
    m, n = Q.shape
    out = _d()
    _check(lib().later_check_orth(ctxt(stream), m, n, Q.data_ptr(), _ld(Q), ctypes.byref(out)),
           "later_check_orth")
    return float(out.value)


def check_backward(A, Q, R, stream=None) -> float:
    """||A - Q triu(R)||_F / ||A||_F with FP64 accumulation (test/test_qr.cu:216-228)."""
    m, n = A.shape
    out = _d()
    _check(lib().later_check_backward(ctxt(stream), m, n, A.data_ptr(), _ld(A), Q.data_ptr(), _ld(Q),
                                      R.data_ptr(), _ld(R), ctypes.byref(out)), "later_check_backward")
    return float(out.value)


# --------------------------------------------------------------------------
# Least squares and serving
# --------------------------------------------------------------------------
def trsm_upper(R, B, stream=None):
    """B[:n] <- R^{-1} B[:n] for upper-triangular R (n x n)."""
    n = R.shape[0]
    k = B.shape[1]
    _check(lib().later_trsm_upper(ctxt(stream), n, k, R.data_ptr(), _ld(R), B.data_ptr(), _ld(B)),
           "later_trsm_upper")
    return B


def gels(A, B, o: later_opts | None = None, stream=None):
    """min ||A X - B||: A <- Q in place; returns (X, R)."""
    torch = _torch()
    m, n = A.shape
    k = B.shape[1]
    lw = _i64()
    _check(lib().later_ls_workspace(m, n, k, ctypes.byref(o) if o is not None else None, ctypes.byref(lw)),
           "later_ls_workspace")
    work = torch.empty(max(lw.value, 1), device=A.device, dtype=torch.float32)
    R = colmajor(n, n, device=A.device)
    X = colmajor(n, k, device=A.device)
    _check(lib().later_gels(ctxt(stream), m, n, k, A.data_ptr(), _ld(A), R.data_ptr(), _ld(R), B.data_ptr(),
                            _ld(B), X.data_ptr(), _ld(X), work.data_ptr(), work.numel(),
                            ctypes.byref(o) if o is not None else None, None), "later_gels")
    return X, R


class Server:
    """Continuous-batching least-squares server (later_server_*)."""

    def __init__(self, slots: int = 8, hbm_budget_bytes: int = 8 << 30, o: later_opts | None = None):
        h = _vp()
        _check(lib().later_server_create(slots, hbm_budget_bytes, ctypes.byref(o) if o is not None else None,
                                         ctypes.byref(h)), "later_server_create")
        self.h = h

    def submit(self, A, B) -> int:
        m, n = A.shape
        k = B.shape[1]
        rid = ctypes.c_uint64()
        rc = lib().later_server_submit(self.h, m, n, k, A.data_ptr(), _ld(A), B.data_ptr(), _ld(B),
                                       ctypes.byref(rid))
        if rc not in (OK, ERR_BUDGET):
            _check(rc, "later_server_submit")
        return int(rid.value)

    def poll(self) -> int:
        a = _i()
        _check(lib().later_server_poll(self.h, ctypes.byref(a)), "later_server_poll")
        return int(a.value)

    def result(self, rid: int) -> dict:
        st, status, it = _i(), _i(), _i()
        X, Q, R = _vp(), _vp(), _vp()
        lat, qs, cond = _d(), _d(), _d()
        _check(lib().later_server_result(self.h, rid, ctypes.byref(st), ctypes.byref(status), ctypes.byref(X),
                                         ctypes.byref(Q), ctypes.byref(R), ctypes.byref(lat), ctypes.byref(qs),
                                         ctypes.byref(cond), ctypes.byref(it)), "later_server_result")
        return {"state": st.value, "status": status.value, "X": X.value, "Q": Q.value, "R": R.value,
                "latency_s": lat.value, "queue_s": qs.value, "cond": cond.value, "iters": it.value}

    def fetch(self, rid: int, m: int, n: int, k: int, q: bool = False):
        """Copies of (X, R[, Q]) of a finished request into new tensors."""
        X = colmajor(n, k)
        R = colmajor(n, n)
        Q = colmajor(m, n) if q else None
        _check(lib().later_server_fetch(self.h, rid, X.data_ptr(), _ld(X), R.data_ptr(), _ld(R),
                                        Q.data_ptr() if q else None, _ld(Q) if q else 0), "later_server_fetch")
        return (X, R, Q) if q else (X, R)

    def release(self, rid: int):
        _check(lib().later_server_release(self.h, rid), "later_server_release")

    def cancel(self, rid: int):
        _check(lib().later_server_cancel(self.h, rid), "later_server_cancel")

    def stats(self) -> dict:
        pb, ub, pk = _i64(), _i64(), _i64()
        c, r = _i(), _i()
        _check(lib().later_server_stats(self.h, ctypes.byref(pb), ctypes.byref(ub), ctypes.byref(pk),
                                        ctypes.byref(c), ctypes.byref(r)), "later_server_stats")
        return {"pool_bytes": pb.value, "used_bytes": ub.value, "peak_bytes": pk.value, "completed": c.value,
                "rejected": r.value}

    def close(self):
        if self.h:
            _check(lib().later_server_destroy(self.h), "later_server_destroy")
            self.h = None


PROF_CLASSES = ("gemm_proj", "gemm_update", "panel_mgs", "other", "gemm_update_x3", "proj_i8")


def prof_begin():
    lib().later_prof_begin()


def prof_end() -> dict:
    """Per kernel class: launches, ms, algorithmic flops and bytes."""
    n = len(PROF_CLASSES)
    la, ms, fl, by = (_i * n)(), (_d * n)(), (_d * n)(), (_d * n)()
    _check(lib().later_prof_end(la, ms, fl, by), "later_prof_end")
    return {c: {"launches": la[k], "ms": ms[k], "flops": fl[k], "bytes": by[k]} for k, c in enumerate(PROF_CLASSES)}


def version() -> str:
    return lib().later_version().decode()


def flops_ref(m: int, n: int) -> float:
    """The reference's GFLOPS numerator 2 n^2 (m - n/3) (test/test_qr.cu:85-86)."""
    return 2.0 * n * n * (m - n / 3.0)


def flops_mgs(m: int, n: int) -> float:
    return 2.0 * m * n * n
