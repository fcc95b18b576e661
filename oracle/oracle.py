"""The oracle -- TEST INFRASTRUCTURE ONLY.

Python face of the CPU checker for the recursive Gram-Schmidt QR path:

* ``liboracle.so`` (``rgsqrf_oracle.c``): a plain-C restatement of the
  reference recursion ``QR/later_rgsqrf.cu:25-60``, its 128-column panel
  ``QR/panel.cu:10-63``, the TSQR tree ``QR/panel.cu:65-134`` and the block MGS
  kernel ``QR/panel.cu:246-325``; the Philox input generator; the RN-even casts
  of ``util/util.cu:24-32``; the ``-check`` metrics of
  ``test/test_qr.cu:216-268`` in FP64.
* FP64 LAPACK (``dgeqrf`` + ``dorgqr`` through numpy) with the R factor
  sign-normalised to a positive diagonal: a full-rank A has exactly one such
  thin QR, so this pins the whole factorisation.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg import this module, and only to check or to time the CPU form.  The product
library (``later_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

PREC = {"f32": 0, "f16": 1, "bf16": 2}


def build() -> None:
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        f32p = ctypes.POINTER(ctypes.c_float)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.oracle_philox4x32_10.restype = None
        for name in ("oracle_gen_uniform", "oracle_gen_normal"):
            fn = getattr(L, name)
            fn.argtypes = [ctypes.c_int64, ctypes.c_int64, f32p, ctypes.c_int64,
                           ctypes.c_uint32, ctypes.c_uint32]
            fn.restype = None
        L.oracle_f32_to_f16.argtypes = [ctypes.c_float]
        L.oracle_f32_to_f16.restype = ctypes.c_uint16
        L.oracle_f32_to_f16_array.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        L.oracle_f32_to_f16_array.restype = None
        L.oracle_f16_mismatches.argtypes = [ctypes.c_uint32, ctypes.c_int64, ctypes.c_void_p]
        L.oracle_f16_mismatches.restype = ctypes.c_int64
        L.oracle_f32_to_bf16.argtypes = [ctypes.c_float]
        L.oracle_f32_to_bf16.restype = ctypes.c_uint16
        L.oracle_rgsqrf.argtypes = [ctypes.c_int, ctypes.c_int, f32p, ctypes.c_int, f32p,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_rgsqrf.restype = ctypes.c_int
        L.oracle_caqr32.argtypes = [ctypes.c_int, ctypes.c_int, f32p, ctypes.c_int, f32p,
                                    ctypes.c_int]
        L.oracle_caqr32.restype = ctypes.c_int
        L.oracle_mgs_block.argtypes = L.oracle_caqr32.argtypes
        L.oracle_mgs_block.restype = ctypes.c_int
        L.oracle_orth_fro.argtypes = [ctypes.c_int, ctypes.c_int, f32p, ctypes.c_int]
        L.oracle_orth_fro.restype = ctypes.c_double
        L.oracle_backward.argtypes = [ctypes.c_int, ctypes.c_int, f32p, ctypes.c_int, f32p,
                                      ctypes.c_int, f32p, ctypes.c_int]
        L.oracle_backward.restype = ctypes.c_double
        _lib = L
    return _lib


def _fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


# --------------------------------------------------------------------------
# Generator
# --------------------------------------------------------------------------
def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def gen_uniform(m: int, n: int, seed: int = 0, mat: int = 0) -> np.ndarray:
    """Entries 2u-1, u from 24 Philox bits keyed by (seed, mat)."""
    A = np.empty((m, n), dtype=np.float32, order="F")
    lib().oracle_gen_uniform(m, n, _fp(A), m, seed, mat)
    return A


def gen_normal(m: int, n: int, seed: int = 0, mat: int = 0) -> np.ndarray:
    """N(0,1) entries by FP64 Box-Muller over two Philox words."""
    A = np.empty((m, n), dtype=np.float32, order="F")
    lib().oracle_gen_normal(m, n, _fp(A), m, seed, mat)
    return A


def f32_to_f16_bits(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty(x.shape, dtype=np.uint16)
    lib().oracle_f32_to_f16_array(x.ctypes.data, y.ctypes.data, x.size)
    return y


# --------------------------------------------------------------------------
# Restatement of the reference path
# --------------------------------------------------------------------------
def rgsqrf(A: np.ndarray, prec: str = "f16", nmin: int = 128):
    """Reference-form recursive GS QR. Returns (Q, R) as new FP32 arrays."""
    m, n = A.shape
    Q = np.array(A, dtype=np.float32, order="F", copy=True)
    R = np.zeros((n, n), dtype=np.float32, order="F")
    rc = lib().oracle_rgsqrf(m, n, _fp(Q), m, _fp(R), n, PREC[prec], nmin)
    if rc != 0:
        raise ValueError(f"oracle_rgsqrf rejected shape {m}x{n}")
    return Q, R


def caqr32(A: np.ndarray):
    m, w = A.shape
    Q = np.array(A, dtype=np.float32, order="F", copy=True)
    R = np.zeros((w, w), dtype=np.float32, order="F")
    if lib().oracle_caqr32(m, w, _fp(Q), m, _fp(R), w) != 0:
        raise ValueError("bad panel shape")
    return Q, R


def mgs_block(A: np.ndarray):
    m, w = A.shape
    Q = np.array(A, dtype=np.float32, order="F", copy=True)
    R = np.zeros((w, w), dtype=np.float32, order="F")
    if lib().oracle_mgs_block(m, w, _fp(Q), m, _fp(R), w) != 0:
        raise ValueError("bad block shape")
    return Q, R


# --------------------------------------------------------------------------
# Metrics (FP64 accumulation)
# --------------------------------------------------------------------------
def orth(Q: np.ndarray) -> float:
    """||I - Q^T Q||_F (raw; the reference prints it divided by n)."""
    Q = np.asfortranarray(Q, dtype=np.float32)
    m, n = Q.shape
    return float(lib().oracle_orth_fro(m, n, _fp(Q), m))


def backward(A: np.ndarray, Q: np.ndarray, R: np.ndarray) -> float:
    """||A - Q triu(R)||_F / ||A||_F."""
    A = np.asfortranarray(A, dtype=np.float32)
    Q = np.asfortranarray(Q, dtype=np.float32)
    R = np.asfortranarray(R, dtype=np.float32)
    m, n = A.shape
    return float(lib().oracle_backward(m, n, _fp(A), m, _fp(Q), m, _fp(R), n))


# --------------------------------------------------------------------------
# FP64 LAPACK oracle
# --------------------------------------------------------------------------
def lapack_qr64(A: np.ndarray):
    """Thin QR by FP64 dgeqrf+dorgqr, sign-normalised so diag(R) > 0."""
    Q, R = np.linalg.qr(np.asarray(A, dtype=np.float64), mode="reduced")
    d = np.sign(np.diag(R))
    d[d == 0] = 1.0
    return Q * d[None, :], R * d[:, None]


def rel_err(X, Y) -> float:
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    den = np.linalg.norm(Y)
    return float(np.linalg.norm(X - Y) / (den if den > 0 else 1.0))


# ------------------------------------------------------------------ FP8
# The e4m3 operand quantisation of the FP8 projections (no counterpart in the
# reference, whose only cast is the FP16 one of util/util.cu:24-32): every
# column is cut into blocks of 128 rows; a block with largest magnitude mx is
# scaled by 2^e with e the largest integer keeping mx 2^e <= 448 (e4m3's
# largest finite value; 448 = 0.875 * 2^9), rounded to the nearest e4m3
# value (ties to even), and carries the E8M0 scale 127 - e.

E4M3_MAX = 448.0


def e4m3_round(y: np.ndarray) -> np.ndarray:
    """Round |y| <= 448 to the nearest OCP e4m3 value, ties to even: 3
    fraction bits, exponents -6 .. 8, subnormal spacing 2^-9."""
    y = np.asarray(y, dtype=np.float64)
    a = np.abs(y)
    with np.errstate(divide="ignore"):
        ex = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    ex = np.maximum(ex, -6.0)
    ulp = np.exp2(ex - 3.0)
    return np.round(y / ulp) * ulp  # np.round: half to even


def fp8_block_exponents(X: np.ndarray) -> np.ndarray:
    """e per (128-row block, column): mx 2^e in (224, 448] (0 for a zero block)."""
    X = np.asarray(X, dtype=np.float64)
    K, n = X.shape
    nb = (K + 127) // 128
    e = np.zeros((nb, n), dtype=np.int64)
    for b in range(nb):
        mx = np.abs(X[128 * b:128 * (b + 1)]).max(axis=0)
        f, fe = np.frexp(mx)  # mx = f 2^fe, f in [0.5, 1)
        eb = np.where(f <= 0.875, 9 - fe, 8 - fe)
        e[b] = np.where(mx > 0, np.clip(eb, -126, 126), 0)
    return e


def fp8_quantize(X: np.ndarray) -> np.ndarray:
    """The FP64 values the FP8 product multiplies: e4m3(x 2^e) 2^-e."""
    X = np.asarray(X, dtype=np.float64)
    e = fp8_block_exponents(X)
    sc = np.repeat(np.exp2(e.astype(np.float64)), 128, axis=0)[: X.shape[0]]
    return e4m3_round(X * sc) / sc
