"""The oracle pinned against what fixes it independently of itself.

* Philox4x32-10: the published known-answer vectors of the Random123
  distribution (Salmon et al., SC'11).
* The binary16 / bfloat16 casts (util/util.cu:24-32): numpy's IEEE float16
  cast over every 32-bit pattern, torch's bfloat16 cast on a large sample.
* The reference restatement (QR/later_rgsqrf.cu:25-60, QR/panel.cu:10-325):
  FP64 LAPACK dgeqrf/dorgqr (sign-normalised; a full-rank A has one thin QR
  with positive diagonal), brute-force FP64 Gram-Schmidt, and closed forms
  (orthonormal, column-scaled and triangular inputs).
* Planted mistakes: a restatement with a dropped update or a transposed
  operand must fail these checks (the mistake is planted in a copy of the
  algorithm on the CPU, never in the product).
"""
import numpy as np
import pytest


def test_philox_kat(oracle):
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, out in kat:
        assert tuple(oracle.philox4x32_10(ctr, key)) == out


def test_uniform_generator_range_and_bits(oracle):
    A = oracle.gen_uniform(1000, 7, seed=0, mat=3)
    assert A.min() >= -1.0 and A.max() < 1.0
    # Every value is 2u-1 with u a multiple of 2^-24.
    u = (A.astype(np.float64) + 1.0) / 2.0 * 2.0 ** 24
    assert np.all(u == np.round(u))
    # Element (i, j) uses counter i + j*m: a column subset regenerates the same bits.
    B = oracle.gen_uniform(1000, 7, seed=0, mat=3)
    assert np.array_equal(A, B)
    C = oracle.gen_uniform(1000, 7, seed=1, mat=3)
    assert not np.array_equal(A, C)
    ctr = 5 + 2 * 1000
    x = oracle.philox4x32_10((ctr, 0, 0, 0), (0, 3))[0]
    assert A[5, 2] == np.float32(2.0 * (x >> 8) * 2.0 ** -24 - 1.0)


def test_normal_generator_moments(oracle):
    A = oracle.gen_normal(200000, 1, seed=0, mat=0).ravel().astype(np.float64)
    assert abs(A.mean()) < 0.01 and abs(A.std() - 1.0) < 0.01
    ctr = 17
    x = oracle.philox4x32_10((ctr, 0, 0, 0), (0, 0))
    u1 = ((x[0] >> 8) + 1) * 2.0 ** -24
    u2 = (x[1] >> 8) * 2.0 ** -24
    assert A[17] == np.float32(np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2))


def test_f16_cast_exhaustive(oracle):
    """Every FP32 bit pattern against torch's CPU float16 cast (an independent
    IEEE RN-even implementation), and a strided sample against numpy's."""
    import torch
    chunk = 1 << 24
    base = torch.arange(chunk, dtype=torch.int64)
    bits = torch.empty(chunk, dtype=torch.int64)
    ref = torch.empty(chunk, dtype=torch.float16)
    lib = oracle.lib()
    bad = 0
    for start in range(0, 1 << 32, chunk):
        torch.add(base, start, out=bits)
        ref.copy_(bits.to(torch.int32).view(torch.float32))
        bad += lib.oracle_f16_mismatches(start, chunk, ref.data_ptr())
    assert bad == 0
    x = np.arange(0, 1 << 32, 4099, dtype=np.uint64).astype(np.uint32).view(np.float32)
    with np.errstate(all="ignore"):
        r = x.astype(np.float16).view(np.uint16)
    nan = np.isnan(x)
    assert np.array_equal(oracle.f32_to_f16_bits(x)[~nan], r[~nan])


def test_bf16_cast_sample(oracle):
    import torch
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 1 << 32, size=1 << 20, dtype=np.uint64).astype(np.uint32)
    bits = np.concatenate([bits, np.array([0x3f808000, 0x3f818000, 0x7f7fffff, 0x00008000, 0x80000001],
                                          dtype=np.uint32)])
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    ref = torch.from_numpy(x.copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    lib = oracle.lib()
    mine = np.array([lib.oracle_f32_to_bf16(float(v)) for v in x[:20000]], dtype=np.uint16)
    assert np.array_equal(mine, ref[:20000])


def _gs64(A):
    """Brute-force FP64 modified Gram-Schmidt."""
    A = np.array(A, dtype=np.float64)
    m, n = A.shape
    R = np.zeros((n, n))
    for k in range(n):
        R[k, k] = np.linalg.norm(A[:, k])
        A[:, k] /= R[k, k]
        R[k, k + 1:] = A[:, k] @ A[:, k + 1:]
        A[:, k + 1:] -= np.outer(A[:, k], R[k, k + 1:])
    return A, R


@pytest.mark.parametrize("m,w", [(256, 32), (4096, 32), (1000, 32), (287, 32), (300, 7), (64, 1)])
def test_caqr32_vs_gs64(oracle, m, w):
    A = oracle.gen_uniform(m, w, 0, 0)
    Q, R = oracle.caqr32(A)
    Qe, Re = _gs64(A)
    assert oracle.rel_err(R, Re) < 2e-6
    assert oracle.rel_err(Q, Qe) < 2e-6
    assert np.all(np.diag(R) > 0)
    assert np.array_equal(R, np.triu(R))


def test_mgs_block_closed_forms(oracle):
    Q0, _ = oracle.lapack_qr64(oracle.gen_uniform(256, 32, 0, 0))
    Q0 = np.asfortranarray(Q0, dtype=np.float32)
    # Orthonormal input: Q = A, R = I.
    Q, R = oracle.mgs_block(Q0)
    assert np.abs(R - np.eye(32)).max() < 1e-6
    assert np.abs(Q - Q0).max() < 1e-6
    # Column scaling: R = diag(d).
    d = np.linspace(0.5, 4.0, 32).astype(np.float32)
    Q, R = oracle.mgs_block(np.asfortranarray(Q0 * d))
    assert np.abs(R - np.diag(d)).max() < 4e-6 * d.max()
    # Triangular input embedded in the top rows: Q = [I; 0], R = T.
    T = np.triu(oracle.gen_uniform(32, 32, 0, 1)) + 4 * np.eye(32, dtype=np.float32)
    A = np.zeros((256, 32), dtype=np.float32, order="F")
    A[:32] = T
    Q, R = oracle.mgs_block(A)
    assert np.abs(R - T).max() < 1e-5
    assert np.abs(Q[:32] - np.eye(32)).max() < 1e-6 and np.abs(Q[32:]).max() == 0


@pytest.mark.parametrize("m,n", [(1024, 128), (2048, 256), (1000, 100), (300, 33), (512, 512)])
def test_rgsqrf_fp32_vs_lapack(oracle, m, n):
    A = oracle.gen_uniform(m, n, 0, 0)
    Q, R = oracle.rgsqrf(A, "f32")
    Q64, R64 = oracle.lapack_qr64(A)
    kappa = np.linalg.cond(A.astype(np.float64))
    assert oracle.rel_err(R, R64) < 1e-6 * max(1.0, kappa / 100)
    assert oracle.rel_err(Q, Q64) < 1e-6 * max(1.0, kappa / 10)
    assert oracle.backward(A, Q, R) < 1e-6
    assert np.array_equal(R, np.triu(R))


def test_rgsqrf_f16_reference_form(oracle):
    """FP16 staging (QR/later_rgsqrf.cu:43-56) limits accuracy to ~u16 = 4.9e-4."""
    A = oracle.gen_uniform(2048, 256, 0, 0)
    Q, R = oracle.rgsqrf(A, "f16")
    Q32, R32 = oracle.rgsqrf(A, "f32")
    Q64, R64 = oracle.lapack_qr64(A)
    e16, e32 = oracle.rel_err(R, R64), oracle.rel_err(R32, R64)
    assert 1e-5 < e16 < 2e-3 and e32 < 1e-6
    assert oracle.backward(A, Q, R) < 5e-4
    # Only the top node (n=256 > nmin=128) uses FP16: the left 128 columns
    # of Q and the R11 block match the FP32 form bit for bit.
    assert np.array_equal(Q[:, :128], Q32[:, :128]) and np.array_equal(R[:128, :128], R32[:128, :128])
    # FP16 operands give the same R12 as an FP64 product of fp16-rounded Q1, A2.
    Q1h = Q[:, :128].astype(np.float16).astype(np.float64)
    # A2 before the update is the original right half.
    A2h = A[:, 128:].astype(np.float16).astype(np.float64)
    R12 = (Q1h.T @ A2h).astype(np.float32)
    assert np.array_equal(R12, R[:128, 128:])


def test_metrics_closed_forms(oracle):
    Q0, R0 = oracle.lapack_qr64(oracle.gen_uniform(300, 20, 0, 0))
    Q = np.asfortranarray(Q0, dtype=np.float32)
    assert oracle.orth(Q) < 1e-6
    # A column scaled by (1+e) gives ||I - Q^T Q||_F = 2e + e^2.
    Qs = Q.copy()
    Qs[:, 3] *= np.float32(1.001)
    e = float(np.float32(1.001)) - 1.0
    assert abs(oracle.orth(Qs) - (2 * e + e * e)) < 1e-6
    A = np.asfortranarray((Q0 @ R0).astype(np.float32))
    assert oracle.backward(A, Q, np.asfortranarray(R0, dtype=np.float32)) < 1e-6
    # The strictly-lower part of R is ignored (the reference's R keeps garbage there).
    Rg = np.asfortranarray(R0, dtype=np.float32)
    Rg[5, 2] = 1e3
    assert oracle.backward(A, Q, Rg) < 1e-6


def test_planted_mistakes_are_caught(oracle):
    """Copies of the recursion with a dropped update or a transposed operand
    fail the checks that pin the restatement."""
    A = oracle.gen_uniform(512, 64, 0, 0).astype(np.float64)

    def rec(A, bug):
        m, n = A.shape
        if n <= 32:
            return _gs64(A)
        n1 = n // 2
        Q1, R11 = rec(A[:, :n1], bug)
        A2 = A[:, n1:].copy()
        R12 = (A2.T @ Q1).T if bug != "transpose" else (A2.T @ Q1)
        if bug != "drop":
            A2 -= Q1 @ R12
        Q2, R22 = rec(A2, bug)
        R = np.block([[R11, R12], [np.zeros((n - n1, n1)), R22]])
        return np.hstack([Q1, Q2]), R

    Q64, R64 = oracle.lapack_qr64(A)
    Q, R = rec(A, None)
    assert oracle.rel_err(R, R64) < 1e-12
    for bug in ("drop", "transpose"):
        Qb, Rb = rec(A, bug)
        assert oracle.rel_err(Rb, R64) > 1e-3, bug


# ---------------------------------------------------------------- FP8 quantisation
def _e4m3_codes():
    vals = []
    for c in range(256):
        s, e, m = c >> 7, (c >> 3) & 15, c & 7
        if e == 15 and m == 7:
            continue  # NaN
        v = (m / 8.0) * 2.0 ** -6 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7)
        vals.append(-v if s else v)
    return np.unique(np.array(vals))


def test_e4m3_round_matches_torch_and_brute_force(oracle):
    import torch
    rng = np.random.default_rng(3)
    y = (rng.standard_normal(400000) * np.exp2(rng.integers(-14, 9, 400000))).clip(-448, 448).astype(np.float32)
    codes = _e4m3_codes()
    # exact ties between neighbouring codes (round half to even)
    mids = ((codes[:-1] + codes[1:]) / 2).astype(np.float32)
    y = np.concatenate([y, mids, codes.astype(np.float32)])
    mine = oracle.e4m3_round(y.astype(np.float64))
    t = torch.from_numpy(y).to(torch.float8_e4m3fn).float().numpy().astype(np.float64)
    assert np.array_equal(mine, t)
    # brute force: the nearest code, ties to the code with an even last bit
    idx = np.searchsorted(codes, y.astype(np.float64))
    lo = codes[np.clip(idx - 1, 0, len(codes) - 1)]
    hi = codes[np.clip(idx, 0, len(codes) - 1)]
    d_lo, d_hi = np.abs(y - lo), np.abs(hi - y)
    near = np.where(d_lo < d_hi, lo, hi)
    tie = d_lo == d_hi
    assert np.array_equal(mine[~tie], near[~tie])
    assert np.all(np.isin(mine[tie], codes))


def test_fp8_block_exponents(oracle):
    rng = np.random.default_rng(4)
    X = rng.standard_normal((1000, 7)) * np.exp2(rng.integers(-40, 40, (1, 7)))
    X[:, 3] = 0.0
    X[128:256, 5] = 448.0  # exactly the top: e = 0
    e = oracle.fp8_block_exponents(X)
    assert e.shape == (8, 7)
    for b in range(8):
        mx = np.abs(X[128 * b:128 * (b + 1)]).max(axis=0)
        s = mx * np.exp2(e[b].astype(float))
        nz = mx > 0
        assert np.all((s[nz] > 224) & (s[nz] <= 448))
    assert np.all(e[:, 3] == 0) and e[1, 5] == 0
    Xq = oracle.fp8_quantize(X)
    # every value moves by at most half an e4m3 spacing of its block's top binade
    for b in range(8):
        blk = slice(128 * b, 128 * (b + 1))
        mx = np.abs(X[blk]).max(axis=0)
        assert np.all(np.abs(Xq[blk] - X[blk]) <= mx * 2.0 ** -4 + 1e-300)
    # the scaled values are e4m3 codes
    codes = _e4m3_codes()
    sc = np.repeat(np.exp2(e.astype(float)), 128, axis=0)[:1000]
    assert np.all(np.isin(Xq * sc, codes))
    # planted mistakes: a scale 2^12 too small (values fall into the
    # subnormals) changes the result; one binade too large for blocks whose
    # top lies in (448, 512) 2^-e leaves the code set
    Y = rng.standard_normal((128, 50))
    ey = oracle.fp8_block_exponents(Y)[0]
    small = oracle.e4m3_round(Y * np.exp2(ey - 12.0)) / np.exp2(ey - 12.0)
    assert not np.array_equal(small, oracle.fp8_quantize(Y))
    mx = np.abs(Y).max(axis=0)
    f, fe = np.frexp(mx)
    big = f > 0.875
    assert big.any()
    top = mx[big] * np.exp2(9.0 - fe[big])
    assert not np.any(np.isin(oracle.e4m3_round(top), codes))
