"""Autotuner of the one-launch batched QR (later_batch_wg_*).

For every size bucket of the tuning table (later_batch_wg_bucket: n class
min(ceil(n/32), 8) - 1, m class clamp(floor(log2 m) - 6, 0, 6)) it factors a
synthetic batch of matrices drawn from that bucket (m log-uniform and n
uniform inside the bucket's ranges, entries 2u - 1 from the Philox
generator) with each variant of the kernel forced, and keeps the fastest:
variant v runs panels of 16 << (v & 1) columns, projected against blocks of
64 << (v >> 1) earlier columns of Q.  The batch of each bucket holds enough matrices for
about 2e10 flops (64 to 2048 of them), timed as the minimum of three
executions after one warm-up, on the mixed-form options.

Writes the table as the C header later_amd/csrc/batch_tuning.h (compiled
into the library as the default table) and as later_amd/batch_tuning.json
(with the timings).  Run on the GPU:  python -m later_amd.autotune
"""
from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
N_CLASSES, M_CLASSES = 8, 7
VARIANTS = (0, 1, 2, 3)


def bucket_ranges(nc: int, mc: int):
    """(m_lo, m_hi, n_lo, n_hi), inclusive bounds of a bucket's shapes."""
    m_lo = 32 if mc == 0 else 2 ** (mc + 6)
    m_hi = 2 ** (mc + 7) - 1
    n_lo, n_hi = 32 * nc + 1, 32 * (nc + 1)
    return m_lo, m_hi, n_lo, min(n_hi, m_hi)


def tune(budget_flops: float = 2e10, log=print):
    import numpy as np
    import torch
    import later_amd as L

    o = L.opts("mixed")
    table = [-1] * (N_CLASSES * M_CLASSES)
    timings = {}
    for nc in range(N_CLASSES):
        for mc in range(M_CLASSES):
            b = M_CLASSES * nc + mc
            m_lo, m_hi, n_lo, n_hi = bucket_ranges(nc, mc)
            if n_lo > n_hi:
                continue  # no shape with n <= m in this bucket
            rng = np.random.default_rng(1000 + b)
            mean = 0.5 * (m_lo + m_hi) * ((n_lo + n_hi) / 2) ** 2 * 4
            cnt = int(min(2048, max(64, budget_flops / mean)))
            m = np.rint(np.exp(rng.uniform(np.log(max(m_lo, n_lo)), np.log(m_hi), cnt))).astype(int)
            n = np.minimum(rng.integers(n_lo, n_hi + 1, cnt), m)
            assert all(L.batch_bucket(int(a), int(c)) == b for a, c in zip(m, n)), b
            As = [L.generate_uniform(int(mi), int(ni), seed=b, mat=i) for i, (mi, ni) in enumerate(zip(m, n))]
            A0 = [a.clone() for a in As]
            Rs = [L.colmajor(int(ni), int(ni)) for ni in n]
            s = torch.cuda.current_stream()
            best, best_ms, row = None, None, {}
            for v in VARIANTS:
                plan = L.BatchWg(As, Rs, o, tuning=[v] * L.BATCH_BUCKETS)
                ts = []
                for it in range(4):
                    for a, a0 in zip(As, A0):
                        a.copy_(a0)
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    torch.cuda.synchronize()
                    e0.record(s)
                    plan.execute()
                    e1.record(s)
                    torch.cuda.synchronize()
                    if it:
                        ts.append(e0.elapsed_time(e1))
                plan.close()
                ms = min(ts)
                row[v] = ms
                if best_ms is None or ms < best_ms:
                    best, best_ms = v, ms
            table[b] = best
            timings[b] = {"m": [m_lo, m_hi], "n": [n_lo, n_hi], "count": cnt, "ms": row, "best": best}
            log(f"[autotune] bucket {b:2d} m {m_lo}-{m_hi} n {n_lo}-{n_hi} ({cnt} matrices): "
                + "  ".join(f"v{v} {row[v]:.3f} ms" for v in VARIANTS) + f"  -> v{best}")
            del As, A0, Rs
            torch.cuda.empty_cache()
    return table, timings


def write(table, timings, device: str):
    # Buckets with no shape keep the widest variant.
    full = [3 if v < 0 else v for v in table]
    lines = ["// Variant of the one-workgroup batched kernel per size bucket (see",
             "// bucket_of): 0 = 16-column panels / 64-column blocks of earlier Q columns, 1 = 32 / 64,",
             "// 2 = 16 / 128, 3 = 32 / 128.  Written by later_amd/autotune.py from",
             f"// timings on {device}; rows are n classes (n <= 32, <= 64, ..., > 224),",
             "// columns m classes (m < 128, < 256, ..., >= 4096).",
             "#pragma once",
             "constexpr int kDefaultTuning[kBuckets] = {"]
    for nc in range(N_CLASSES):
        lines.append("    " + ", ".join(str(full[M_CLASSES * nc + mc]) for mc in range(M_CLASSES)) + ",")
    lines.append("};")
    with open(os.path.join(HERE, "csrc", "batch_tuning.h"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(HERE, "batch_tuning.json"), "w") as f:
        json.dump({"device": device, "table": full, "buckets": timings}, f, indent=1, sort_keys=True)


def main():
    import torch
    t0 = time.time()
    table, timings = tune(log=lambda *a: print(*a, flush=True))
    write(table, timings, torch.cuda.get_device_name(0) + " (gfx950)")
    print(f"[autotune] done in {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(HERE))
    main()
