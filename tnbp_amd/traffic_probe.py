"""Reference streaming kernel for calibrating HBM traffic counters.

Runs the library's elementwise scale kernel (k_scale: one read and one write
of every complex<double> element) over a tensor of known size, so that a
`rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` run gives the counter value
of exactly `16 n` bytes read and `16 n` bytes written per dispatch.  The ratio
of the counter to those bytes is the factor applied to the same counters read
for the hot kernels.

    python -m tnbp_amd.traffic_probe [--mib 512] [--repeat 5]
"""
from __future__ import annotations

import argparse
import json
import time

from ._lib import check
from .tci import Context, Tensor


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=512)
    ap.add_argument("--repeat", type=int, default=5)
    a = ap.parse_args()
    n = a.mib * (1 << 20) // 16
    ctx = Context(0)
    t = Tensor(ctx, (n,))
    check(ctx.lib.hbp_fill(ctx.h, t.h, 1.0, 0.0))
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.repeat):
        check(ctx.lib.hbp_scale(ctx.h, t.h, 1.0, 0.0))
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / a.repeat
    print(json.dumps({"kernel": "k_scale (complex<double>)", "elements": n,
                      "bytes_read_per_dispatch": 16 * n, "bytes_written_per_dispatch": 16 * n,
                      "seconds_per_dispatch_host": dt, "GBps_host": 32 * n / dt / 1e9}))
    del t
    ctx.close()


if __name__ == "__main__":
    main()
