#!/usr/bin/env python3
"""Prepares one two-site step of the L = 32, N = 32, 2S = 0 random-D2h problem at
D = 2000 with synthetic operators of the reference's shapes and applies Heff ten times
(five timed with HIP events): a short program for kernel counters of one site.

  python -m chemps2_amd.site_probe [site]      (default 8: chained terms dominate)
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    import chemps2_amd
    from chemps2_amd.problems import mx_table, random_abinitio

    L, D = 32, 2000
    site = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    irr, T, V = random_abinitio(L, 0, group=7)
    c = chemps2_amd.Context(0)
    c.set_problem(L, 7, irr, L, 0, 0, mx_table(T, V, L))
    c.set_sectors_auto(D)
    c.fill_synthetic(site - 1, 1, 1)
    c.fill_synthetic(site + 1, 0, 2)
    n = c.prepare(site)
    print("site", site, "V", n, c.plan_stats(), flush=True)
    x = np.random.default_rng(0).normal(size=n)
    for _ in range(5):
        c.apply(x)
    c.timing(True)
    for _ in range(5):
        c.apply(x)
    ms, cnt = c.timing_read()
    print("ms per application: GEMM tiles %.3f, explicit %.3f, chained kernels %.3f" % tuple(v / cnt for v in ms), flush=True)
    c.close()


if __name__ == "__main__":
    main()
