"""GPU DMRG against two more of the reference's own known-answer tests, both of which
hand the solver a raw matrix-element table (Problem::setMxElement, Problem.cpp:357-361),
which is exactly what chm_set_problem takes:

* tests/test12.cpp.in:39-99: the reduced BCS Hamiltonian on 8 levels (pairing g = -1),
  E = -25.5134137600604;
* tests/test9.cpp.in:38-127: the 3 x 3 periodic Hubbard model (U = 5, N = 9, 2S = 1) in
  the site basis and in the momentum basis (a table with only four-fold symmetry): the
  two energies agree to 1e-8.
"""
import numpy as np
import pytest

import chemps2_amd
from chemps2_amd.problems import mx_table

pytestmark = pytest.mark.gpu


def _solve(L, N, TwoS, mx, scheme, seed=7):
    c = chemps2_amd.Context(0)
    try:
        c.set_problem(L, 0, np.zeros(L, dtype=np.int32), N, TwoS, 0, mx, 0.0)
        c.dmrg_init(scheme[0][0], seed)
        return c.dmrg_solve(scheme)
    finally:
        c.close()


def test_reference_test12_bcs():
    L = N = 8
    eps = np.array([-3.5, -2.5, -1.5, -0.5, 0.5, 1.5, 2.5, 3.5])
    g, power = -1.0, 0.0
    mx = np.zeros((L, L, L, L))  # [a, b, c, d] -> element a + L(b + L(c + L d))
    for o1 in range(L):
        for o2 in range(L):
            eri = g * abs(eps[o1] * eps[o2]) ** power
            oei = (eps[o1] + eps[o2]) / (N - 1)
            if o1 == o2:
                mx[o1, o1, o2, o2] = eri + oei
            else:
                mx[o1, o1, o2, o2] = eri
                mx[o1, o2, o1, o2] = oei
    flat = np.ascontiguousarray(mx.transpose(3, 2, 1, 0)).ravel()
    print("case test12 BCS L=8", flush=True)
    # tests/test12.cpp.in:48-50: D = 100 (noise 0.5), then D = 1000; Davidson tolerance 1e-5
    e = _solve(L, N, 0, flat, [(100, 1e-10, 10, 0.5, 1e-5), (1000, 1e-10, 10, 0.0, 1e-5)])
    print(f"test12: gpu {e!r}", flush=True)
    assert abs(e + 25.5134137600604) < 1e-8


def test_reference_test9_site_vs_momentum():
    n, U, t = 3, 5.0, -1.0
    L, N, TwoS = n * n, 9, 1
    T = np.zeros((L, L))
    V = np.zeros((L, L, L, L))
    for i in range(L):
        V[i, i, i, i] = U
    for ix in range(n):
        for iy in range(n):
            a = ix + n * iy
            for b in (((ix + 1) % n) + n * iy, ix + n * ((iy + 1) % n)):
                T[a, b] = T[b, a] = t  # Hamiltonian::setTmat fills both triangles
    scheme = [(500, 1e-10, 3, 0.05, 1e-5), (1000, 1e-10, 10, 0.0, 1e-5)]
    print("case test9 3x3 Hubbard, site basis", flush=True)
    e_site = _solve(L, N, TwoS, mx_table(T, V, N), scheme)
    mom = np.zeros((L, L, L, L))
    for o1 in range(L):
        k1x, k1y = o1 % n, o1 // n
        t1 = 2 * t * (np.cos(2 * np.pi * k1x / n) + np.cos(2 * np.pi * k1y / n))
        for o2 in range(L):
            k2x, k2y = o2 % n, o2 // n
            t2 = 2 * t * (np.cos(2 * np.pi * k2x / n) + np.cos(2 * np.pi * k2y / n))
            for o3 in range(L):
                k3x, k3y = o3 % n, o3 // n
                for o4 in range(L):
                    k4x, k4y = o4 % n, o4 // n
                    v = 0.0
                    if (k1x + k2x) % n == (k3x + k4x) % n and (k1y + k2y) % n == (k3y + k4y) % n:
                        v += U / L
                    if o1 == o3 and o2 == o4:
                        v += (t1 + t2) / (N - 1)
                    mom[o1, o2, o3, o4] = v
    print("case test9 3x3 Hubbard, momentum basis", flush=True)
    e_mom = _solve(L, N, TwoS, np.ascontiguousarray(mom.transpose(3, 2, 1, 0)).ravel(), scheme)
    print(f"test9: site {e_site!r} momentum {e_mom!r}", flush=True)
    assert abs(e_site - e_mom) < 1e-8
