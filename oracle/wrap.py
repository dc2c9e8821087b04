"""ctypes wrapper around the CPU oracle library ``oracle/liboracle.so``.

Test infrastructure only: the test-suite, the smoke check and the CPU-baseline leg of
the benchmark use it as the checker; the product library never loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

KINDS = {"L": 0, "F0": 1, "F1": 2, "S0": 3, "S1": 4, "A": 5, "B": 6, "C": 7, "D": 8, "Q": 9, "X": 10}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force=False):
    """Compile the oracle with its Makefile (g++ only)."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-B" if force else "liboracle.so"], cwd=_HERE, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        # the oracle parallelises over blocks with OpenMP; MKL runs sequentially inside
        os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
        l = C.CDLL(_LIB)
        vp, i, d, L_ = C.c_void_p, C.c_int, C.c_double, C.c_long
        sig = {
            "orc_create": (vp, [i, i, _ip, i, i, i, _dp, _dp, d, i]),
            "orc_destroy": (None, [vp]),
            "orc_init": (None, [vp, i, C.c_ulonglong]),
            "orc_solve": (d, [vp, i, _ip, _dp, _ip, _dp, _dp]),
            "orc_matvec_count": (L_, [vp]),
            "orc_matvec_flops": (d, [vp]),
            "orc_sweep_left_to": (d, [vp, i, i, d, i]),
            "orc_step_left": (d, [vp, i, i, d, i]),
            "orc_synthetic": (L_, [vp, i, i, C.c_ulonglong, C.c_ulonglong]),
            "orc_bookkeeper": (None, [vp, i]),
            "orc_split_test": (d, [vp, i, i, i, i, _dp]),
            "orc_threads": (i, []),
            "orc_site_apply_range": (d, [vp, _dp, _dp, i, i, C.POINTER(C.c_double)]),
            "orc_site": (L_, [vp, i]),
            "orc_load_mps": (L_, [vp, i, i, C.POINTER(C.c_void_p), C.POINTER(C.c_long)]),
            "orc_site_x": (None, [vp, _dp]),
            "orc_site_apply": (d, [vp, _dp, _dp]),
            "orc_site_apply_timed": (d, [vp, _dp, _dp, i]),
            "orc_site_diag": (None, [vp, _dp]),
            "orc_site_davidson": (d, [vp, d, _dp, C.POINTER(C.c_int)]),
            "orc_site_nkappa": (i, [vp]),
            "orc_site_kappa": (None, [vp, _ip, _ip]),
            "orc_L": (i, [vp]),
            "orc_nirrep": (i, [vp]),
            "orc_irrep": (i, [vp, i]),
            "orc_mx": (None, [vp, _dp]),
            "orc_sectors": (i, [vp, i, C.c_void_p]),
            "orc_op": (L_, [vp, i, i, i, i, C.c_void_p, C.POINTER(C.POINTER(C.c_double))]),
            "orc_op_qsite": (i, [vp, i, i]),
            "orc_mps": (L_, [vp, i, C.POINTER(C.POINTER(C.c_double))]),
            "orc_wigner6j": (d, [i] * 6),
            "orc_wigner9j": (d, [i] * 9),
            "orc_davidson_dense": (d, [i, _dp, _dp, d, _dp, C.POINTER(C.c_int)]),
            "orc_2rdm": (d, [vp, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


class OracleDMRG:
    """One oracle DMRG instance (problem + bookkeeper + MPS + boundary operators)."""

    def __init__(self, L, group, irreps, N, TwoS, Irrep, T, V, econst=0.0, reorder=False):
        self.lib = lib()
        irreps = np.ascontiguousarray(irreps, dtype=np.int32)
        T = np.ascontiguousarray(T, dtype=np.float64).reshape(L, L)
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(L, L, L, L)
        # flat index i + L(j + L(k + L l)) == Fortran order of V[i,j,k,l]
        Tf = np.ascontiguousarray(T.T).ravel()
        Vf = np.ascontiguousarray(V.transpose(3, 2, 1, 0)).ravel()
        self.h = self.lib.orc_create(L, group, irreps, N, TwoS, Irrep, Tf, Vf, float(econst), int(reorder))
        self.L = L
        self.group, self.N, self.TwoS, self.Irrep = group, N, TwoS, Irrep

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def init(self, D0, seed=1):
        self.lib.orc_init(self.h, int(D0), int(seed))

    def solve(self, scheme):
        """scheme: list of (D, econv, max_sweeps, noise, rtol)."""
        n = len(scheme)
        D = np.array([s[0] for s in scheme], dtype=np.int32)
        ec = np.array([s[1] for s in scheme], dtype=np.float64)
        ms = np.array([s[2] for s in scheme], dtype=np.int32)
        nz = np.array([s[3] for s in scheme], dtype=np.float64)
        rt = np.array([s[4] for s in scheme], dtype=np.float64)
        return self.lib.orc_solve(self.h, n, D, ec, ms, nz, rt)

    def sweep_left_to(self, stop, D, rtol=1e-5, change=True):
        return self.lib.orc_sweep_left_to(self.h, int(stop), int(D), float(rtol), int(change))

    def synthetic(self, D, site, seed_l, seed_r):
        """Benchmark workload: dimensions at D and pseudo-random operators around site."""
        self.index = site
        self.V = int(self.lib.orc_synthetic(self.h, int(D), int(site), int(seed_l), int(seed_r)))
        return self.V

    def split_test(self, index, D, moving_right, change, s_prog):
        return self.lib.orc_split_test(self.h, int(index), int(D), int(moving_right), int(change),
                                       np.ascontiguousarray(s_prog, dtype=np.float64))

    def bookkeeper(self, D):
        self.lib.orc_bookkeeper(self.h, int(D))

    def apply_range(self, x, k0, k1):
        """Heff x on output blocks [k0, k1): (y, flops, seconds)."""
        y = np.zeros(self.V)
        sec = C.c_double(0.0)
        fl = self.lib.orc_site_apply_range(self.h, np.ascontiguousarray(x, dtype=np.float64), y, int(k0), int(k1),
                                           C.byref(sec))
        return y, fl, sec.value

    def step_left(self, index, D, rtol=1e-5, change=True):
        return self.lib.orc_step_left(self.h, int(index), int(D), float(rtol), int(change))

    def load_mps(self, D, site, mps):
        """Bookkeeper at D, the given site tensors, the operators around `site` built by
        the oracle's updates, and the two-site step at `site` (returns V)."""
        arrs = [np.ascontiguousarray(m, dtype=np.float64) for m in mps]
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        sizes = (C.c_long * len(arrs))(*[a.size for a in arrs])
        v = int(self.lib.orc_load_mps(self.h, int(D), int(site), ptrs, sizes))
        if v < 0:
            raise ValueError("MPS tensor sizes do not match the bookkeeper at D")
        self.index = site
        self.V = v
        return v

    def site(self, index):
        self.index = index
        self.V = int(self.lib.orc_site(self.h, int(index)))
        return self.V

    def site_x(self):
        x = np.zeros(self.V)
        self.lib.orc_site_x(self.h, x)
        return x

    def apply(self, x):
        y = np.zeros(self.V)
        fl = self.lib.orc_site_apply(self.h, np.ascontiguousarray(x, dtype=np.float64), y)
        return y, fl

    def apply_timed(self, x, reps):
        y = np.zeros(self.V)
        t = self.lib.orc_site_apply_timed(self.h, np.ascontiguousarray(x, dtype=np.float64), y, int(reps))
        return y, t

    def diag(self):
        d = np.zeros(self.V)
        self.lib.orc_site_diag(self.h, d)
        return d

    def davidson(self, rtol=1e-10):
        x = np.zeros(self.V)
        nm = C.c_int(0)
        e = self.lib.orc_site_davidson(self.h, float(rtol), x, C.byref(nm))
        return e, x, nm.value

    def kappa(self):
        nk = self.lib.orc_site_nkappa(self.h)
        lab = np.zeros(9 * nk, dtype=np.int32)
        k2i = np.zeros(nk + 1, dtype=np.int32)
        self.lib.orc_site_kappa(self.h, lab, k2i)
        return lab.reshape(nk, 9), k2i

    def irreps(self):
        return np.array([self.lib.orc_irrep(self.h, s) for s in range(self.L)], dtype=np.int32)

    def nirrep(self):
        return self.lib.orc_nirrep(self.h)

    def mx(self):
        out = np.zeros(self.L ** 4)
        self.lib.orc_mx(self.h, out)
        return out

    def sectors(self, b):
        n = self.lib.orc_sectors(self.h, int(b), None)
        out = np.zeros(4 * n, dtype=np.int32)
        self.lib.orc_sectors(self.h, int(b), out.ctypes.data)
        return out.reshape(n, 4)

    def op(self, cnt, kind, k=0, m=0):
        """Returns (meta, data copy) or None if the tensor is absent."""
        meta = np.zeros(8, dtype=np.int32)
        ptr = C.POINTER(C.c_double)()
        n = self.lib.orc_op(self.h, int(cnt), KINDS[kind], int(k), int(m), meta.ctypes.data, C.byref(ptr))
        if n < 0:
            return None
        data = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0)
        return meta, data

    def mps(self, site):
        """Copy of the MPS tensor of `site` (TensorT storage)."""
        ptr = C.POINTER(C.c_double)()
        n = self.lib.orc_mps(self.h, int(site), C.byref(ptr))
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0)

    def two_rdm(self):
        """2-RDMs (A, B) of the converged state as L x L x L x L arrays indexed [i, j, k, l]
        (DMRG orbital order) and the energy 0.5 trace(A h) + Econst. Call after solve()."""
        L = self.L
        A = np.zeros(L ** 4)
        B = np.zeros(L ** 4)
        e = self.lib.orc_2rdm(self.h, A, B)
        return A.reshape(L, L, L, L, order="F"), B.reshape(L, L, L, L, order="F"), e

    def matvec_stats(self):
        return self.lib.orc_matvec_count(self.h), self.lib.orc_matvec_flops(self.h)


def threads():
    return lib().orc_threads()


def wigner6j(*a):
    return lib().orc_wigner6j(*[int(x) for x in a])


def wigner9j(*a):
    return lib().orc_wigner9j(*[int(x) for x in a])


def davidson_dense(A, x0, rtol=1e-10):
    n = A.shape[0]
    Af = np.ascontiguousarray(A.T).ravel()
    x = np.zeros(n)
    nm = C.c_int(0)
    e = lib().orc_davidson_dense(n, Af, np.ascontiguousarray(x0, dtype=np.float64), float(rtol), x, C.byref(nm))
    return e, x, nm.value
