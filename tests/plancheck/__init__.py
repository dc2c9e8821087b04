"""CPU executor of the chemps2_amd Heff work list (test infrastructure).

The generator is built host-only (``tests/plancheck/Makefile``); its terms,
combinations and diagonal terms are executed here with numpy so the work list can be
compared with the oracle without a GPU.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libplancheck.so")
_lib = None
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
KIND_NAMES = ["L", "F0", "F1", "S0", "S1", "A", "B", "C", "D", "Q", "X"]


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s"], cwd=HERE, check=True)
        l = C.CDLL(LIB)
        vp, i, d, lg = C.c_void_p, C.c_int, C.c_double, C.c_long
        sig = {
            "pc_create": (vp, [i, i, _ip, i, i, i, _dp]),
            "pc_destroy": (None, [vp]),
            "pc_sectors": (None, [vp, i, i, _ip, _ip, _ip, _ip]),
            "pc_sectors_auto": (None, [vp, i]),
            "pc_nsectors": (i, [vp, i, C.c_void_p]),
            "pc_fci": (i, [vp, i, C.c_void_p]),
            "pc_tensor": (i, [vp, i, i, i, i, i, i, i, i, C.POINTER(C.c_long)]),
            "pc_synthetic_set": (i, [vp, i, i]),
            "pc_plan": (lg, [vp, i]),
            "pc_error": (C.c_char_p, [vp]),
            "pc_flops": (d, [vp]),
            "pc_nk": (i, [vp]),
            "pc_k2i": (None, [vp, _lp, _ip]),
            "pc_terms": (None, [vp, _ip, _lp, _dp]),
            "pc_ncombos": (lg, [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
            "pc_combos": (None, [vp, _lp, _ip, _lp, _dp]),
            "pc_ndiag": (lg, [vp]),
            "pc_diag": (None, [vp, _ip, _lp, _dp]),
            "pc_mps": (i, [vp, i, C.POINTER(C.c_long)]),
            "pc_update": (lg, [vp, i, i]),
            "pc_update_newt": (None, [vp, _lp]),
            "pc_update_nterms": (lg, [vp, i]),
            "pc_update_terms": (None, [vp, i, _ip, _lp, _dp]),
            "pc_update_ncombos": (lg, [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
            "pc_update_combos": (None, [vp, _lp, _ip, _lp, _dp]),
            "pc_update_flops": (d, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def oracle_tensors(o, cnt):
    """Every operator tensor of oracle slot cnt: list of (kind, k, m, meta, data)."""
    out = []
    for kind in KIND_NAMES:
        for k in range(o.L + 1):
            for m in range(o.L + 1):
                r = o.op(cnt, kind, k, m)
                if r is not None:
                    out.append((kind, k, m, r[0], r[1]))
                if kind in ("L", "Q", "X"):
                    break
            if kind == "X":
                break
    return out


class Plan:
    """Work list for the oracle's current site, with the oracle's tensor data."""

    def __init__(self, o, site):
        l = lib()
        self.l = l
        L = o.L
        self.h = l.pc_create(L, int(o.group), o.irreps(), o.N, o.TwoS, o.Irrep, o.mx())
        for b in range(L + 1):
            s = o.sectors(b)
            l.pc_sectors(self.h, b, len(s), *[np.ascontiguousarray(s[:, c]) for c in range(4)])
        self.bufs = {}
        for cnt, mr in ((site - 1, 1), (site + 1, 0)):
            if cnt < 0 or cnt > L - 2:
                continue
            for kind, k, m, meta, data in oracle_tensors(o, cnt):
                size = C.c_long(0)
                buf = l.pc_tensor(self.h, cnt, mr, KIND_NAMES.index(kind), k, m, int(meta[1]), int(meta[2]),
                                  int(meta[3]), C.byref(size))
                assert size.value == len(data), (kind, k, m, size.value, len(data))
                self.bufs[buf] = data
        n = l.pc_plan(self.h, site)
        if n < 0:
            raise RuntimeError(l.pc_error(self.h).decode())
        self.flops = l.pc_flops(self.h)
        nk = l.pc_nk(self.h)
        self.k2i = np.zeros(nk + 1, dtype=np.int64)
        self.dims = np.zeros(2 * nk, dtype=np.int32)
        l.pc_k2i(self.h, self.k2i, self.dims)
        self.dims = self.dims.reshape(nk, 2)
        self.ti = np.zeros(8 * n, dtype=np.int32)
        self.to = np.zeros(2 * n, dtype=np.int64)
        self.ta = np.zeros(n, dtype=np.float64)
        l.pc_terms(self.h, self.ti, self.to, self.ta)
        self.ti = self.ti.reshape(n, 8)
        self.to = self.to.reshape(n, 2)
        ns, ar = C.c_long(0), C.c_long(0)
        nc = l.pc_ncombos(self.h, C.byref(ns), C.byref(ar))
        meta = np.zeros(3 * max(nc, 1), dtype=np.int64)
        sb = np.zeros(max(ns.value, 1), dtype=np.int32)
        so = np.zeros(max(ns.value, 1), dtype=np.int64)
        sc = np.zeros(max(ns.value, 1), dtype=np.float64)
        l.pc_combos(self.h, meta, sb, so, sc)
        arena = np.zeros(max(ar.value, 1))
        pos = 0
        for q in range(nc):
            off, size, nsrc = meta[3 * q: 3 * q + 3]
            acc = np.zeros(size)
            for e in range(pos, pos + nsrc):
                acc += sc[e] * self.bufs[sb[e]][so[e]: so[e] + size]
            arena[off: off + size] = acc
            pos += nsrc
        self.bufs[1] = arena
        nd = l.pc_ndiag(self.h)
        self.di = np.zeros(4 * max(nd, 1), dtype=np.int32)
        self.do = np.zeros(2 * max(nd, 1), dtype=np.int64)
        self.dc = np.zeros(max(nd, 1))
        l.pc_diag(self.h, self.di, self.do, self.dc)
        self.nd = nd
        self.V = int(self.k2i[-1])

    def __del__(self):
        if getattr(self, "h", None):
            self.l.pc_destroy(self.h)

    def _blk(self, buf, off, shape):
        n = shape[0] * shape[1]
        return self.bufs[buf][off: off + n].reshape(shape[1], shape[0]).T

    def apply(self, x):
        y = np.zeros(self.V)
        for q in range(len(self.ta)):
            out, inn, tl, tr, xr, lb, rb, _ = self.ti[q]
            dL, dR = self.dims[out]
            dLi, dRi = self.dims[inn]
            X = x[self.k2i[inn]: self.k2i[inn + 1]].reshape(dRi, dLi).T
            M = X
            if chr(tl) != "I":
                Lm = self._blk(lb, self.to[q, 0], (dL, dLi) if chr(tl) == "N" else (dLi, dL))
                M = (Lm if chr(tl) == "N" else Lm.T) @ M
            if chr(tr) != "I":
                Rm = self._blk(rb, self.to[q, 1], (dRi, dR) if chr(tr) == "N" else (dR, dRi))
                M = M @ (Rm if chr(tr) == "N" else Rm.T)
            y[self.k2i[out]: self.k2i[out + 1]] += (self.ta[q] * M).T.ravel()
        return y

    def diag(self):
        d = np.zeros(self.V)
        for q in range(self.nd):
            k, kind, lb, rb = self.di[4 * q: 4 * q + 4]
            dL, dR = self.dims[k]
            blk = d[self.k2i[k]: self.k2i[k + 1]].reshape(dR, dL).T
            c = self.dc[q]
            if kind == 0:
                blk += c
            elif kind == 1:
                blk += c * self.bufs[lb][self.do[2 * q]: self.do[2 * q] + dL * dL:dL + 1][:, None]
            elif kind == 2:
                blk += c * self.bufs[rb][self.do[2 * q + 1]: self.do[2 * q + 1] + dR * dR:dR + 1][None, :]
            else:
                dl_ = self.bufs[lb][self.do[2 * q]: self.do[2 * q] + dL * dL:dL + 1]
                dr_ = self.bufs[rb][self.do[2 * q + 1]: self.do[2 * q + 1] + dR * dR:dR + 1]
                blk += c * np.outer(dl_, dr_)
        return d


def _op(M, t, rows, cols, ld):
    """op(M) as a (rows x cols) view of a column-major block with leading dimension ld."""
    if t:
        return M[: cols * ld].reshape(rows, ld).T[:cols, :].T if False else M[: (rows - 1) * ld + cols].copy()
    return None


def _blk(buf, off, rows, cols, ld, trans):
    """The (rows x cols) operand op(B) where B is stored column-major with leading dimension ld."""
    if not trans:
        n = (cols - 1) * ld + rows
        flat = buf[off: off + n]
        return np.lib.stride_tricks.as_strided(flat, shape=(rows, cols), strides=(8, 8 * ld))
    n = (rows - 1) * ld + cols
    flat = buf[off: off + n]
    return np.lib.stride_tricks.as_strided(flat, shape=(rows, cols), strides=(8 * ld, 8))


class UpdatePlan:
    """Operator-update work list for slot `index`, executed with numpy on the oracle's data."""

    def __init__(self, o, index, moving_right):
        l = lib()
        self.l = l
        L = o.L
        self.h = l.pc_create(L, int(o.group), o.irreps(), o.N, o.TwoS, o.Irrep, o.mx())
        for b in range(L + 1):
            s = o.sectors(b)
            l.pc_sectors(self.h, b, len(s), *[np.ascontiguousarray(s[:, c]) for c in range(4)])
        bufs = {}
        prev = index - 1 if moving_right else index + 1
        if 0 <= prev <= L - 2:
            for kind, k, m, meta, data in oracle_tensors(o, prev):
                size = C.c_long(0)
                buf = l.pc_tensor(self.h, prev, 1 if moving_right else 0, KIND_NAMES.index(kind), k, m, int(meta[1]),
                                  int(meta[2]), int(meta[3]), C.byref(size))
                assert size.value == len(data)
                bufs[buf] = data
        tsite = index if moving_right else index + 1
        size = C.c_long(0)
        buf = l.pc_mps(self.h, tsite, C.byref(size))
        T = o.mps(tsite)
        assert size.value == len(T), (size.value, len(T))
        bufs[buf] = T
        nt = l.pc_update(self.h, index, 1 if moving_right else 0)
        if nt < 0:
            raise RuntimeError(l.pc_error(self.h).decode())
        newt = np.zeros(5 * nt, dtype=np.int64)
        l.pc_update_newt(self.h, newt)
        self.newt = newt.reshape(nt, 5)
        for kind, k, m, b, sz in self.newt:
            bufs[int(b)] = np.zeros(int(sz))
        ns, ar = C.c_long(0), C.c_long(0)
        nc = l.pc_update_ncombos(self.h, C.byref(ns), C.byref(ar))
        meta = np.zeros(3 * max(nc, 1), dtype=np.int64)
        sb = np.zeros(max(ns.value, 1), dtype=np.int32)
        so = np.zeros(max(ns.value, 1), dtype=np.int64)
        sc = np.zeros(max(ns.value, 1))
        l.pc_update_combos(self.h, meta, sb, so, sc)
        arena = np.zeros(max(ar.value, 1))
        pos = 0
        for q in range(nc):
            off, sz, nsrc = meta[3 * q: 3 * q + 3]
            acc = np.zeros(sz)
            for e in range(pos, pos + nsrc):
                acc += sc[e] * bufs[sb[e]][so[e]: so[e] + sz]
            arena[off: off + sz] = acc
            pos += nsrc
        bufs[1] = arena
        for level in (1, 2):
            n = l.pc_update_nterms(self.h, level)
            ints = np.zeros(16 * max(n, 1), dtype=np.int32)
            offs = np.zeros(4 * max(n, 1), dtype=np.int64)
            al = np.zeros(max(n, 1))
            l.pc_update_terms(self.h, level, ints, offs, al)
            ints = ints.reshape(-1, 16)
            offs = offs.reshape(-1, 4)
            for q in range(n):
                ob, m_, n_, k1, k2, ab, bb, cb, ta, tb, tc, lda, ldb, ldc, kind, _ = ints[q]
                oo, ao, bo, co = offs[q]
                if kind == 2:
                    M = al[q] * _blk(bufs[ab], ao, m_, n_, lda, ta)
                elif kind == 0:
                    M = al[q] * (_blk(bufs[ab], ao, m_, k1, lda, ta) @ _blk(bufs[bb], bo, k1, n_, ldb, tb))
                else:
                    W = al[q] * (_blk(bufs[ab], ao, m_, k1, lda, ta) @ _blk(bufs[bb], bo, k1, k2, ldb, tb))
                    M = W @ _blk(bufs[cb], co, k2, n_, ldc, tc)
                out = bufs[ob]
                out[oo: oo + m_ * n_] += M.T.ravel()
        self.bufs = bufs
        self.flops = l.pc_update_flops(self.h)

    def tensors(self):
        """{(kind name, k, m): data} of the new slot."""
        return {(KIND_NAMES[int(kd)], int(k), int(m)): self.bufs[int(b)] for kd, k, m, b, sz in self.newt}

    def __del__(self):
        if getattr(self, "h", None):
            self.l.pc_destroy(self.h)
