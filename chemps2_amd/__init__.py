"""chemps2_amd: MI355X implementation of CheMPS2's two-site effective-Hamiltonian path.

Python access to the C ABI of ``libchemps2_amd.so`` (declared in
``include/chemps2_amd.h``). ``Heff`` mirrors the reference's ``CheMPS2::Heff``
(Heff.h:50-89): construct it for a prepared two-site step, then ``makeHeff``,
``fillHeffDiag`` and ``SolveDAVIDSON`` run on the GPU. There is no CPU fallback: loading
fails loudly when the library has not been built or no GPU is present.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchemps2_amd.so")
TUNING_PATH = os.path.join(HERE, "data", "autotune.json")

KINDS = {"L": 0, "F0": 1, "F1": 2, "S0": 3, "S1": 4, "A": 5, "B": 6, "C": 7, "D": 8, "Q": 9, "X": 10}
EXPORTS = [
    "chm_ctx_create", "chm_ctx_destroy", "chm_last_error", "chm_set_problem", "chm_set_sectors",
    "chm_set_sectors_auto", "chm_get_sectors", "chm_tensor_upload", "chm_tensor_fill_synthetic",
    "chm_tensor_clear", "chm_heff_prepare", "chm_heff_veclen", "chm_heff_apply", "chm_heff_diag",
    "chm_davidson", "chm_flops", "chm_plan_stats", "chm_stream", "chm_timing_enable", "chm_timing_read",
    "chm_mps_upload", "chm_mps_download", "chm_tensor_download", "chm_update_operators",
    "chm_join", "chm_s_get", "chm_s_set", "chm_split", "chm_get_problem", "chm_mps_layout_size", "chm_dmrg_init", "chm_dmrg_presolve", "chm_dmrg_solve", "chm_dmrg_sweep", "chm_dmrg_heff_stats", "chm_dmrg_stats",
    "chm_set_heff_form", "chm_heff_apply_f32", "chm_mixed_stats", "chm_heff_set_lower",
    "chm_precond_mode", "chm_mlp_nparams", "chm_mlp_set", "chm_mlp_get", "chm_mlp_forward", "chm_mlp_loss_grad",
    "chm_mlp_train", "chm_harvest", "chm_harvest_take", "chm_dmrg_sweep_range",
    "chm_dmrg_new_excitation", "chm_dmrg_nstates", "chm_dmrg_site_log", "chm_dmrg_2rdm", "chm_heff_apply_noalloc", "chm_davidson_stats",
    "chm_set_heff_knobs", "chm_set_tuning", "chm_tuning_used", "chm_precond_stats",
]
PRECOND_DIAGONAL, PRECOND_LEARNED = 0, 1
MLP_IN, MLP_H = 8, 128
FORM_FP64, FORM_MIXED = 0, 1

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class ChmError(RuntimeError):
    pass


def load():
    """Load libchemps2_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ChmError(f"{LIB_PATH} is missing: build it with chemps2_amd/build.py (hipcc, gfx950)")
    l = C.CDLL(LIB_PATH)
    vp, i, d, i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
    sig = {
        "chm_ctx_create": (i, [i, C.POINTER(vp)]),
        "chm_ctx_destroy": (i, [vp]),
        "chm_last_error": (C.c_char_p, [vp]),
        "chm_set_problem": (i, [vp, i, i, _ip, i, i, i, _dp, d]),
        "chm_set_sectors": (i, [vp, i, i, _ip, _ip, _ip, _ip]),
        "chm_set_sectors_auto": (i, [vp, i]),
        "chm_get_sectors": (i, [vp, i, C.POINTER(i), vp, vp, vp, vp]),
        "chm_tensor_upload": (i, [vp, i, i, i, i, i, i, i, i, _dp, i64]),
        "chm_tensor_fill_synthetic": (i, [vp, i, i, C.c_uint64]),
        "chm_tensor_clear": (i, [vp, i]),
        "chm_mps_upload": (i, [vp, i, _dp, i64]),
        "chm_mps_download": (i, [vp, i, _dp, i64]),
        "chm_tensor_download": (i, [vp, i, i, i, i, _dp, i64]),
        "chm_update_operators": (i, [vp, i, i]),
        "chm_join": (i, [vp, i, C.POINTER(i64)]),
        "chm_s_get": (i, [vp, vp, i64, i]),
        "chm_s_set": (i, [vp, vp, i64, i]),
        "chm_split": (i, [vp, i, i, i, i, C.POINTER(d)]),
        "chm_dmrg_init": (i, [vp, i, C.c_uint64]),
        "chm_dmrg_new_excitation": (i, [vp, d, i, C.c_uint64]),
        "chm_dmrg_nstates": (i, [vp, C.POINTER(i)]),
        "chm_dmrg_2rdm": (i, [vp, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
        "chm_heff_apply_noalloc": (i, [vp, C.c_void_p, C.c_void_p]),
        "chm_davidson_stats": (i, [vp, C.POINTER(C.c_int64)]),
        "chm_dmrg_site_log": (i, [vp, C.POINTER(i), vp, vp, vp]),
        "chm_set_heff_knobs": (i, [vp, i64, i, i, i, i]),
        "chm_set_tuning": (i, [vp, i, vp]),
        "chm_tuning_used": (i, [vp, vp]),
        "chm_precond_stats": (i, [vp, C.POINTER(i64)]),
        "chm_dmrg_solve": (i, [vp, i, _ip, _dp, _ip, _dp, _dp, C.POINTER(d)]),
        "chm_dmrg_sweep": (i, [vp, i, i, d, d, i, C.POINTER(d)]),
        "chm_get_problem": (i, [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
        "chm_dmrg_presolve": (i, [vp]),
        "chm_mps_layout_size": (i, [vp, i, C.POINTER(i64)]),
        "chm_dmrg_stats": (i, [vp, C.POINTER(d), C.POINTER(i64)]),
        "chm_dmrg_heff_stats": (i, [vp, C.POINTER(d), C.POINTER(d)]),
        "chm_heff_prepare": (i, [vp, i]),
        "chm_heff_veclen": (i, [vp, C.POINTER(i64)]),
        "chm_heff_apply": (i, [vp, vp, vp, i]),
        "chm_heff_apply_f32": (i, [vp, vp, vp, i]),
        "chm_set_heff_form": (i, [vp, i]),
        "chm_heff_set_lower": (i, [vp, i, vp, i]),
        "chm_mixed_stats": (i, [vp, C.POINTER(i64), C.POINTER(i64)]),
        "chm_precond_mode": (i, [vp, i]),
        "chm_mlp_nparams": (i, [C.POINTER(i64)]),
        "chm_mlp_set": (i, [vp, vp, i64]),
        "chm_mlp_get": (i, [vp, vp, i64]),
        "chm_mlp_forward": (i, [vp, vp, i64, vp]),
        "chm_mlp_loss_grad": (i, [vp, vp, vp, vp, vp, i64, vp, C.POINTER(d)]),
        "chm_mlp_train": (i, [vp, vp, vp, vp, vp, i64, i, i, d, C.c_uint64, C.POINTER(d), C.POINTER(d)]),
        "chm_harvest": (i, [vp, i]),
        "chm_dmrg_sweep_range": (i, [vp, i, i, i, i, d, d, i, C.POINTER(d)]),
        "chm_harvest_take": (i, [vp, C.POINTER(i64), vp, vp, vp, vp]),
        "chm_heff_diag": (i, [vp, vp, i]),
        "chm_davidson": (i, [vp, vp, i, d, C.POINTER(d), C.POINTER(i)]),
        "chm_flops": (i, [vp, C.POINTER(d), C.POINTER(d)]),
        "chm_plan_stats": (i, [vp, C.POINTER(i64)]),
        "chm_stream": (i, [vp, C.POINTER(vp)]),
        "chm_timing_enable": (i, [vp, i]),
        "chm_timing_read": (i, [vp, C.POINTER(d), C.POINTER(i64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(l, name)
        f.restype = res
        f.argtypes = args
    _lib = l
    return l


def mlp_nparams():
    n = C.c_int64(0)
    load().chm_mlp_nparams(C.byref(n))
    return n.value


def mlp_init(seed=0):
    """Initial parameters: He-normal hidden layers, zero output layer (m = 0, the plain
    diagonal preconditioner)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in [(MLP_IN, MLP_H), (MLP_H, MLP_H), (MLP_H, MLP_H)]:
        parts.append(rng.normal(scale=np.sqrt(2.0 / fan_in), size=(fan_in, fan_out)).ravel())
        parts.append(np.zeros(fan_out))
    parts.append(np.zeros(MLP_H))
    parts.append(np.zeros(1))
    return np.concatenate(parts).astype(np.float32)


def _samples(X, p, e, w):
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_IN)
    p, e, w = (np.ascontiguousarray(a, dtype=np.float32) for a in (p, e, w))
    if not (len(X) == len(p) == len(e) == len(w)):
        raise ChmError("sample arrays differ in length")
    return X, p, e, w


IO_LIB_PATH = os.path.join(HERE, "libchemps2_amd_io.so")
IO_EXPORTS = ["chm_mps_save", "chm_mps_load_dims", "chm_mps_load", "chm_io_write", "chm_io_read_dims", "chm_io_read_site"]
_io = None


def load_io():
    """Load libchemps2_amd_io.so (MPS checkpoints in the reference's HDF5 layout)."""
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(IO_LIB_PATH):
        raise ChmError(f"{IO_LIB_PATH} is missing: build it with chemps2_amd/build.py (needs HDF5)")
    load()  # the checkpoint library resolves the core library's symbols
    l = C.CDLL(IO_LIB_PATH)
    vp, i, i64 = C.c_void_p, C.c_int, C.c_int64
    sig = {
        "chm_mps_save": (i, [vp, C.c_char_p, i]),
        "chm_mps_load_dims": (i, [vp, C.c_char_p]),
        "chm_mps_load": (i, [vp, C.c_char_p, C.POINTER(i)]),
        "chm_io_write": (i, [C.c_char_p, i, i, i, i, i, _ip, C.POINTER(C.c_void_p), np.ctypeslib.ndpointer(dtype=np.int64)]),
        "chm_io_read_dims": (i, [C.c_char_p, i, i, i, i, vp, C.POINTER(i)]),
        "chm_io_read_site": (i, [C.c_char_p, i, vp, i64, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(l, name)
        f.restype = res
        f.argtypes = args
    _io = l
    return l


def io_write(path, L, N, two_s, nirrep, converged, dims, mps):
    """Host-only checkpoint writer: dims over SyBookkeeper's sector ranges, mps a list of
    L arrays in TensorT order (DMRG::saveMPS layout)."""
    l = load_io()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in mps]
    ptrs = (C.c_void_p * L)(*[a.ctypes.data for a in arrs])
    sizes = np.array([a.size for a in arrs], dtype=np.int64)
    rc = l.chm_io_write(path.encode(), L, N, two_s, nirrep, int(converged), np.ascontiguousarray(dims, dtype=np.int32),
                        ptrs, sizes)
    if rc != 0:
        raise ChmError(f"chm_io_write failed ({rc})")


def io_read_dims(path, L, N, two_s, nirrep):
    l = load_io()
    n = C.c_int(0)
    l.chm_io_read_dims(path.encode(), L, N, two_s, nirrep, None, C.byref(n))
    out = np.zeros(n.value, dtype=np.int32)
    rc = l.chm_io_read_dims(path.encode(), L, N, two_s, nirrep, out.ctypes.data, C.byref(n))
    if rc != 0:
        raise ChmError(f"chm_io_read_dims failed ({rc})")
    return out


def io_read_site(path, site, size):
    l = load_io()
    out = np.zeros(max(size, 1))
    conv = C.c_int(-1)
    rc = l.chm_io_read_site(path.encode(), int(site), out.ctypes.data, int(size), C.addressof(conv))
    if rc != 0:
        raise ChmError(f"chm_io_read_site failed ({rc})")
    return out[:size], conv.value


def _ptr(a):
    """Address of a numpy array or a torch tensor (device pointers for on_device calls)."""
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def _dev(t, device, n, what):
    """Device address of a torch tensor handed to an on_device call: float64, contiguous,
    on the context's GPU and holding at least n elements. Work already queued on torch's
    current stream is waited for, since the library reads on its own stream."""
    import torch

    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what}: a numpy array or a torch tensor is required")
    if t.dtype != torch.float64:
        raise TypeError(f"{what}: float64 required, got {t.dtype}")
    if not t.is_cuda or t.device.index != device:
        raise ValueError(f"{what}: the tensor must live on cuda:{device}, not {t.device}")
    if not t.is_contiguous() or t.numel() < n:
        raise ValueError(f"{what}: a contiguous tensor of at least {n} elements is required")
    torch.cuda.current_stream(t.device).synchronize()
    return t.data_ptr()


class Context:
    """One library context on one GPU (chm_ctx)."""

    def __init__(self, device=0, tuning=True):
        """tuning: install the autotuner's table (data/autotune.json, written by
        chemps2_amd.autotune on the GPU) when it exists."""
        self.lib = load()
        self.h = C.c_void_p()
        self.device = int(device)
        rc = self.lib.chm_ctx_create(int(device), C.byref(self.h))
        if rc != 0:
            raise ChmError(f"chm_ctx_create failed ({rc}): no usable HIP device {device}")
        if tuning and os.path.exists(TUNING_PATH):
            import json

            with open(TUNING_PATH) as f:
                self.set_tuning(json.load(f)["rows"])

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.chm_last_error(self.h)
            raise ChmError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.chm_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_problem(self, L, group, orb_irreps, N, two_s, irrep, mx, econst=0.0):
        mx = np.ascontiguousarray(mx, dtype=np.float64).ravel()
        self.L = L
        self._check(self.lib.chm_set_problem(self.h, L, group, np.ascontiguousarray(orb_irreps, dtype=np.int32), N,
                                             two_s, irrep, mx, float(econst)), "chm_set_problem")

    def set_sectors(self, boundary, sectors):
        s = np.asarray(sectors, dtype=np.int32).reshape(-1, 4)
        cols = [np.ascontiguousarray(s[:, c]) for c in range(4)]
        self._check(self.lib.chm_set_sectors(self.h, boundary, len(s), *cols), "chm_set_sectors")

    def set_sectors_auto(self, D):
        self._check(self.lib.chm_set_sectors_auto(self.h, int(D)), "chm_set_sectors_auto")

    def get_sectors(self, boundary):
        n = C.c_int(0)
        self._check(self.lib.chm_get_sectors(self.h, boundary, C.byref(n), None, None, None, None), "chm_get_sectors")
        a = [np.zeros(n.value, dtype=np.int32) for _ in range(4)]
        self._check(self.lib.chm_get_sectors(self.h, boundary, C.byref(n), *[x.ctypes.data for x in a]), "chm_get_sectors")
        return np.stack(a, axis=1)

    def upload(self, cnt, moving_right, kind, k, m, two_j, n_elec, n_irrep, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        self._check(self.lib.chm_tensor_upload(self.h, cnt, int(moving_right), KINDS[kind] if isinstance(kind, str) else kind,
                                               k, m, two_j, n_elec, n_irrep, data, data.size), "chm_tensor_upload")

    def fill_synthetic(self, cnt, moving_right, seed):
        self._check(self.lib.chm_tensor_fill_synthetic(self.h, cnt, int(moving_right), int(seed)), "chm_tensor_fill_synthetic")

    def clear(self, cnt):
        self._check(self.lib.chm_tensor_clear(self.h, cnt), "chm_tensor_clear")

    def mps_upload(self, site, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        self._check(self.lib.chm_mps_upload(self.h, int(site), data, data.size), "chm_mps_upload")

    def mps_layout_size(self, site):
        """TensorT storage size of a site under the current sectors."""
        n = C.c_int64(0)
        self._check(self.lib.chm_mps_layout_size(self.h, int(site), C.byref(n)), "chm_mps_layout_size")
        return n.value

    def mps_download(self, site, n=None):
        out = np.zeros(int(self.mps_layout_size(site) if n is None else n))
        self._check(self.lib.chm_mps_download(self.h, int(site), out, out.size), "chm_mps_download")
        return out

    def tensor_download(self, cnt, kind, k, m, n):
        out = np.zeros(int(n))
        self._check(self.lib.chm_tensor_download(self.h, int(cnt), KINDS[kind] if isinstance(kind, str) else kind,
                                                 int(k), int(m), out, out.size), "chm_tensor_download")
        return out

    def update_operators(self, index, moving_right):
        self._check(self.lib.chm_update_operators(self.h, int(index), int(moving_right)), "chm_update_operators")

    def join(self, site):
        v = C.c_int64(0)
        self._check(self.lib.chm_join(self.h, int(site), C.byref(v)), "chm_join")
        self.SV = v.value
        return v.value

    def s_get(self):
        out = np.zeros(self.SV)
        self._check(self.lib.chm_s_get(self.h, _ptr(out), out.size, 0), "chm_s_get")
        return out

    def s_set(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        self._check(self.lib.chm_s_set(self.h, _ptr(x), x.size, 0), "chm_s_set")

    def split(self, site, D, moving_right, change=True):
        dw = C.c_double(0)
        self._check(self.lib.chm_split(self.h, int(site), int(D), int(moving_right), int(change), C.byref(dw)), "chm_split")
        return dw.value

    def dmrg_init(self, D, seed=1):
        self._check(self.lib.chm_dmrg_init(self.h, int(D), int(seed)), "chm_dmrg_init")

    def mps_save(self, path, converged=False):
        """DMRG::saveMPS: sectors and site tensors in the reference's HDF5 layout."""
        rc = load_io().chm_mps_save(self.h, str(path).encode(), int(converged))
        self._check(rc, "chm_mps_save")

    def mps_load_dims(self, path):
        """DMRG::loadDIM: the file's virtual dimensions become the current sectors."""
        self._check(load_io().chm_mps_load_dims(self.h, str(path).encode()), "chm_mps_load_dims")

    def mps_load(self, path):
        """DMRG::loadMPS: every site tensor from the file; returns the converged flag."""
        conv = C.c_int(0)
        self._check(load_io().chm_mps_load(self.h, str(path).encode(), C.byref(conv)), "chm_mps_load")
        return bool(conv.value)

    def dmrg_presolve(self):
        """DMRG::PreSolve on the current MPS."""
        self._check(self.lib.chm_dmrg_presolve(self.h), "chm_dmrg_presolve")

    def get_problem(self):
        v = [C.c_int(0) for _ in range(5)]
        self._check(self.lib.chm_get_problem(self.h, *[C.byref(x) for x in v]), "chm_get_problem")
        return dict(zip(["L", "N", "two_s", "irrep", "nirrep"], [x.value for x in v]))

    def new_excitation(self, eshift, D, seed=2):
        """DMRG::newExcitation: keep the current state as a lower state with level shift
        eshift and start the next one (SyBookkeeper dimensions at D, random MPS)."""
        self._check(self.lib.chm_dmrg_new_excitation(self.h, float(eshift), int(D), int(seed)), "chm_dmrg_new_excitation")

    def nstates(self):
        n = C.c_int(0)
        self._check(self.lib.chm_dmrg_nstates(self.h, C.byref(n)), "chm_dmrg_nstates")
        return n.value

    def two_rdm(self):
        """TwoDM of the current state (DMRG::calc_rdms_and_correlations, 2-RDM part): the
        spin-summed 2-RDMs A and B as L x L x L x L arrays indexed [i, j, k, l] (DMRG
        orbital order) and Econst + 0.5 sum A h. Call after dmrg_solve."""
        L = self.get_problem()["L"]
        A = np.zeros(L ** 4)
        B = np.zeros(L ** 4)
        e = C.c_double(0.0)
        self._check(self.lib.chm_dmrg_2rdm(self.h, A.ctypes.data, B.ctypes.data, C.byref(e)), "chm_dmrg_2rdm")
        return A.reshape(L, L, L, L, order="F"), B.reshape(L, L, L, L, order="F"), e.value

    def dmrg_solve(self, scheme):
        """scheme: list of (D, econv, max_sweeps, noise, rtol), as ConvergenceScheme."""
        D = np.array([s[0] for s in scheme], dtype=np.int32)
        ec = np.array([s[1] for s in scheme], dtype=np.float64)
        ms = np.array([s[2] for s in scheme], dtype=np.int32)
        nz = np.array([s[3] for s in scheme], dtype=np.float64)
        rt = np.array([s[4] for s in scheme], dtype=np.float64)
        e = C.c_double(0)
        self._check(self.lib.chm_dmrg_solve(self.h, len(scheme), D, ec, ms, nz, rt, C.byref(e)), "chm_dmrg_solve")
        return e.value

    def dmrg_sweep_range(self, moving_right, first, last, D, rtol=1e-5, noise=0.0, change=True):
        """Sites first..last of a half sweep (continuing it when first is not the chain end)."""
        e = C.c_double(0)
        self._check(self.lib.chm_dmrg_sweep_range(self.h, int(moving_right), int(first), int(last), int(D), float(rtol),
                                                  float(noise), int(change), C.byref(e)), "chm_dmrg_sweep_range")
        return e.value

    def dmrg_sweep(self, moving_right, D, rtol=1e-5, noise=0.0, change=True):
        """One half sweep (DMRG::sweepleft / sweepright); returns the last site's energy."""
        e = C.c_double(0)
        self._check(self.lib.chm_dmrg_sweep(self.h, int(moving_right), int(D), float(rtol), float(noise), int(change),
                                            C.byref(e)), "chm_dmrg_sweep")
        return e.value

    def dmrg_heff_stats(self):
        """(sum of F_H over the sweeps' Heff applications, their device time in ms)."""
        f, ms = C.c_double(0), C.c_double(0)
        self._check(self.lib.chm_dmrg_heff_stats(self.h, C.byref(f), C.byref(ms)), "chm_dmrg_heff_stats")
        return f.value, ms.value

    def set_heff_knobs(self, explicit_max=-1, item_terms=0, group_terms=0, chain_mfma=-1, fchain_min=-1):
        self._check(self.lib.chm_set_heff_knobs(self.h, int(explicit_max), int(item_terms), int(group_terms),
                                                int(chain_mfma), int(fchain_min)), "chm_set_heff_knobs")

    def set_tuning(self, rows):
        """Autotuner table: rows of (dmax_lo, dmax_hi, V_lo, V_hi, explicit_max, item_terms,
        group_terms, chain_mfma, fchain_min)."""
        a = np.ascontiguousarray(np.asarray(rows, dtype=np.int64).reshape(-1, 9))
        self._check(self.lib.chm_set_tuning(self.h, a.shape[0], a.ctypes.data if a.size else None), "chm_set_tuning")

    def tuning_used(self):
        out = np.zeros(7, dtype=np.int64)
        self._check(self.lib.chm_tuning_used(self.h, out.ctypes.data), "chm_tuning_used")
        return dict(zip(["row", "dmax", "explicit_max", "item_terms", "group_terms", "chain_mfma", "fchain_min"],
                        out.tolist()))

    def site_log(self):
        """Per two-site step since the last read: list of dicts (site, energy, products,
        seconds, F_H, V, heff_ms); clears the records."""
        n = C.c_int(0)
        self._check(self.lib.chm_dmrg_site_log(self.h, C.byref(n), None, None, None), "chm_dmrg_site_log")
        sites = np.zeros(n.value, dtype=np.int32)
        prods = np.zeros(n.value, dtype=np.int32)
        vals = np.zeros(5 * n.value)
        if n.value:
            self._check(self.lib.chm_dmrg_site_log(self.h, C.byref(n), sites.ctypes.data, prods.ctypes.data,
                                                   vals.ctypes.data), "chm_dmrg_site_log")
        v = vals.reshape(-1, 5)
        return [{"site": int(sites[q]), "energy": float(v[q, 0]), "products": int(prods[q]), "seconds": float(v[q, 1]),
                 "F_H": float(v[q, 2]), "V": int(v[q, 3]), "heff_ms": float(v[q, 4])} for q in range(n.value)]

    def dmrg_stats(self):
        t = (C.c_double * 6)()
        n = C.c_int64(0)
        self._check(self.lib.chm_dmrg_stats(self.h, t, C.byref(n)), "chm_dmrg_stats")
        return dict(zip(["join", "prepare", "davidson", "split", "update", "total"], list(t))), n.value

    def prepare(self, site):
        self._check(self.lib.chm_heff_prepare(self.h, int(site)), "chm_heff_prepare")
        v = C.c_int64(0)
        self._check(self.lib.chm_heff_veclen(self.h, C.byref(v)), "chm_heff_veclen")
        self.V = v.value
        return self.V

    def apply(self, x, y=None, f32=False):
        """y = Heff x; f32: with the hi floats of a site prepared in the mixed form."""
        fn, name = (self.lib.chm_heff_apply_f32, "chm_heff_apply_f32") if f32 else (self.lib.chm_heff_apply, "chm_heff_apply")
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, dtype=np.float64)
            y = np.zeros(self.V) if y is None else y
            self._check(fn(self.h, _ptr(x), _ptr(y), 0), name)
            return y
        if y is None:
            raise ValueError("a device x needs a device y of the same length")
        xp, yp = _dev(x, self.device, self.V, "x"), _dev(y, self.device, self.V, "y")
        self._check(fn(self.h, xp, yp, 1), name)
        return y

    def apply_noalloc(self, x, y):
        """y = Heff x through chm_heff_apply_noalloc (device tensors of length V; partial
        sums in the library's static device scratch, no device allocation)."""
        xp, yp = _dev(x, self.device, self.V, "x"), _dev(y, self.device, self.V, "y")
        self._check(self.lib.chm_heff_apply_noalloc(self.h, xp, yp), "chm_heff_apply_noalloc")
        return y

    def davidson_stats(self):
        """(deflations, iteration graphs captured, graph launches) of this context's Davidson solves."""
        s = (C.c_int64 * 3)()
        self._check(self.lib.chm_davidson_stats(self.h, s), "chm_davidson_stats")
        return tuple(s)

    def set_lower(self, veff):
        """Lower states' VeffTilde vectors for the prepared site (n_lower x V, host array
        or device tensor); None or an empty array removes them."""
        if veff is None or len(veff) == 0:
            self._check(self.lib.chm_heff_set_lower(self.h, 0, None, 0), "chm_heff_set_lower")
            return
        if isinstance(veff, np.ndarray):
            v = np.ascontiguousarray(np.atleast_2d(veff), dtype=np.float64)
            if v.shape[1] != self.V:
                raise ValueError(f"lower-state vectors of length {v.shape[1]}, the site has V = {self.V}")
            self._check(self.lib.chm_heff_set_lower(self.h, v.shape[0], _ptr(v), 0), "chm_heff_set_lower")
            return
        if veff.dim() != 2 or veff.shape[1] != self.V:
            raise ValueError("lower-state vectors must be an n_lower x V tensor")
        vp = _dev(veff, self.device, veff.shape[0] * self.V, "lower-state vectors")
        self._check(self.lib.chm_heff_set_lower(self.h, veff.shape[0], vp, 1), "chm_heff_set_lower")

    def set_heff_form(self, form):
        """FORM_FP64 or FORM_MIXED; applies from the next prepare."""
        self._check(self.lib.chm_set_heff_form(self.h, int(form)), "chm_set_heff_form")

    # ---- learned preconditioner
    def precond_mode(self, mode):
        self._check(self.lib.chm_precond_mode(self.h, int(mode)), "chm_precond_mode")

    def precond_fallbacks(self):
        """Davidson solves whose learned correction was switched off by the residual guard."""
        n = C.c_int64(0)
        self._check(self.lib.chm_precond_stats(self.h, C.byref(n)), "chm_precond_stats")
        return n.value

    def mlp_set(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.lib.chm_mlp_set(self.h, p.ctypes.data, p.size), "chm_mlp_set")

    def mlp_get(self):
        p = np.zeros(mlp_nparams(), dtype=np.float32)
        self._check(self.lib.chm_mlp_get(self.h, p.ctypes.data, p.size), "chm_mlp_get")
        return p

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_IN)
        out = np.zeros(len(X), dtype=np.float32)
        self._check(self.lib.chm_mlp_forward(self.h, X.ctypes.data, len(X), out.ctypes.data), "chm_mlp_forward")
        return out

    def mlp_loss_grad(self, X, p, e, w):
        X, p, e, w = _samples(X, p, e, w)
        g = np.zeros(mlp_nparams(), dtype=np.float32)
        loss = C.c_double(0.0)
        self._check(self.lib.chm_mlp_loss_grad(self.h, X.ctypes.data, p.ctypes.data, e.ctypes.data, w.ctypes.data, len(p),
                                               g.ctypes.data, C.byref(loss)), "chm_mlp_loss_grad")
        return loss.value, g

    def mlp_train(self, X, p, e, w, batch=4096, epochs=4, lr=1e-3, seed=1):
        X, p, e, w = _samples(X, p, e, w)
        l0, l1 = C.c_double(0.0), C.c_double(0.0)
        self._check(self.lib.chm_mlp_train(self.h, X.ctypes.data, p.ctypes.data, e.ctypes.data, w.ctypes.data, len(p),
                                           int(batch), int(epochs), float(lr), int(seed), C.byref(l0), C.byref(l1)),
                    "chm_mlp_train")
        return l0.value, l1.value

    def harvest(self, per_iteration):
        self._check(self.lib.chm_harvest(self.h, int(per_iteration)), "chm_harvest")

    def harvest_take(self):
        """(X, p, e, w) of the samples collected since the last take."""
        n = C.c_int64(0)
        self._check(self.lib.chm_harvest_take(self.h, C.byref(n), None, None, None, None), "chm_harvest_take")
        X = np.zeros((n.value, MLP_IN), dtype=np.float32)
        p, e, w = (np.zeros(n.value, dtype=np.float32) for _ in range(3))
        self._check(self.lib.chm_harvest_take(self.h, C.byref(n), X.ctypes.data, p.ctypes.data, e.ctypes.data,
                                              w.ctypes.data), "chm_harvest_take")
        return X, p, e, w

    def mixed_stats(self):
        lo, hi = C.c_int64(0), C.c_int64(0)
        self._check(self.lib.chm_mixed_stats(self.h, C.byref(lo), C.byref(hi)), "chm_mixed_stats")
        return lo.value, hi.value

    def diag(self):
        d = np.zeros(self.V)
        self._check(self.lib.chm_heff_diag(self.h, _ptr(d), 0), "chm_heff_diag")
        return d

    def davidson(self, x, rtol):
        """x: program-convention guess (numpy, updated in place). Returns (eigenvalue, n_matvec)."""
        e = C.c_double(0.0)
        nm = C.c_int(0)
        if isinstance(x, np.ndarray):
            if x.dtype != np.float64 or not x.flags.c_contiguous or x.size != self.V:
                raise ValueError(f"x must be a contiguous float64 array of length V = {self.V}")
            on_dev, xp = 0, _ptr(x)
        else:
            on_dev, xp = 1, _dev(x, self.device, self.V, "x")
        self._check(self.lib.chm_davidson(self.h, xp, on_dev, float(rtol), C.byref(e), C.byref(nm)), "chm_davidson")
        return e.value, nm.value

    def flops(self):
        f, b = C.c_double(0), C.c_double(0)
        self._check(self.lib.chm_flops(self.h, C.byref(f), C.byref(b)), "chm_flops")
        return f.value, b.value

    def plan_stats(self):
        s = (C.c_int64 * 16)()
        self._check(self.lib.chm_plan_stats(self.h, s), "chm_plan_stats")
        keys = ["terms", "combos", "combo_doubles", "W_doubles", "tiles1", "tiles2", "explicit_pairs", "chain_terms",
                "blocks", "V", "diag_terms", "explicit_doubles", "explicit_cols", "explicit_threshold", "explicit_items", "chain_kernel_terms"]
        return dict(zip(keys, list(s)))

    def stream(self):
        p = C.c_void_p()
        self._check(self.lib.chm_stream(self.h, C.byref(p)), "chm_stream")
        return p.value

    def timing(self, enable):
        self._check(self.lib.chm_timing_enable(self.h, int(enable)), "chm_timing_enable")

    def timing_read(self):
        """Accumulated milliseconds per kernel (chain phase 1, explicit product, chain
        phase 2) and the number of timed applications."""
        ms = (C.c_double * 3)()
        n = C.c_int64(0)
        self._check(self.lib.chm_timing_read(self.h, ms, C.byref(n)), "chm_timing_read")
        return list(ms), n.value


class Heff:
    """Mirror of CheMPS2::Heff (Heff.h:50-89) for one prepared two-site step."""

    def __init__(self, ctx, site, dvdson_rtol=1e-5):
        self.ctx = ctx
        self.site = site
        self.rtol = dvdson_rtol
        self.V = ctx.prepare(site)

    def makeHeff(self, x):
        """y = Heff x in the symmetric convention (Heff.cpp:43-248)."""
        return self.ctx.apply(x)

    def fillHeffDiag(self):
        """diag(Heff) (Heff.cpp:250-315)."""
        return self.ctx.diag()

    def SolveDAVIDSON(self, s_storage):
        """Lowest eigenpair from the Sobject storage (program convention, updated in
        place); returns the eigenvalue (Heff.cpp:330-386)."""
        e, self.n_matvec = self.ctx.davidson(s_storage, self.rtol)
        return e
