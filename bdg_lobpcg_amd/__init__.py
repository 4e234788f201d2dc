"""bdg_lobpcg_amd: MI355X (gfx950) BdG iLOBPCG hot path.

The product is the in-tree shared library ``libbdg_amd.so`` (HIP kernels +
host C): the reference's public C API (``include/bdg/bdg.h``), its
extensions (``include/bdg/bdg_amd.h``) and the device C ABI
(``include/bdgx.h``).  This module only binds it with ctypes so that Python
tests, the smoke check and the benchmark can drive it; it never computes
anything itself and raises if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbdg_amd.so")
_lib = None

# operator codes (include/bdgx.h)
OP = {"kinetic": 0, "K": 1, "M": 2, "Lrep": 3, "Swap": 4, "precondK": 5,
      "precondM": 6, "precondLrep": 7, "dipolar_conv": 8}
FP64, FP32 = 0, 1
FORM_FIRST_ORDER, FORM_PRODUCT = 0, 1
LP_NKNOT, LP_HIDDEN = 16, 16
LP_NPARAM = 8 + 2 * LP_NKNOT + 29 * LP_HIDDEN
PREC_FP64, PREC_MIXED = 0, 1
INIT_PLANEWAVE, INIT_WF_WEIGHTED, INIT_REUSE, INIT_CUSTOM = 0, 1, 2, 3
GEOM_AUTO, GEOM_ELONGATED, GEOM_RING = 0, 1, 2

V_TRAP_FN = C.CFUNCTYPE(C.c_double, C.c_uint64, C.POINTER(C.c_double), C.c_void_p)
U_INT_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double)
INIT_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p)


class BdgError(RuntimeError):
    pass


class Graph:
    """A captured device chain of a DeviceContext (bdgx_graph)."""

    def __init__(self, ctx, handle):
        self.ctx, self.h = ctx, handle

    def launch(self):
        check(lib().bdgx_graph_launch(self.ctx.h, self.h), "bdgx_graph_launch")

    def close(self):
        if self.h:
            lib().bdgx_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_uint64), ("column_applies", C.c_uint64),
                ("column_applies_fp32", C.c_uint64), ("image_refreshes", C.c_uint64),
                ("restarts", C.c_uint64), ("learned_iterations", C.c_uint64),
                ("learned_fallback_at", C.c_int), ("graph_launches", C.c_uint64),
                ("graph_captures", C.c_uint64), ("seconds_total", C.c_double),
                ("seconds_rr", C.c_double), ("status", C.c_int),
                ("frozen_skips", C.c_uint64), ("iterations_fp32", C.c_uint64),
                ("fp32_worst_residual", C.c_double), ("fp64_sweep_only", C.c_int),
                ("fp64_group_size", C.c_int)]


def build() -> None:
    """Compile libbdg_amd.so for gfx950 with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "-j8"], check=True)


def lib():
    """Load the native library; raise loudly when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BdgError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P, U64, I, D = C.c_void_p, C.c_uint64, C.c_int, C.c_double
    sig = {
        # device ABI
        "bdgx_ctx_create": (I, [C.POINTER(P), U64, P, P, I]),
        "bdgx_ctx_destroy": (None, [P]),
        "bdgx_ctx_size": (U64, [P]),
        "bdgx_ctx_ksize": (U64, [P]),
        "bdgx_error_string": (C.c_char_p, [I]),
        "bdgx_set_fields": (I, [P, P, P, P, P, P, D]),
        "bdgx_derive_precond": (I, [P, D]),
        "bdgx_dipolar_set_kernel": (I, [P, D, P, D]),
        "bdgx_dipolar_get_kernel": (I, [P, P]),
        "bdgx_dipolar_meanfield": (I, [P, P]),
        "bdgx_apply": (I, [P, I, I, P, P, U64, U64]),
        "bdgx_apply_host": (I, [P, I, P, P, U64, U64]),
        "bdgx_gram": (I, [P, I, P, U64, U64, P, U64, U64, U64, I, P]),
        "bdgx_gram_noalloc": (I, [P, I, P, U64, U64, P, U64, U64, U64, I, P]),
        "bdgx_combine": (I, [P, I, P, U64, U64, P, P, U64, U64, U64]),
        "bdgx_combine_cols": (I, [P, I, P, U64, U64, P, P, U64, P, U64, U64]),
        "bdgx_residual": (I, [P, I, P, P, P, U64, U64, U64, P, I, P, U64, P]),
        "bdgx_convert": (I, [P, I, P, U64, I, P, U64, U64, U64]),
        "bdgx_fill_uniform_xorshift": (I, [P, P, U64, U64, U64, U64, C.c_uint32]),
        "bdgx_gram_pair": (I, [P, I, P, U64, U64, P, U64, P, U64, U64, U64, P]),
        "bdgx_itp_groundstate": (I, [P, P, D, D, P, D, D, U64, P, P]),
        "bdgx_malloc": (I, [C.POINTER(P), C.c_size_t]),
        "bdgx_free": (I, [P]),
        "bdgx_memcpy_h2d": (I, [P, P, P, C.c_size_t]),
        "bdgx_memcpy_d2h": (I, [P, P, P, C.c_size_t]),
        "bdgx_memcpy_d2h_pair": (I, [P, P, P, P, P, C.c_size_t]),
        "bdgx_memset": (I, [P, P, I, C.c_size_t]),
        "bdgx_sync": (I, [P]),
        "bdgx_stream": (P, [P]),
        "bdgx_timer_begin": (I, [P]),
        "bdgx_timer_end": (I, [P, C.POINTER(C.c_float)]),
        "bdgx_workspace_bytes": (U64, [P]),
        "bdgx_bench_pass": (I, [P, I, I, U64, I, C.POINTER(C.c_float), C.POINTER(D)]),
        "bdgx_autotune": (I, [P, I, U64, I]),
        "bdgx_get_launch_variant": (I, [P, I, I, C.POINTER(I)]),
        "bdgx_set_launch_variant": (I, [P, I, I, I]),
        "bdgx_lp_set_params": (I, [P, P]),
        "bdgx_lp_apply": (I, [P, I, P, P, U64, U64]),
        "bdgx_lp_backward": (I, [P, P, P, U64, U64, P]),
        "bdgx_lp_apply_layer": (I, [P, I, P, P, U64, U64]),
        "bdgx_mem_info": (I, [P, P]),
        "bdgx_lp_backward_layer": (I, [P, I, P, P, U64, U64, P]),
        # public API
        "bdg_alloc": (P, [U64, P, P, I]),
        "bdg_free": (None, [C.POINTER(P)]),
        "bdg_reset": (None, [P]),
        "bdg_set_system": (None, [P]),
        "bdg_set_trap": (None, [P, P, P]),
        "bdg_set_wavefunction": (None, [P, P, U64]),
        "bdg_set_local_interactions": (None, [P, P, P, P]),
        "bdg_set_mu": (None, [P, D]),
        "bdg_set_dipolar": (None, [P, D, P, D]),
        "bdg_set_solver_params": (None, [P, U64, U64, U64, D]),
        "bdg_set_init_mode": (None, [P, I, P, P]),
        "bdg_reuse_modes": (I, [P, D]),
        "bdg_solve": (I, [P]),
        "bdg_converged": (U64, [P]),
        "bdg_eigenvalues": (C.POINTER(D), [P]),
        "bdg_modes_u": (P, [P]),
        "bdg_modes_v": (P, [P]),
        "bdg_load_wavefunction": (I, [P, C.c_char_p]),
        "bdg_save_wavefunction_bin": (I, [P, C.c_char_p]),
        "bdg_load_wavefunction_bin": (I, [P, C.c_char_p]),
        "bdg_write_modes_bin": (I, [P, C.c_char_p]),
        "bdg_write_eigenvalues": (I, [P, C.c_char_p]),
        "bdg_write_mode_u": (I, [P, U64, C.c_char_p]),
        "bdg_write_mode_v": (I, [P, U64, C.c_char_p]),
        # extensions
        "bdg_set_precision": (None, [P, I]),
        "bdg_get_solve_stats": (None, [P, C.POINTER(SolveStats)]),
        "bdg_set_formulation": (None, [P, I]),
        "bdg_set_fp32_phase": (None, [P, D, U64]),
        "bdg_set_locking": (None, [P, I]),
        "bdg_set_lock_guards": (None, [P, I]),
        "bdg_set_device_sweeps": (None, [P, I]),
        "bdg_set_fp64_sweep_only": (None, [P, I]),
        "bdg_set_fp64_groups": (None, [P, I]),
        "bdg_residuals": (C.POINTER(D), [P]),
        "bdg_zero_modes": (C.POINTER(I), [P]),
        "bdg_device_context": (P, [P]),
        "bdg_local_term_K": (C.POINTER(D), [P]),
        "bdg_local_term_M": (C.POINTER(D), [P]),
        "bdg_precond_weight_K": (C.POINTER(D), [P]),
        "bdg_precond_weight_M": (C.POINTER(D), [P]),
        "bdg_mu": (D, [P]),
        "bdg_set_mixed_params": (None, [P, D, U64]),
        "bdg_set_verbosity": (None, [P, I]),
        "bdg_set_graph_capture": (None, [P, I]),
        "bdg_set_freeze": (None, [P, D]),
        "bdg_solve_batch": (I, [C.POINTER(P), I, I, C.POINTER(I)]),
        "bdgx_graph_begin": (I, [P]),
        "bdgx_graph_end": (I, [P, C.POINTER(P)]),
        "bdgx_graph_launch": (I, [P, P]),
        "bdgx_graph_destroy": (None, [P]),
        "bdgx_reserve": (I, [P, U64, U64]),
        "bdgx_set_chunk_bytes": (I, [P, U64]),
        "bdg_set_learned_preconditioner": (None, [P, P]),
        "bdg_set_learned_fallback": (None, [P, I, D]),
        "bdg_start_block": (None, [P, P]),
        "bdg_load_wavefunction_fmt": (I, [P, C.c_char_p]),
        "bdgx_reset_fields": (I, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def error_string(rc: int) -> str:
    """Text of a library status code."""
    return lib().bdgx_error_string(rc).decode() if rc > -100 else "rayleigh-ritz failure"


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        raise BdgError(f"{what}: error {rc} ({error_string(rc)})")


# ---------------------------------------------------------------- device --
class DeviceBuffer:
    """Raw device allocation owned by Python (freed on close/GC)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        check(lib().bdgx_malloc(C.byref(p), int(nbytes)), "bdgx_malloc")
        self.ptr = p
        self.nbytes = int(nbytes)

    def close(self):
        if getattr(self, "ptr", None):
            lib().bdgx_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


class DeviceContext:
    """bdgx context: grid, FFT plans, physics fields, HIP stream."""

    def __init__(self, N, L, complex_psi0: bool = False):
        self.N = tuple(int(n) for n in N)
        self.L = tuple(float(v) for v in L)
        self.cplx = bool(complex_psi0)
        h = C.c_void_p()
        Na = np.ascontiguousarray(self.N, dtype=np.uint64)
        La = np.ascontiguousarray(self.L, dtype=np.float64)
        check(lib().bdgx_ctx_create(C.byref(h), len(self.N), _p(Na), _p(La), int(self.cplx)),
              "bdgx_ctx_create")
        self.h = h
        self._owned = True
        self._init_sizes()

    def _init_sizes(self):
        self.size = int(lib().bdgx_ctx_size(self.h))
        self.k_size = int(lib().bdgx_ctx_ksize(self.h))
        self.dtype64 = np.complex128 if self.cplx else np.float64
        self.dtype32 = np.complex64 if self.cplx else np.float32

    @classmethod
    def wrap(cls, handle, N, L, complex_psi0=False):
        """Non-owning view of an existing context (e.g. a Problem's)."""
        self = cls.__new__(cls)
        self.N = tuple(int(n) for n in N)
        self.L = tuple(float(v) for v in L)
        self.cplx = bool(complex_psi0)
        self.h = C.c_void_p(handle)
        self._owned = False
        self._init_sizes()
        return self

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                lib().bdgx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def dtype(self, prec):
        return self.dtype64 if prec == FP64 else self.dtype32

    def set_fields(self, ltK=None, ltM=None, wf=None, pK=None, pM=None, mu=0.0):
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return _p(a)

        check(lib().bdgx_set_fields(self.h, arr(ltK, np.float64), arr(ltM, np.float64),
                                    arr(wf, self.dtype64), arr(pK, np.float64),
                                    arr(pM, np.float64), float(mu)), "bdgx_set_fields")

    def derive_precond(self, floor):
        check(lib().bdgx_derive_precond(self.h, float(floor)), "bdgx_derive_precond")

    def set_dipolar(self, g_ddi, direction, cutoff):
        d = np.ascontiguousarray(direction, dtype=np.float64)
        check(lib().bdgx_dipolar_set_kernel(self.h, float(g_ddi), _p(d), float(cutoff)),
              "bdgx_dipolar_set_kernel")

    def dipolar_kernel(self):
        out = np.zeros(self.k_size)
        check(lib().bdgx_dipolar_get_kernel(self.h, _p(out)), "bdgx_dipolar_get_kernel")
        return out

    def dipolar_meanfield(self):
        out = np.zeros(self.size)
        check(lib().bdgx_dipolar_meanfield(self.h, _p(out)), "bdgx_dipolar_meanfield")
        return out

    def upload(self, a: np.ndarray) -> DeviceBuffer:
        a = np.asfortranarray(a)
        b = DeviceBuffer(a.nbytes)
        check(lib().bdgx_memcpy_h2d(self.h, b.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes), "h2d")
        return b

    def fill_uniform(self, nrows: int, ncols: int, steps_per_col: int, seed: int = 42) -> np.ndarray:
        """bdgx_fill_uniform_xorshift into a fresh device block, downloaded."""
        dt = np.complex128 if self.cplx else np.float64
        b = DeviceBuffer(np.dtype(dt).itemsize * nrows * ncols)
        check(lib().bdgx_fill_uniform_xorshift(self.h, b.ptr, nrows, nrows, ncols, steps_per_col,
                                               seed), "bdgx_fill_uniform_xorshift")
        return self.download(b, (nrows, ncols), dt)

    def download(self, b: DeviceBuffer, shape, dtype) -> np.ndarray:
        out = np.zeros(shape, dtype=dtype, order="F")
        check(lib().bdgx_memcpy_d2h(self.h, out.ctypes.data_as(C.c_void_p), b.ptr, out.nbytes), "d2h")
        return out

    def apply(self, op: str, X: np.ndarray, prec: int = FP64) -> np.ndarray:
        """Upload a column block, apply `op` on the device, download."""
        X = np.asarray(X)
        one = X.ndim == 1
        Xf = np.asfortranarray(X[:, None] if one else X, dtype=self.dtype(prec))
        dX = self.upload(Xf)
        dY = DeviceBuffer(Xf.nbytes)
        check(lib().bdgx_memset(self.h, dY.ptr, 0, Xf.nbytes), "memset")
        check(lib().bdgx_apply(self.h, OP[op], prec, dX.ptr, dY.ptr, Xf.shape[1], Xf.shape[0]),
              f"bdgx_apply({op})")
        Y = self.download(dY, Xf.shape, Xf.dtype)
        return Y[:, 0] if one else Y

    def apply_device(self, op, prec, dX, dY, ncols, ld):
        check(lib().bdgx_apply(self.h, OP[op], prec, dX.ptr, dY.ptr, int(ncols), int(ld)),
              f"bdgx_apply({op})")

    def sync(self):
        check(lib().bdgx_sync(self.h), "sync")

    def timer_begin(self):
        check(lib().bdgx_timer_begin(self.h), "timer")

    def timer_end(self) -> float:
        ms = C.c_float()
        check(lib().bdgx_timer_end(self.h, C.byref(ms)), "timer")
        return float(ms.value)

    def autotune(self, prec: int, nbatch: int = 32, force: bool = False):
        """Time the launch variants of the fixed-size FFT kernels and keep the
        fastest (bdgx_autotune); returns the chosen variants."""
        check(lib().bdgx_autotune(self.h, int(prec), int(nbatch), int(force)), "bdgx_autotune")
        return self.launch_variants(prec)

    def launch_variants(self, prec: int) -> dict:
        out = {}
        for kind, name in enumerate(("plane_fwd", "plane_inv", "z", "combine", "gram")):
            v = C.c_int()
            check(lib().bdgx_get_launch_variant(self.h, kind, int(prec), C.byref(v)),
                  "bdgx_get_launch_variant")
            out[name] = int(v.value)
        return out

    def set_launch_variant(self, kind: int, prec: int, variant: int):
        check(lib().bdgx_set_launch_variant(self.h, int(kind), int(prec), int(variant)),
              "bdgx_set_launch_variant")

    def capture(self, fn):
        """Record the device work that fn() issues on this context into a
        graph (bdgx_graph_begin / _end); returns a Graph (launch(), close())."""
        check(lib().bdgx_graph_begin(self.h), "bdgx_graph_begin")
        try:
            fn()
        except Exception:
            lib().bdgx_graph_end(self.h, None)
            raise
        g = C.c_void_p()
        check(lib().bdgx_graph_end(self.h, C.byref(g)), "bdgx_graph_end")
        return Graph(self, g)

    def bench_pass(self, pass_id: int, prec: int, nbatch: int, iters: int = 20):
        """(ms per launch, algorithmic bytes per launch) of one FFT pass kind."""
        ms = C.c_float()
        by = C.c_double()
        check(lib().bdgx_bench_pass(self.h, int(pass_id), int(prec), int(nbatch), int(iters),
                                    C.byref(ms), C.byref(by)), "bdgx_bench_pass")
        return float(ms.value), float(by.value)

    def gram(self, A: np.ndarray, B: np.ndarray, prec=FP64, swap_b=False) -> np.ndarray:
        A = np.asfortranarray(A, dtype=self.dtype(prec))
        B = np.asfortranarray(B, dtype=self.dtype(prec))
        dA, dB = self.upload(A), self.upload(B)
        w = 2 if self.cplx else 1
        dG = DeviceBuffer(8 * w * A.shape[1] * B.shape[1])
        check(lib().bdgx_gram(self.h, prec, dA.ptr, A.shape[0], A.shape[1], dB.ptr, B.shape[0],
                              B.shape[1], A.shape[0], int(swap_b), dG.ptr), "gram")
        G = self.download(dG, (A.shape[1], B.shape[1]), self.dtype64)
        return G

    def gram_noalloc_device(self, prec, dA, lda, p, dB, ldb, q, nrows, swap_b, dG) -> int:
        """bdgx_gram_noalloc on device buffers; returns the status code."""
        return int(lib().bdgx_gram_noalloc(self.h, prec, dA.ptr, lda, p, dB.ptr, ldb, q, nrows,
                                           int(swap_b), dG.ptr))

    def workspace_bytes(self) -> int:
        return int(lib().bdgx_workspace_bytes(self.h))

    def gram_pair(self, A: np.ndarray, B1: np.ndarray, B2: np.ndarray, prec=FP64) -> np.ndarray:
        """bdgx_gram_pair: A^H [B1 | B2] with A read once (real contexts)."""
        A, B1, B2 = (np.asfortranarray(x, dtype=self.dtype(prec)) for x in (A, B1, B2))
        dA, d1, d2 = self.upload(A), self.upload(B1), self.upload(B2)
        q = B1.shape[1]
        dG = DeviceBuffer(8 * A.shape[1] * 2 * q)
        check(lib().bdgx_gram_pair(self.h, prec, dA.ptr, A.shape[0], A.shape[1], d1.ptr,
                                   B1.shape[0], d2.ptr, B2.shape[0], q, A.shape[0], dG.ptr),
              "gram_pair")
        return self.download(dG, (A.shape[1], 2 * q), np.float64)

    def combine(self, A: np.ndarray, Cm: np.ndarray, prec=FP64) -> np.ndarray:
        A = np.asfortranarray(A, dtype=self.dtype(prec))
        Cm = np.asfortranarray(Cm, dtype=self.dtype64)
        dA, dC = self.upload(A), self.upload(Cm)
        dY = DeviceBuffer(A.shape[0] * Cm.shape[1] * A.itemsize)
        check(lib().bdgx_combine(self.h, prec, dA.ptr, A.shape[0], A.shape[1], dC.ptr, dY.ptr,
                                 A.shape[0], Cm.shape[1], A.shape[0]), "combine")
        return self.download(dY, (A.shape[0], Cm.shape[1]), A.dtype)

    def combine_cols(self, A: np.ndarray, Cm: np.ndarray, cols, Y0: np.ndarray) -> np.ndarray:
        """Y = Y0 with Y[:, cols[j]] = A @ Cm[:, j] (FP64, at most 8 columns)."""
        A = np.asfortranarray(A, dtype=self.dtype64)
        Cm = np.asfortranarray(Cm, dtype=self.dtype64)
        Y0 = np.asfortranarray(Y0, dtype=self.dtype64)
        idx = np.ascontiguousarray(cols, dtype=np.int32)
        dA, dC, dY = self.upload(A), self.upload(Cm), self.upload(Y0)
        check(lib().bdgx_combine_cols(self.h, FP64, dA.ptr, A.shape[0], A.shape[1], dC.ptr,
                                      dY.ptr, Y0.shape[0], idx.ctypes.data, idx.size,
                                      A.shape[0]), "combine_cols")
        return self.download(dY, Y0.shape, Y0.dtype)

    # -- learned preconditioner (FP32, real psi0) ---------------------------
    def lp_set_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        if th.size != LP_NPARAM:
            raise ValueError(f"theta must have {LP_NPARAM} entries")
        check(lib().bdgx_lp_set_params(self.h, th.ctypes.data_as(C.c_void_p)), "bdgx_lp_set_params")

    def lp_apply_device(self, dR, dW, ncols):
        check(lib().bdgx_lp_apply(self.h, FP32, dR.ptr, dW.ptr, int(ncols), self.size),
              "bdgx_lp_apply")

    def lp_backward_device(self, dR, dG, ncols) -> np.ndarray:
        g = np.zeros(LP_NPARAM, dtype=np.float32)
        check(lib().bdgx_lp_backward(self.h, dR.ptr, dG.ptr, int(ncols), self.size,
                                     g.ctypes.data_as(C.c_void_p)), "bdgx_lp_backward")
        return g

    def lp_apply_layer_device(self, layer, dR, dW, ncols, r_off=0, w_off=0):
        """dW = T'_layer dR (FP32 columns; offsets in columns)"""
        check(lib().bdgx_lp_apply_layer(self.h, int(layer), dR.ptr.value + 4 * self.size * r_off,
                                        dW.ptr.value + 4 * self.size * w_off, int(ncols),
                                        self.size), "bdgx_lp_apply_layer")

    def lp_backward_layer_device(self, layer, dR, dG, ncols, g_off=0) -> np.ndarray:
        g = np.zeros(LP_NPARAM, dtype=np.float32)
        check(lib().bdgx_lp_backward_layer(self.h, int(layer), dR.ptr,
                                           dG.ptr.value + 4 * self.size * g_off, int(ncols),
                                           self.size, g.ctypes.data_as(C.c_void_p)),
              "bdgx_lp_backward_layer")
        return g

    def lp_apply_layer(self, layer, R: np.ndarray) -> np.ndarray:
        R = np.asfortranarray(R, dtype=np.float32)
        dR = self.upload(R)
        dW = DeviceBuffer(R.nbytes)
        self.lp_apply_layer_device(layer, dR, dW, R.shape[1])
        return self.download(dW, R.shape, np.float32)

    def lp_apply(self, R: np.ndarray) -> np.ndarray:
        R = np.asfortranarray(R, dtype=np.float32)
        dR = self.upload(R)
        dW = DeviceBuffer(R.nbytes)
        self.lp_apply_device(dR, dW, R.shape[1])
        return self.download(dW, R.shape, np.float32)

    def lp_backward(self, R: np.ndarray, G: np.ndarray) -> np.ndarray:
        R = np.asfortranarray(R, dtype=np.float32)
        G = np.asfortranarray(G, dtype=np.float32)
        dR, dG = self.upload(R), self.upload(G)
        return self.lp_backward_device(dR, dG, R.shape[1])

    def itp_groundstate(self, psi, omega, g, g_dd, direction, cutoff, dt, nsteps):
        psi = np.ascontiguousarray(psi, dtype=self.dtype64).copy()
        om = np.ascontiguousarray(omega, dtype=np.float64)
        d = np.ascontiguousarray(direction, dtype=np.float64)
        mu = C.c_double(0.0)
        check(lib().bdgx_itp_groundstate(self.h, _p(om), float(g), float(g_dd), _p(d),
                                         float(cutoff), float(dt), int(nsteps), _p(psi),
                                         C.cast(C.pointer(mu), C.c_void_p)), "itp")
        return psi, mu.value


# --------------------------------------------------------------- problem --
class Problem:
    """The reference's bdg_t lifecycle (include/bdg/bdg.h) driven from Python."""

    def __init__(self, N, L, complex_psi0: bool = False):
        self.N = tuple(int(n) for n in N)
        self.L = tuple(float(v) for v in L)
        self.cplx = bool(complex_psi0)
        Na = np.ascontiguousarray(self.N, dtype=np.uint64)
        La = np.ascontiguousarray(self.L, dtype=np.float64)
        self.h = C.c_void_p(lib().bdg_alloc(len(self.N), _p(Na), _p(La), int(self.cplx)))
        self.size = int(np.prod(self.N))
        self._keep = []

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().bdg_free(C.byref(self.h))
            self.h = None

    def __del__(self):
        self.close()

    def reset(self):
        """bdg_reset: physics cleared, grid / device context / start-block
        strategy (and a reuse block) kept."""
        lib().bdg_reset(self.h)

    # setup, in the reference's order
    def set_system(self):
        lib().bdg_set_system(self.h)

    def set_trap_harmonic(self, omega):
        om = np.ascontiguousarray(omega, dtype=np.float64)
        self._keep.append(om)
        fn = C.cast(lib().bdg_cb_harmonic_trap, C.c_void_p)
        lib().bdg_set_trap(self.h, fn, _p(om))

    def set_trap(self, fn_py):
        cb = V_TRAP_FN(lambda dim, r, p: fn_py([r[i] for i in range(dim)]))
        self._keep.append(cb)
        lib().bdg_set_trap(self.h, C.cast(cb, C.c_void_p), None)

    def set_wavefunction(self, wf):
        a = np.ascontiguousarray(wf, dtype=np.complex128 if self.cplx else np.float64)
        lib().bdg_set_wavefunction(self.h, _p(a), a.size)

    def set_contact(self, gK, gM=None):
        gK = np.array([float(gK)])
        gM = np.array([float(gK[0] if gM is None else gM)])
        self._keep += [gK, gM]
        fn = C.cast(lib().bdg_cb_contact, C.c_void_p)
        # U_intK and U_intM share the callback; the parameter selects g
        self._keep.append(fn)
        # two different parameters are needed: use Python trampolines only if
        # gK != gM, else the native callback with one parameter
        if gK[0] == gM[0]:
            lib().bdg_set_local_interactions(self.h, fn, fn, _p(gK))
        else:
            cbK = U_INT_FN(lambda p, n: gK[0] * n)
            cbM = U_INT_FN(lambda p, n: gM[0] * n)
            self._keep += [cbK, cbM]
            lib().bdg_set_local_interactions(self.h, C.cast(cbK, C.c_void_p),
                                             C.cast(cbM, C.c_void_p), None)

    def set_local_interactions(self, fK, fM):
        cbK = U_INT_FN(lambda p, n: fK(n))
        cbM = U_INT_FN(lambda p, n: fM(n))
        self._keep += [cbK, cbM]
        lib().bdg_set_local_interactions(self.h, C.cast(cbK, C.c_void_p),
                                         C.cast(cbM, C.c_void_p), None)

    def set_dipolar(self, g_ddi, direction, cutoff):
        d = np.ascontiguousarray(direction, dtype=np.float64)
        lib().bdg_set_dipolar(self.h, float(g_ddi), _p(d), float(cutoff))

    def set_mu(self, mu):
        lib().bdg_set_mu(self.h, float(mu))

    def set_solver_params(self, nev, sizeSub, maxIter, tol):
        lib().bdg_set_solver_params(self.h, int(nev), int(sizeSub), int(maxIter), float(tol))

    def set_precision(self, mode):
        lib().bdg_set_precision(self.h, int(mode))

    def set_mixed_params(self, img_tol=0.0, refresh_every=20):
        lib().bdg_set_mixed_params(self.h, float(img_tol), int(refresh_every))

    def set_formulation(self, form):
        """FORM_FIRST_ORDER (default) or FORM_PRODUCT."""
        lib().bdg_set_formulation(self.h, int(form))

    def set_fp32_phase(self, fp32_tol=1e-5, stall_window=20):
        lib().bdg_set_fp32_phase(self.h, float(fp32_tol), int(stall_window))

    def set_locking(self, max_active=-1):
        """Lock converged pairs once at most max_active columns are active
        (-1 automatic, 0 never)."""
        lib().bdg_set_locking(self.h, int(max_active))

    def set_lock_guards(self, guards=4):
        """Guard columns carried by the small block after locking (first-order
        form): Ritz pairs just beyond the block, iterated, never returned."""
        lib().bdg_set_lock_guards(self.h, int(guards))

    def set_device_sweeps(self, on: bool):
        """Keep the modes on the device after a solve so that reuse_modes
        builds the next start block there (default on)."""
        lib().bdg_set_device_sweeps(self.h, int(bool(on)))

    def set_fp64_sweep_only(self, on: bool):
        """Mixed first-order solves: the FP64 phase is one refinement sweep
        (automatic where the FP64 basis does not fit in device memory)."""
        lib().bdg_set_fp64_sweep_only(self.h, int(bool(on)))

    def set_fp64_groups(self, group: int):
        """Mixed first-order solves: FP64 phase in groups of `group` modes
        (automatic where the block's FP64 basis does not fit the device)."""
        lib().bdg_set_fp64_groups(self.h, int(group))

    def reuse_modes(self, noise_frac: float) -> int:
        return int(lib().bdg_reuse_modes(self.h, float(noise_frac)))

    def save_wavefunction_bin(self, path) -> int:
        return int(lib().bdg_save_wavefunction_bin(self.h, str(path).encode()))

    def load_wavefunction_bin(self, path) -> int:
        return int(lib().bdg_load_wavefunction_bin(self.h, str(path).encode()))

    def write_modes_bin(self, path) -> int:
        return int(lib().bdg_write_modes_bin(self.h, str(path).encode()))

    def set_verbosity(self, level):
        lib().bdg_set_verbosity(self.h, int(level))

    def set_graph_capture(self, on: bool):
        lib().bdg_set_graph_capture(self.h, int(bool(on)))

    def set_freeze(self, tol: float):
        """Converged columns moved by less than tol keep their vectors (0: off)."""
        lib().bdg_set_freeze(self.h, float(tol))

    def set_learned_fallback(self, window=30, factor=0.5):
        lib().bdg_set_learned_fallback(self.h, int(window), float(factor))

    def set_learned(self, theta):
        """Learned preconditioner parameters (LP_NPARAM floats), or None."""
        if theta is None:
            lib().bdg_set_learned_preconditioner(self.h, None)
            return
        th = np.ascontiguousarray(theta, dtype=np.float32)
        if th.size != LP_NPARAM:
            raise ValueError(f"theta must have {LP_NPARAM} entries")
        lib().bdg_set_learned_preconditioner(self.h, th.ctypes.data_as(C.c_void_p))

    def set_init_custom_block(self, X: np.ndarray):
        """BDG_INIT_CUSTOM with a fixed n x sizeSub block (copied in the callback)."""
        Xf = np.asfortranarray(X, dtype=np.complex128 if self.cplx else np.float64)

        def fill(bdg, ptr, n, m, param):
            C.memmove(ptr, Xf.ctypes.data, Xf.nbytes)

        cb = INIT_FN(fill)
        self._keep += [Xf, cb]
        lib().bdg_set_init_mode(self.h, INIT_CUSTOM, C.cast(cb, C.c_void_p), None)

    def set_init_uniform(self, seed: int = 42, host: bool = False):
        """BDG_INIT_CUSTOM with the built-in xorshift32 uniform block (u = v).
        bdg_solve generates this block on the device; host=True passes the
        same C function through a Python callback instead, which bdg_solve
        runs on the host like any user callback (same values)."""
        s = np.array([seed], dtype=np.uint32)
        self._keep.append(s)
        if not host:
            lib().bdg_set_init_mode(self.h, INIT_CUSTOM,
                                    C.cast(lib().bdg_init_uniform_xorshift, C.c_void_p), _p(s))
            return
        fn = C.cast(lib().bdg_init_uniform_xorshift, INIT_FN)

        def fill(bdg, ptr, n, m, param):
            fn(bdg, ptr, n, m, param)

        cb = INIT_FN(fill)
        self._keep += [fn, cb]
        lib().bdg_set_init_mode(self.h, INIT_CUSTOM, C.cast(cb, C.c_void_p), _p(s))

    def set_init_mode(self, mode, hint=None):
        lib().bdg_set_init_mode(self.h, int(mode), None,
                                C.c_void_p(int(hint)) if hint else None)

    def start_block(self, nrows_cols) -> np.ndarray:
        n, m = nrows_cols
        X = np.zeros((n, m), dtype=np.complex128 if self.cplx else np.float64, order="F")
        lib().bdg_start_block(self.h, _p(X))
        return X

    def solve(self) -> int:
        return int(lib().bdg_solve(self.h))

    @staticmethod
    def solve_batch(problems, nthreads: int = 0):
        """bdg_solve_batch: solve independent problems concurrently (one
        worker thread and HIP stream each); returns the per-problem codes."""
        n = len(problems)
        hs = (C.c_void_p * n)(*[p.h for p in problems])
        rc = (C.c_int * n)()
        lib().bdg_solve_batch(hs, n, int(nthreads), rc)
        return [int(x) for x in rc]

    # results
    def converged(self) -> int:
        return int(lib().bdg_converged(self.h))

    def nev_arr(self, ptr, n, dt=np.float64):
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt).copy()

    @staticmethod
    def _result(p, nev, what):
        if not p:
            raise RuntimeError(f"{what}: no results (the last solve did not finish)")
        return np.ctypeslib.as_array(p, shape=(nev,))

    def eigenvalues(self, nev) -> np.ndarray:
        return self._result(lib().bdg_eigenvalues(self.h), nev, "eigenvalues").copy()

    def residuals(self, nev) -> np.ndarray:
        return self._result(lib().bdg_residuals(self.h), nev, "residuals").copy()

    def zero_modes(self, nev) -> np.ndarray:
        return self._result(lib().bdg_zero_modes(self.h), nev, "zero_modes").astype(bool)

    def modes(self, nev):
        dt = np.complex128 if self.cplx else np.float64
        out = []
        for fn in (lib().bdg_modes_u, lib().bdg_modes_v):
            p = fn(self.h)
            a = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)),
                                      shape=(self.size * nev * (2 if self.cplx else 1),))
            out.append(a.view(dt).reshape(nev, self.size).T.copy())
        return out[0], out[1]

    def stats(self) -> dict:
        s = SolveStats()
        lib().bdg_get_solve_stats(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in SolveStats._fields_}

    def arrays(self):
        S = self.size
        g = lambda fn: np.ctypeslib.as_array(fn(self.h), shape=(S,)).copy()
        return dict(ltK=g(lib().bdg_local_term_K), ltM=g(lib().bdg_local_term_M),
                    pK=g(lib().bdg_precond_weight_K), pM=g(lib().bdg_precond_weight_M),
                    mu=float(lib().bdg_mu(self.h)))

    def device_context_handle(self):
        return lib().bdg_device_context(self.h)

    def device_context(self):
        """Non-owning DeviceContext of this problem (fields as of the last solve)."""
        return DeviceContext.wrap(self.device_context_handle(), self.N, self.L, self.cplx)


def read_bin(path):
    """Read a file of bdg_save_wavefunction_bin / bdg_write_modes_bin:
    returns (header dict, psi0) for kind 1 and (header, eigenvalues, U, V)
    with U, V of shape (S, nev) for kind 2."""
    with open(path, "rb") as f:
        raw = f.read(64)
        if len(raw) != 64 or raw[:8] != b"BDGBIN1\0":
            raise ValueError(f"{path}: not a BDGBIN1 file")
        kind, cplx = np.frombuffer(raw, np.uint32, 2, 8)
        dim, n0, n1, n2, count = (int(v) for v in np.frombuffer(raw, np.uint64, 5, 16))
        hdr = dict(kind=int(kind), complex=bool(cplx), dim=dim, N=(n0, n1, n2)[:dim],
                   count=count)
        S = n0 * n1 * n2
        dt = np.complex128 if cplx else np.float64
        if kind == 1:
            return hdr, np.fromfile(f, dt, S)
        w = np.fromfile(f, np.float64, count)
        U = np.fromfile(f, dt, S * count).reshape(count, S).T
        V = np.fromfile(f, dt, S * count).reshape(count, S).T
        return hdr, w, U, V
