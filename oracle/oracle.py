"""ctypes wrapper around the FP64 CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` as the checker.  The product library
(``bdg_lobpcg_amd``) never imports it.

The C side restates the reference operators one vector at a time
(pstuermer/BDG_LOBPCG src/kinetic_impl.inc, src/matmulK_impl.inc,
src/matmulM_impl.inc, src/matmulLrep_impl.inc, src/matmulSwap_impl.inc,
src/precond_impl.inc, src/dipolar_conv_impl.inc, src/dipolar_meanfield_impl.inc,
src/dipolar.c, src/setup.c, src/matmul_ctx.c).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

OPS = {"kinetic": 0, "K": 1, "M": 2, "Lrep": 3, "Swap": 4,
       "precondK": 5, "precondM": 6, "precondLrep": 7, "dipolar_conv": 8}


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
        L.orc_ctx_new.restype = P
        L.orc_ctx_new.argtypes = [C.c_int, i64p, np.ctypeslib.ndpointer(np.float64, flags="C"), C.c_int]
        L.orc_ctx_free.argtypes = [P]
        L.orc_ctx_size.restype = C.c_int64
        L.orc_ctx_size.argtypes = [P]
        L.orc_ctx_ksize.restype = C.c_int64
        L.orc_ctx_ksize.argtypes = [P]
        for fn in ("orc_ctx_get_k2", "orc_dipolar_get_kernel", "orc_dipolar_meanfield"):
            getattr(L, fn).argtypes = [P, P]
        L.orc_ctx_get_kx2.argtypes = [P, C.c_int, P]
        L.orc_ctx_set_fields.argtypes = [P, P, P, P, P, P, C.c_double]
        L.orc_dipolar_set_kernel.argtypes = [P, C.c_double, P, C.c_double]
        L.orc_apply_block.argtypes = [P, C.c_int, P, P, C.c_int64, C.c_int64]
        L.orc_fft_c2c.argtypes = [C.c_int, P, P, C.c_int]
        L.orc_fft_r2c.argtypes = [C.c_int, P, P, P]
        L.orc_fft_c2r.argtypes = [C.c_int, P, P, P]
        L.orc_setup_trap_harmonic.argtypes = [P, P, P, P]
        L.orc_setup_contact.argtypes = [P, P, C.c_double, C.c_double, P, P]
        L.orc_setup_mu.argtypes = [C.c_int64, C.c_double, P, P, P, P]
        L.orc_itp_groundstate.argtypes = [P, P, C.c_double, C.c_double, P, C.c_double,
                                          C.c_double, C.c_int64, P, P]
        L.orc_xorshift32.restype = C.c_uint32
        L.orc_xorshift32.argtypes = [C.POINTER(C.c_uint32)]
        L.orc_xrand.restype = C.c_double
        L.orc_xrand.argtypes = [C.POINTER(C.c_uint32)]
        L.orc_start_block_uniform.argtypes = [P, C.c_uint32, C.c_int64, P]
        L.orc_reuse_block.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int, P, P, P,
                                      C.c_double, C.c_uint32, P]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _n3(N):
    return np.ascontiguousarray(np.asarray(N, dtype=np.int64))


# ---------------------------------------------------------------- FFT ------
def fft_c2c(x: np.ndarray, sign: int = -1) -> np.ndarray:
    """Unnormalised c2c of an array shaped (Nz, Ny, Nx) / (Ny, Nx) / (Nx,)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    N = _n3(a.shape[::-1])
    lib().orc_fft_c2c(len(N), _p(N), _p(a), sign)
    return a


def fft_r2c(x: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.float64)
    N = _n3(a.shape[::-1])
    out = np.zeros(a.shape[:-1] + (a.shape[-1] // 2 + 1,), dtype=np.complex128)
    lib().orc_fft_r2c(len(N), _p(N), _p(a), _p(out))
    return out


def fft_c2r(h: np.ndarray, nx: int) -> np.ndarray:
    a = np.ascontiguousarray(h, dtype=np.complex128).copy()
    shape = a.shape[:-1] + (nx,)
    N = _n3(shape[::-1])
    out = np.zeros(shape, dtype=np.float64)
    lib().orc_fft_c2r(len(N), _p(N), _p(a), _p(out))
    return out


# ------------------------------------------------------------ context ------
class Ctx:
    """One oracle context: grid, k-arrays and the physics fields."""

    def __init__(self, N, L, complex_psi0: bool = False):
        self.N = tuple(int(n) for n in N)
        self.L = tuple(float(v) for v in L)
        self.dim = len(self.N)
        self.cplx = bool(complex_psi0)
        Ls = np.ascontiguousarray(self.L, dtype=np.float64)
        self._h = lib().orc_ctx_new(self.dim, _n3(self.N), Ls, int(self.cplx))
        self.size = int(lib().orc_ctx_size(self._h))
        self.k_size = int(lib().orc_ctx_ksize(self._h))
        self.dtype = np.complex128 if self.cplx else np.float64

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_ctx_free(h)
            self._h = None

    def k2(self):
        out = np.zeros(self.k_size)
        lib().orc_ctx_get_k2(self._h, _p(out))
        return out

    def kx2(self, d: int):
        out = np.zeros(self.N[d])
        lib().orc_ctx_get_kx2(self._h, d, _p(out))
        return out

    def set_fields(self, ltK=None, ltM=None, wf=None, pK=None, pM=None, mu=0.0):
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return _p(a)

        lib().orc_ctx_set_fields(self._h, arr(ltK, np.float64), arr(ltM, np.float64),
                                 arr(wf, self.dtype), arr(pK, np.float64),
                                 arr(pM, np.float64), float(mu))

    def set_dipolar(self, g_ddi, direction, cutoff):
        d = np.ascontiguousarray(direction, dtype=np.float64)
        lib().orc_dipolar_set_kernel(self._h, float(g_ddi), _p(d), float(cutoff))

    def dipolar_kernel(self):
        out = np.zeros(self.k_size)
        lib().orc_dipolar_get_kernel(self._h, _p(out))
        return out

    def dipolar_meanfield(self):
        out = np.zeros(self.size)
        lib().orc_dipolar_meanfield(self._h, _p(out))
        return out

    def apply(self, op: str, X: np.ndarray) -> np.ndarray:
        """Apply operator `op` to a vector (1D) or column-major block (2D, rows x cols)."""
        X = np.asarray(X)
        one = X.ndim == 1
        Xf = np.asfortranarray(X.reshape(X.shape[0], -1) if not one else X[:, None], dtype=self.dtype)
        Y = np.zeros_like(Xf)
        lib().orc_apply_block(self._h, OPS[op], _p(Xf), _p(Y), Xf.shape[1], Xf.shape[0])
        return Y[:, 0] if one else Y

    # -------- setup helpers
    def trap_harmonic(self, omega, ltK, ltM):
        om = np.ascontiguousarray(omega, dtype=np.float64)
        lib().orc_setup_trap_harmonic(self._h, _p(om), _p(ltK), _p(ltM))

    def contact(self, wf, gK, gM, ltK, ltM):
        w = np.ascontiguousarray(wf, dtype=self.dtype)
        lib().orc_setup_contact(self._h, _p(w), float(gK), float(gM), _p(ltK), _p(ltM))

    def itp_groundstate(self, psi, omega, g, g_dd, direction, cutoff, dt, nsteps):
        psi = np.ascontiguousarray(psi, dtype=self.dtype).copy()
        om = np.ascontiguousarray(omega, dtype=np.float64)
        d = np.ascontiguousarray(direction, dtype=np.float64)
        mu = C.c_double(0.0)
        lib().orc_itp_groundstate(self._h, _p(om), float(g), float(g_dd), _p(d), float(cutoff),
                                  float(dt), int(nsteps), _p(psi), C.cast(C.pointer(mu), C.c_void_p))
        return psi, mu.value

    def start_block_uniform(self, seed: int, ncols: int) -> np.ndarray:
        X = np.zeros((2 * self.size, ncols), dtype=self.dtype, order="F")
        lib().orc_start_block_uniform(self._h, int(seed), int(ncols), _p(X))
        return X


def setup_mu(mu, ltK, ltM):
    S = ltK.size
    pK = np.zeros(S)
    pM = np.zeros(S)
    lib().orc_setup_mu(S, float(mu), _p(ltK), _p(ltM), _p(pK), _p(pM))
    return pK, pM


def xrand_stream(seed: int, count: int) -> np.ndarray:
    s = C.c_uint32(seed)
    return np.array([lib().orc_xrand(C.byref(s)) for _ in range(count)])


def reuse_block(U, V, m, noise, wf=None, seed=12345):
    """Reuse start block [u; v] (2S x m) of a sweep from modes U, V (S x nev)
    (reference src/bdg.c:112-207, bdg_reuse_modes)."""
    cplx = np.iscomplexobj(U) or np.iscomplexobj(V)
    dt = np.complex128 if cplx else np.float64
    U = np.asfortranarray(U, dtype=dt)
    V = np.asfortranarray(V, dtype=dt)
    S, nev = U.shape
    X = np.zeros((2 * S, m), dtype=dt, order="F")
    w = None if wf is None else np.ascontiguousarray(wf, dtype=dt)
    lib().orc_reuse_block(S, nev, int(m), int(cplx), _p(U), _p(V),
                          None if w is None else _p(w), float(noise), int(seed), _p(X))
    return X


def num_threads() -> int:
    return int(lib().orc_num_threads())
