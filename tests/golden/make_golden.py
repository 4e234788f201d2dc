"""Writes tests/golden/fixtures.npz: golden vectors for the CPU oracle.

Independent of the oracle's C code: FFTs come from numpy.fft, the operators
from a separate numpy restatement of the reference formulas
(pstuermer/BDG_LOBPCG src/kinetic_impl.inc:17-30, src/matmulK_impl.inc:11-28,
src/matmulM_impl.inc:14-44, src/precond_impl.inc:21-118,
src/dipolar_conv_impl.inc:22-52, src/dipolar_meanfield_impl.inc:20-55,
src/dipolar.c:21-72, src/setup.c:17-177), and the eigenvalues from dense
scipy eigensolves of the explicitly assembled K and M on tiny grids
(independent of any iterative solver).

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np
import scipy.linalg as sla

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.npz")


# ------------------------------------------------------------ restatement
class NumpyBdG:
    """Arrays are flat, x fastest; shape (Nz, Ny, Nx) for numpy.fft."""

    def __init__(self, N, L, cplx):
        self.N, self.L, self.cplx = tuple(N), tuple(L), cplx
        self.shape = tuple(reversed(self.N))
        self.S = int(np.prod(N))
        # k_d(j) = (2 pi / L_d) (j <= N_d/2 ? j : j - N_d): the Nyquist index is
        # +N/2 (reference src/matmul_ctx.c:106-112, src/dipolar.c:267-276),
        # which matters for a tilted dipole where k.d mixes signs
        ks = []
        for n, l in zip(self.N, self.L):
            j = np.arange(n)
            ks.append(2 * np.pi / l * np.where(j <= n // 2, j, j - n))
        if not cplx:
            ks[0] = ks[0][: self.N[0] // 2 + 1].copy()
        grids = np.meshgrid(*reversed(ks), indexing="ij")  # (kz, ky, kx) order
        self.kvec = list(reversed(grids))                   # [kx, ky, kz]
        self.k2 = sum(k ** 2 for k in self.kvec)

    def fwd(self, x):
        a = x.reshape(self.shape)
        return np.fft.fftn(a) if self.cplx else np.fft.rfftn(a)

    def inv(self, F):
        # numpy's inverse includes 1/S; the reference applies 1/S itself, so
        # multiply back to keep the reference's order of operations explicit
        if self.cplx:
            return (np.fft.ifftn(F) * self.S).ravel()
        return (np.fft.irfftn(F, s=self.shape, axes=tuple(range(len(self.shape)))) * self.S).ravel()

    def kernel(self, direction, R):
        d = np.asarray(direction, float)
        if len(self.N) != 3:
            raise ValueError("dipolar needs 3D")
        kd = self.kvec[0] * d[0] + self.kvec[1] * d[1] + self.kvec[2] * d[2]
        with np.errstate(divide="ignore", invalid="ignore"):
            c2 = kd ** 2 / (self.k2 * (d @ d))
            kR = np.sqrt(self.k2) * R
            f = 1 + 3 * np.cos(kR) / kR ** 2 - 3 * np.sin(kR) / kR ** 3
            U = (3 * c2 - 1) * f
        U[self.k2 < 1e-50] = 0.0
        return U

    def setup(self, psi, omega, g, gdd=0.0, direction=(0, 0, 1), R=None, mu=0.0):
        xs = [(np.arange(n) - n / 2) * l / n for n, l in zip(self.N, self.L)]
        grids = np.meshgrid(*reversed(xs), indexing="ij")
        r = list(reversed(grids))
        V = sum(0.5 * om ** 2 * ri ** 2 for om, ri in zip(omega, r)).ravel()
        n = np.abs(psi) ** 2
        ltK = V + g * n
        ltM = V + g * n + 2 * g * n
        self.psi = psi
        self.gdd = gdd
        self.U = None
        if gdd:
            self.U = self.kernel(direction, R)
            phi = gdd * np.real(self.inv(self.U * self.fwd(n + 0j if self.cplx else n) / self.S))
            ltK = ltK + phi
            ltM = ltM + phi
            self.gphi = phi
        self.pK = 1 / np.sqrt(np.maximum(1e-8, ltK))
        self.pM = 1 / np.sqrt(np.maximum(1e-8, ltM))
        self.ltK, self.ltM, self.mu = ltK - mu, ltM - mu, mu

    def kinetic(self, x):
        return self.inv(self.fwd(x) * self.k2 * 0.5 / self.S)

    def dipconv(self, x):
        b = np.conj(self.psi) * x
        return self.gdd * self.psi * self.inv(self.fwd(b) * self.U / self.S)

    def K(self, x):
        return self.kinetic(x) + self.ltK * x

    def M(self, x):
        y = self.kinetic(x) + self.ltM * x
        if self.U is not None:
            y = y + 2 * self.dipconv(x)
        return y

    def precond(self, p, x):
        d = self.mu + 0.5 * self.k2
        d = np.where(np.abs(d) < 1e-8, 1e-8, d)
        return p * self.inv(self.fwd(p * x) / (self.S * d))


def dense(op, S, cplx):
    dt = np.complex128 if cplx else np.float64
    E = np.eye(S, dtype=dt)
    return np.stack([op(E[:, j]) for j in range(S)], axis=1)


def bdg_dense_omega(nb, count):
    K = dense(nb.K, nb.S, nb.cplx)
    M = dense(nb.M, nb.S, nb.cplx)
    K = 0.5 * (K + K.conj().T)
    M = 0.5 * (M + M.conj().T)
    w2 = sla.eigh(M @ K @ M, M, eigvals_only=True)
    return np.sqrt(np.maximum(w2[:count], 0.0)), w2[:count]


def main():
    rng = np.random.default_rng(20261203)
    fx = {}
    # ---- FFT golden vectors (numpy.fft)
    shapes = [(8, 8, 8), (12, 12, 12), (10, 10, 10), (4, 8, 16), (6, 5, 4), (64,), (12, 10)]
    for i, sh in enumerate(shapes):
        xr = rng.standard_normal(sh)
        xc = rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
        fx[f"fft{i}_shape"] = np.array(sh)
        fx[f"fft{i}_xr"] = xr
        fx[f"fft{i}_xc"] = xc
        fx[f"fft{i}_r2c"] = np.fft.rfftn(xr)
        fx[f"fft{i}_c2c"] = np.fft.fftn(xc)
        fx[f"fft{i}_c2c_inv"] = np.fft.ifftn(xc) * xc.size
    # ---- operator outputs (numpy restatement)
    cases = [
        ("op0", (8, 6, 10), (6.0, 5.0, 7.0), False),
        ("op1", (6, 8, 5), (5.0, 6.0, 4.0), True),
    ]
    for tag, N, L, cplx in cases:
        nb = NumpyBdG(N, L, cplx)
        xs = [(np.arange(n) - n / 2) * l / n for n, l in zip(N, L)]
        Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
        psi = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 2).ravel()
        if cplx:
            psi = psi * np.exp(0.4j * X.ravel() - 0.2j * Z.ravel())
        nb.setup(psi, (1.0, 0.8, 1.3), g=7.0, gdd=3.0, direction=(0.3, 0.0, 1.0), R=2.5, mu=1.1)
        v = rng.standard_normal(nb.S) + (1j * rng.standard_normal(nb.S) if cplx else 0)
        fx[f"{tag}_N"] = np.array(N)
        fx[f"{tag}_L"] = np.array(L)
        fx[f"{tag}_cplx"] = np.array(int(cplx))
        fx[f"{tag}_psi"] = psi
        fx[f"{tag}_v"] = v
        fx[f"{tag}_ltK"] = nb.ltK
        fx[f"{tag}_ltM"] = nb.ltM
        fx[f"{tag}_pK"] = nb.pK
        fx[f"{tag}_pM"] = nb.pM
        fx[f"{tag}_gphi"] = nb.gphi
        fx[f"{tag}_U"] = nb.U.ravel()
        fx[f"{tag}_kinetic"] = nb.kinetic(v)
        fx[f"{tag}_K"] = nb.K(v)
        fx[f"{tag}_M"] = nb.M(v)
        fx[f"{tag}_dipconv"] = nb.dipconv(v)
        fx[f"{tag}_precondK"] = nb.precond(nb.pK, v)
        fx[f"{tag}_precondM"] = nb.precond(nb.pM, v)
    # ---- dense BdG spectra on tiny grids (mu = 0: K, M positive definite)
    dcases = [
        ("eig0", (6, 6, 6), (6.0, 6.0, 6.0), False, (1.0, 1.0, 1.0), 15.0, 5.0, 3.0),
        ("eig1", (6, 8, 6), (6.0, 8.0, 6.0), True, (1.0, 0.7, 1.2), 12.0, 4.0, 3.0),
        ("eig2", (10, 8), (8.0, 6.0), False, (1.0, 1.3), 6.0, 0.0, 0.0),
        ("eig3", (48,), (10.0,), True, (1.0,), 4.0, 0.0, 0.0),
    ]
    for tag, N, L, cplx, om, g, gdd, R in dcases:
        nb = NumpyBdG(N, L, cplx)
        xs = [(np.arange(n) - n / 2) * l / n for n, l in zip(N, L)]
        grids = np.meshgrid(*reversed(xs), indexing="ij")
        r = list(reversed(grids))
        psi = np.exp(-sum(ri ** 2 for ri in r) / 2).ravel()
        psi = psi / np.sqrt(np.sum(psi ** 2) * np.prod([l / n for n, l in zip(N, L)]))
        if cplx:
            psi = psi * np.exp(0.5j * r[0].ravel())
        nb.setup(psi, om, g=g, gdd=gdd, direction=(0, 0, 1), R=R if gdd else None, mu=0.0)
        w, w2 = bdg_dense_omega(nb, 12)
        fx[f"{tag}_N"] = np.array(N)
        fx[f"{tag}_L"] = np.array(L)
        fx[f"{tag}_cplx"] = np.array(int(cplx))
        fx[f"{tag}_omega"] = np.array(om)
        fx[f"{tag}_g"] = np.array(g)
        fx[f"{tag}_gdd"] = np.array(gdd)
        fx[f"{tag}_R"] = np.array(R)
        fx[f"{tag}_psi"] = psi
        fx[f"{tag}_w"] = w
    np.savez_compressed(OUT, **fx)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
