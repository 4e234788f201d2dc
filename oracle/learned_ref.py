"""CPU restatement (numpy, FP64) of the learned preconditioner network.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the device
kernels (forward and parameter gradient); the product library never imports
it.  The network has no counterpart in the reference; its analytic base is the
reference's spectral preconditioner (reference src/precond_impl.inc:21-111,
weights from src/setup.c:150-177):

    T_X v = p_X * F^-1[ F(p_X * v) / (S * safe(mu + k^2/2)) ],   X in {K, M}

and the network generalises it as

    N(r) = T_M'(V),  V = U + s0 sig * sum_h v_h tanh(sum_nb w_{h,nb} U(x+nb)/(s0 sig) + d_h),
    U = T_K'(r),  sig = rms(r) per column,  s0 = mean(p_K^2) / safe(mu)
    T_X' v = a_X * S * irfft(g_X * rfft(b_X * v))
    a_X = p_X exp(alpha_X . phi),  b_X = p_X exp(beta_X . phi)
    g_X(k) = exp(interp(kappa_X, |k|)) / (S * safe(mu + k^2/2))
    phi = (rho, rho^2),  rho = |psi0|^2 / max |psi0|^2

i.e. spectral layer, 3x3x3 stencil layer with HID hidden channels and a
residual connection, spectral layer.  The stencil acts on the already
smoothed U and its output is smoothed again by T_M', so it cannot add
unfiltered high-k content to the search direction; N is positively
homogeneous in r (sig depends on r only, s0 on the fields only).
with nb over the 27 offsets of a 3x3x3 stencil (periodic), in the order
idx = (dz+1)*9 + (dy+1)*3 + (dx+1), and kappa interpolated linearly on
NKNOT knots spanning |k| in [0, max |k|].  At theta = 0 (except the stencil
weights, which only act through v = 0) the network is exactly T_M T_K.

Parameter vector layout (float32 on the device, float64 here):
  [0:2) alpha_K  [2:4) beta_K  [4:6) alpha_M  [6:8) beta_M
  [8:8+K) kappa_K  [8+K:8+2K) kappa_M
  then w (HID x 27, row-major), d (HID), v (HID)

Pinned by: theta = 0 reproducing the oracle's precondK/precondM product
(tests/test_learned.py), and the analytic gradient against central finite
differences of the loss on small grids.
"""
from __future__ import annotations

import numpy as np

NKNOT = 16
HID = 16
NPARAM = 8 + 2 * NKNOT + 29 * HID


def offsets():
    return [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64)
    K, H = NKNOT, HID
    o = 8 + 2 * K
    return dict(aK=t[0:2], bK=t[2:4], aM=t[4:6], bM=t[6:8], kK=t[8:8 + K], kM=t[8 + K:o],
                w=t[o:o + 27 * H].reshape(H, 27), d=t[o + 27 * H:o + 28 * H],
                v=t[o + 28 * H:o + 29 * H])


def init_theta(seed=0, wscale=0.1):
    rng = np.random.default_rng(seed)
    th = np.zeros(NPARAM)
    o = 8 + 2 * NKNOT
    th[o:o + 27 * HID] = rng.standard_normal(27 * HID) * wscale / np.sqrt(27.0)
    return th


class LearnedRef:
    """Grid-bound network: N = (Nx, Ny, Nz), x fastest, r2c along x."""

    def __init__(self, N, kx2, pK, pM, wf, mu):
        self.N = tuple(int(n) for n in N)
        nx, ny, nz = self.N
        self.S = nx * ny * nz
        self.shape = (nz, ny, nx)
        k2 = (kx2[2][:, None, None] + kx2[1][None, :, None] + kx2[0][None, None, :nx // 2 + 1])
        self.k2 = k2
        kabs = np.sqrt(k2)
        kmax = kabs.max()
        t = kabs / kmax * (NKNOT - 1) if kmax > 0 else np.zeros_like(kabs)
        i = np.minimum(np.floor(t).astype(int), NKNOT - 2)
        self.ki, self.kf = i, t - i
        d = mu + 0.5 * k2
        d = np.where(np.abs(d) < 1e-8, 1e-8, d)
        self.base = 1.0 / (self.S * d)
        self.pK = np.asarray(pK, float).reshape(self.shape)
        self.pM = np.asarray(pM, float).reshape(self.shape)
        smu = mu if abs(mu) >= 1e-8 else 1e-8
        self.s0 = float(np.mean(self.pK ** 2)) / smu
        rho = np.asarray(wf, float).reshape(self.shape) ** 2
        rho = rho / rho.max()
        self.phi = np.stack([rho, rho ** 2])
        # r2c weights: entries with 0 < kx < Nx/2 stand for k and -k
        wt = np.full(nx // 2 + 1, 2.0)
        wt[0] = 1.0
        if nx % 2 == 0:
            wt[nx // 2] = 1.0
        self.wt = wt[None, None, :]

    # -- pieces -------------------------------------------------------------
    def interp(self, kap):
        return (1 - self.kf) * kap[self.ki] + self.kf * kap[self.ki + 1]

    def tables(self, P):
        e = lambda c: np.exp(c[0] * self.phi[0] + c[1] * self.phi[1])
        return dict(aK=self.pK * e(P["aK"]), bK=self.pK * e(P["bK"]),
                    aM=self.pM * e(P["aM"]), bM=self.pM * e(P["bM"]),
                    gK=self.base * np.exp(self.interp(P["kK"])),
                    gM=self.base * np.exp(self.interp(P["kM"])))

    def filt(self, g, v):
        return self.S * np.fft.irfftn(g * np.fft.rfftn(v), s=self.shape, axes=(0, 1, 2))

    def conv_z(self, P, r):
        z = np.zeros((HID,) + self.shape)
        for n, (dz, dy, dx) in enumerate(offsets()):
            rs = np.roll(r, shift=(-dz, -dy, -dx), axis=(0, 1, 2))  # rs[x] = r[x + nb]
            z += P["w"][:, n][:, None, None, None] * rs[None]
        return z + P["d"][:, None, None, None]

    # -- network ------------------------------------------------------------
    def stencil(self, P, U, sig):
        un = U / (self.s0 * sig)
        th = np.tanh(self.conv_z(P, un))
        return un, th, U + self.s0 * sig * np.tensordot(P["v"], th, axes=(0, 0))

    def forward(self, theta, R):
        """R: (S, ncols) -> W: (S, ncols)."""
        P = unpack(theta)
        T = self.tables(P)
        out = np.zeros_like(np.asarray(R, float))
        for j in range(R.shape[1]):
            r = R[:, j].reshape(self.shape)
            sig = np.sqrt(np.mean(r * r))
            u = T["aK"] * self.filt(T["gK"], T["bK"] * r)
            _, _, v = self.stencil(P, u, sig)
            out[:, j] = (T["aM"] * self.filt(T["gM"], T["bM"] * v)).ravel()
        return out

    def grad(self, theta, R, dW):
        """Gradient of sum(dW * forward(theta, R)) with respect to theta."""
        P = unpack(theta)
        T = self.tables(P)
        g = dict((k, np.zeros_like(v)) for k, v in P.items())
        for j in range(R.shape[1]):
            r = R[:, j].reshape(self.shape)
            dw = dW[:, j].reshape(self.shape)
            sig = np.sqrt(np.mean(r * r))
            u = T["aK"] * self.filt(T["gK"], T["bK"] * r)
            un, th, v = self.stencil(P, u, sig)
            wsp = T["aM"] * self.filt(T["gM"], T["bM"] * v)
            # M layer
            g["aM"] += [np.sum(dw * wsp * f) for f in self.phi]
            tM = self.filt(T["gM"], T["aM"] * dw)
            g["bM"] += [np.sum(v * tM * T["bM"] * f) for f in self.phi]
            A = np.fft.rfftn(T["aM"] * dw)
            B = np.fft.rfftn(T["bM"] * v)
            g["kM"] += self.knot_sum(self.wt * np.real(np.conj(A) * B) * T["gM"])
            dv = T["bM"] * tM
            # stencil layer (residual connection): parameters and dU
            ds = self.s0 * sig * dv
            g["v"] += np.tensordot(th, ds, axes=([1, 2, 3], [0, 1, 2]))
            q = P["v"][:, None, None, None] * (1 - th ** 2) * dv[None]
            g["d"] += self.s0 * sig * q.sum(axis=(1, 2, 3))
            du = dv.copy()
            for n, (dz, dy, dx) in enumerate(offsets()):
                rs = np.roll(un, shift=(-dz, -dy, -dx), axis=(0, 1, 2))
                g["w"][:, n] += self.s0 * sig * np.tensordot(q, rs, axes=([1, 2, 3], [0, 1, 2]))
                # dU(y) += sum_h w_{h,nb} q_h(y - nb)
                qs = np.roll(q, shift=(dz, dy, dx), axis=(1, 2, 3))
                du += np.tensordot(P["w"][:, n], qs, axes=(0, 0))
            # K layer
            g["aK"] += [np.sum(du * u * f) for f in self.phi]
            tK = self.filt(T["gK"], T["aK"] * du)
            g["bK"] += [np.sum(r * tK * T["bK"] * f) for f in self.phi]
            A = np.fft.rfftn(T["aK"] * du)
            B = np.fft.rfftn(T["bK"] * r)
            g["kK"] += self.knot_sum(self.wt * np.real(np.conj(A) * B) * T["gK"])
        return np.concatenate([g["aK"], g["bK"], g["aM"], g["bM"], g["kK"], g["kM"],
                               g["w"].ravel(), g["d"], g["v"]])

    # -- one spectral layer alone (the first-order solver's per-half
    #    preconditioner: T'_K on R_U, T'_M on R_V) --------------------------
    def layer_forward(self, theta, R, layer):
        X = "K" if layer == 0 else "M"
        T = self.tables(unpack(theta))
        out = np.zeros_like(np.asarray(R, float))
        for j in range(R.shape[1]):
            r = R[:, j].reshape(self.shape)
            out[:, j] = (T["a" + X] * self.filt(T["g" + X], T["b" + X] * r)).ravel()
        return out

    def layer_grad(self, theta, R, dW, layer):
        """Gradient of sum(dW * layer_forward(theta, R, layer))."""
        X = "K" if layer == 0 else "M"
        T = self.tables(unpack(theta))
        oa, ob, ok = (0, 2, 8) if layer == 0 else (4, 6, 8 + NKNOT)
        g = np.zeros(NPARAM)
        for j in range(R.shape[1]):
            r = R[:, j].reshape(self.shape)
            dw = dW[:, j].reshape(self.shape)
            w = T["a" + X] * self.filt(T["g" + X], T["b" + X] * r)
            g[oa:oa + 2] += [np.sum(dw * w * f) for f in self.phi]
            t = self.filt(T["g" + X], T["a" + X] * dw)
            g[ob:ob + 2] += [np.sum(r * t * T["b" + X] * f) for f in self.phi]
            A = np.fft.rfftn(T["a" + X] * dw)
            B = np.fft.rfftn(T["b" + X] * r)
            g[ok:ok + NKNOT] += self.knot_sum(self.wt * np.real(np.conj(A) * B) * T["g" + X])
        return g

    def knot_sum(self, q):
        out = np.zeros(NKNOT)
        np.add.at(out, self.ki.ravel(), ((1 - self.kf) * q).ravel())
        np.add.at(out, self.ki.ravel() + 1, (self.kf * q).ravel())
        return out


def cosine_loss(W, E):
    """mean_j (1 - <w,e>^2 / (|w|^2 |e|^2)) and its gradient with respect to W."""
    c = np.sum(W * E, axis=0)
    ww = np.sum(W * W, axis=0)
    ee = np.sum(E * E, axis=0)
    n = W.shape[1]
    loss = float(np.mean(1 - c * c / (ww * ee)))
    dW = (-2 * c / (ww * ee))[None, :] * E + (2 * c * c / (ww * ww * ee))[None, :] * W
    return loss, dW / n
