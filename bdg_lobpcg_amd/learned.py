"""Training of the learned preconditioner (the network of include/bdgx.h,
`bdgx_lp_*`) on (residual, error-correction) pairs that the solver itself
produces, entirely on the device through the bdgx C ABI.

Pairs.  A training problem is solved to convergence (FP64) and again with
small iteration budgets from the same start block; every truncated solve is
a snapshot X_i of the iteration.  For each snapshot column (normalised to
x^T M x = 1, as the Rayleigh-Ritz leaves them):

    r = K M x - w^2 x                               (what the solver feeds T)
    e = P_i Pi x,  Pi = projector on the converged block,
                   P_i = projector orthogonal to span(X_i)

so e is the part of the correction toward the converged eigenspace that a
search direction can add to the current subspace.

Loss.  For the network output w = N(r), projected the same way (w_p = P_i w),
mean over columns of 1 - <w_p, e>^2 / (|w_p|^2 |e|^2): scale-free, since the
Rayleigh-Ritz chooses the coefficients.  Forward, projection, loss, the
adjoint of the projection and the network's parameter gradient
(bdgx_lp_backward) run on the GPU; Adam runs on the 504 host parameters.
"""
from __future__ import annotations

import time

import numpy as np

from . import FP32, FP64, LP_HIDDEN, LP_NKNOT, LP_NPARAM, DeviceBuffer, check, lib


def init_theta(seed: int = 0, wscale: float = 0.1) -> np.ndarray:
    """theta = 0 except small random stencil weights; with v = 0 the network
    is exactly the analytic preconditioner T_M T_K."""
    rng = np.random.default_rng(seed)
    th = np.zeros(LP_NPARAM, dtype=np.float64)
    o = 8 + 2 * LP_NKNOT
    th[o:o + 27 * LP_HIDDEN] = rng.standard_normal(27 * LP_HIDDEN) * wscale / np.sqrt(27.0)
    return th.astype(np.float32)


class _Dev:
    """Block helpers on a device context (column-major, ld = size)."""

    def __init__(self, d):
        self.d = d
        self.S = d.size

    def gram(self, prec, A, p, B, q, a_off=0, b_off=0):
        es = 8 if prec == FP64 else 4
        dG = DeviceBuffer(8 * p * q)
        check(lib().bdgx_gram(self.d.h, prec, A.ptr.value + a_off * self.S * es, self.S, p,
                              B.ptr.value + b_off * self.S * es, self.S, q, self.S, 0, dG.ptr),
              "gram")
        G = np.zeros((p, q), order="F")
        check(lib().bdgx_memcpy_d2h(self.d.h, G.ctypes.data, dG.ptr, G.nbytes), "d2h")
        return G

    def combine(self, prec, A, k, Cm, Y, a_off=0, y_off=0):
        es = 8 if prec == FP64 else 4
        Cm = np.asfortranarray(Cm, dtype=np.float64)
        dC = self.d.upload(Cm)
        check(lib().bdgx_combine(self.d.h, prec, A.ptr.value + a_off * self.S * es, self.S, k,
                                 dC.ptr, Y.ptr.value + y_off * self.S * es, self.S, Cm.shape[1],
                                 self.S), "combine")

    def apply(self, op, prec, X, Y, ncols):
        self.d.apply_device(op, prec, X, Y, ncols, self.S)

    def convert(self, ps, X, pd, Y, ncols, x_off=0, y_off=0):
        esx = 8 if ps == FP64 else 4
        esy = 8 if pd == FP64 else 4
        check(lib().bdgx_convert(self.d.h, ps, X.ptr.value + x_off * self.S * esx, self.S, pd,
                                 Y.ptr.value + y_off * self.S * esy, self.S, self.S, ncols),
              "convert")


class Snapshot:
    """Device-resident training batch: columns of R, E and an orthonormal
    basis Q of span(X), FP32, packed as [Q | W | E | W_p] (4m columns) plus R."""

    def __init__(self, S, m):
        self.m = m
        self.B4 = DeviceBuffer(4 * 4 * S * m)
        self.R = DeviceBuffer(4 * S * m)


def _chol_inv_t(G):
    L = np.linalg.cholesky(0.5 * (G + G.T))
    return np.linalg.inv(L).T  # X L^-T has orthonormal columns


def make_snapshot(d, Xs, ws2, Xstar) -> Snapshot:
    """Xs: snapshot block (S x m, host FP64), ws2: its w^2, Xstar: converged
    block (S x n). Builds r, e, Q on the device."""
    dev = _Dev(d)
    S, m = Xs.shape
    n = Xstar.shape[1]
    dX = d.upload(Xs)
    dMX = DeviceBuffer(8 * S * m)
    dev.apply("M", FP64, dX, dMX, m)
    # x^T M x = 1
    xm = np.diag(dev.gram(FP64, dX, m, dMX, m))
    sc = 1.0 / np.sqrt(np.maximum(np.abs(xm), 1e-300))
    dev.combine(FP64, dX, m, np.diag(sc), dMX)  # dMX <- X diag(sc) (scratch use)
    dX, dMX = dMX, dX
    dev.apply("M", FP64, dX, dMX, m)
    dKMX = DeviceBuffer(8 * S * m)
    dev.apply("K", FP64, dMX, dKMX, m)
    # R = KMX - w^2 X  via [KMX | X] [I; -diag w2]
    dRX = DeviceBuffer(8 * S * 2 * m)
    check(lib().bdgx_convert(d.h, FP64, dKMX.ptr, S, FP64, dRX.ptr, S, S, m), "convert")
    check(lib().bdgx_convert(d.h, FP64, dX.ptr, S, FP64, dRX.ptr.value + 8 * S * m, S, S, m),
          "convert")
    dR64 = DeviceBuffer(8 * S * m)
    dev.combine(FP64, dRX, 2 * m, np.vstack([np.eye(m), -np.diag(ws2)]), dR64)
    # Q = X L^-T
    G = dev.gram(FP64, dX, m, dX, m)
    dQ = DeviceBuffer(8 * S * m)
    dev.combine(FP64, dX, m, _chol_inv_t(G), dQ)
    # Pi x = X* (X*^T X*)^-1 X*^T x, then e = (I - Q Q^T) Pi x
    dXs = d.upload(Xstar)
    Gs = dev.gram(FP64, dXs, n, dXs, n)
    Z = dev.gram(FP64, dXs, n, dX, m)
    dPi = DeviceBuffer(8 * S * m)
    dev.combine(FP64, dXs, n, np.linalg.solve(Gs, Z), dPi)
    QtP = dev.gram(FP64, dQ, m, dPi, m)
    dQP = DeviceBuffer(8 * S * 2 * m)
    check(lib().bdgx_convert(d.h, FP64, dQ.ptr, S, FP64, dQP.ptr, S, S, m), "convert")
    check(lib().bdgx_convert(d.h, FP64, dPi.ptr, S, FP64, dQP.ptr.value + 8 * S * m, S, S, m),
          "convert")
    dE64 = DeviceBuffer(8 * S * m)
    dev.combine(FP64, dQP, 2 * m, np.vstack([-QtP, np.eye(m)]), dE64)
    snap = Snapshot(S, m)
    dev.convert(FP64, dQ, FP32, snap.B4, m, 0, 0)
    dev.convert(FP64, dE64, FP32, snap.B4, m, 0, 2 * m)
    dev.convert(FP64, dR64, FP32, snap.R, m)
    d.sync()
    return snap


def loss_and_grad(d, theta, snap: Snapshot):
    """Cosine loss of the projected network output and its theta gradient."""
    dev = _Dev(d)
    S, m = d.size, snap.m
    d.lp_set_params(theta)
    # W into block 1
    check(lib().bdgx_lp_apply(d.h, FP32, snap.R.ptr, snap.B4.ptr.value + 4 * S * m, m, S),
          "lp_apply")
    QtW = dev.gram(FP32, snap.B4, m, snap.B4, m, 0, m)
    # W_p = W - Q Q^T W into block 3
    dev.combine(FP32, snap.B4, 2 * m, np.vstack([-QtW, np.eye(m)]), snap.B4, 0, 3 * m)
    EW = dev.gram(FP32, snap.B4, 2 * m, snap.B4, 2 * m, 2 * m, 2 * m)  # [E W_p]^T [E W_p]
    ee = np.diag(EW)[:m]
    ww = np.diag(EW)[m:]
    c = np.diag(EW[:m, m:])
    ww = np.maximum(ww, 1e-300)
    ee = np.maximum(ee, 1e-300)
    loss = float(np.mean(1.0 - c * c / (ww * ee)))
    alpha = (-2.0 * c / (ww * ee)) / m
    beta = (2.0 * c * c / (ww * ww * ee)) / m
    # dW_p = E diag(alpha) + W_p diag(beta) into block 1
    dev.combine(FP32, snap.B4, 2 * m, np.vstack([np.diag(alpha), np.diag(beta)]), snap.B4, 2 * m, m)
    # dW = (I - Q Q^T) dW_p into block 3
    QtD = dev.gram(FP32, snap.B4, m, snap.B4, m, 0, m)
    dev.combine(FP32, snap.B4, 2 * m, np.vstack([-QtD, np.eye(m)]), snap.B4, 0, 3 * m)
    g = np.zeros(LP_NPARAM, dtype=np.float32)
    check(lib().bdgx_lp_backward(d.h, snap.R.ptr, snap.B4.ptr.value + 12 * S * m, m, S,
                                 g.ctypes.data), "lp_backward")
    return loss, g.astype(np.float64)


def train(d, snaps, theta0=None, steps=60, lr=0.02, seed=0, reg=0.0, log=None):
    """Adam over the snapshot batches (one batch per step, cycling), with an
    optional pull reg * |theta - theta0|^2 toward the starting point (the
    analytic preconditioner), which keeps the network a perturbation of it."""
    th = np.asarray(init_theta(seed) if theta0 is None else theta0, dtype=np.float64)
    th0 = th.copy()
    mom = np.zeros_like(th)
    vel = np.zeros_like(th)
    b1, b2, eps = 0.9, 0.999, 1e-8
    hist = []
    t0 = time.perf_counter()
    for it in range(1, steps + 1):
        snap = snaps[(it - 1) % len(snaps)]
        loss, g = loss_and_grad(d, th.astype(np.float32), snap)
        hist.append(loss)
        if reg > 0:
            g = g + 2.0 * reg * (th - th0)
        mom = b1 * mom + (1 - b1) * g
        vel = b2 * vel + (1 - b2) * g * g
        th -= lr * (mom / (1 - b1 ** it)) / (np.sqrt(vel / (1 - b2 ** it)) + eps)
        if log and (it % 10 == 0 or it == 1):
            log(f"step {it}: loss {loss:.5f} ({time.perf_counter() - t0:.1f} s)")
    return th.astype(np.float32), hist


def evaluate(d, theta, snaps):
    return float(np.mean([loss_and_grad(d, theta, s)[0] for s in snaps]))


def collect(make_problem, nev, snapshots=(4, 8, 14, 22, 32), log=None):
    """Converged solve plus truncated solves of the same problem; returns the
    problem (whose device context carries the fields) and the snapshots."""
    ref = make_problem(maxiter=2000)
    ref.solve()
    _, Xstar = ref.modes(nev)
    snaps = []
    d = ref.device_context()
    for k in snapshots:
        p = make_problem(maxiter=k)
        p.solve()
        w = p.eigenvalues(nev)
        _, X = p.modes(nev)
        snaps.append(make_snapshot(d, np.asfortranarray(X), np.sign(w) * w * w, Xstar))
        p.close()
        if log:
            log(f"snapshot after {k} iterations")
    return ref, d, snaps


# ---------------------------------------------------------------------------
# First-order form: one spectral layer per half.  The first-order solver
# (K u = w v, M v = w u, separate U and V search spaces) takes its new
# directions W_U = T'_K R_U, W_V = T'_M R_V; the pairs of a snapshot are,
# per half X in {U, V} (columns of the Ritz block, [u; v] of unit norm):
#
#     R_U = K u - |w| v,   R_V = M v - w u          (the solver's residuals)
#     E_X = (I - Q_X Q_X^T) Pi_X x                  (x = u or v)
#
# Q_X an orthonormal basis of the snapshot's X block, Pi_X the projector on
# the converged X block.  Loss: the mean over both halves of the cosine loss
# above, each half through its own layer.

class SnapshotFO:
    """Per half: [Q | W | E | W_p] (4m FP32 columns) and R (m FP32 columns)."""

    def __init__(self, S, m):
        self.m = m
        self.B4 = [DeviceBuffer(4 * 4 * S * m) for _ in range(2)]
        self.R = [DeviceBuffer(4 * S * m) for _ in range(2)]


def _pack2(d, A, B, m):
    """[A | B] (S x 2m FP64) from two FP64 S x m blocks."""
    S = d.size
    out = DeviceBuffer(8 * S * 2 * m)
    check(lib().bdgx_convert(d.h, FP64, A.ptr, S, FP64, out.ptr, S, S, m), "convert")
    check(lib().bdgx_convert(d.h, FP64, B.ptr, S, FP64, out.ptr.value + 8 * S * m, S, S, m),
          "convert")
    return out


def make_snapshot_fo(d, U, V, w, Ustar, Vstar) -> SnapshotFO:
    """U, V, w: a (truncated) first-order solve's modes (M v = w u) and
    frequencies; Ustar, Vstar: the converged modes.  Builds R, E, Q of both
    halves on the device."""
    dev = _Dev(d)
    S, m = U.shape
    n = Ustar.shape[1]
    dU, dV = d.upload(np.asfortranarray(U)), d.upload(np.asfortranarray(V))
    dKU, dMV = DeviceBuffer(8 * S * m), DeviceBuffer(8 * S * m)
    dev.apply("K", FP64, dU, dKU, m)
    dev.apply("M", FP64, dV, dMV, m)
    aw = np.abs(np.asarray(w, dtype=np.float64))
    snap = SnapshotFO(S, m)
    for h, (dX, dImg, dOther, coef, Xs) in enumerate(
            ((dU, dKU, dV, aw, Ustar), (dV, dMV, dU, np.asarray(w, np.float64), Vstar))):
        dR64 = DeviceBuffer(8 * S * m)
        dev.combine(FP64, _pack2(d, dImg, dOther, m), 2 * m, np.vstack([np.eye(m), -np.diag(coef)]),
                    dR64)
        G = dev.gram(FP64, dX, m, dX, m)
        dQ = DeviceBuffer(8 * S * m)
        dev.combine(FP64, dX, m, _chol_inv_t(G), dQ)
        dXs = d.upload(np.asfortranarray(Xs))
        Gs = dev.gram(FP64, dXs, n, dXs, n)
        Z = dev.gram(FP64, dXs, n, dX, m)
        dPi = DeviceBuffer(8 * S * m)
        dev.combine(FP64, dXs, n, np.linalg.solve(Gs, Z), dPi)
        QtP = dev.gram(FP64, dQ, m, dPi, m)
        dE64 = DeviceBuffer(8 * S * m)
        dev.combine(FP64, _pack2(d, dQ, dPi, m), 2 * m, np.vstack([-QtP, np.eye(m)]), dE64)
        dev.convert(FP64, dQ, FP32, snap.B4[h], m, 0, 0)
        dev.convert(FP64, dE64, FP32, snap.B4[h], m, 0, 2 * m)
        dev.convert(FP64, dR64, FP32, snap.R[h], m)
    d.sync()
    return snap


def loss_and_grad_fo(d, theta, snap: SnapshotFO, weighted=False):
    """Mean cosine loss of the two projected layer outputs and its gradient.
    weighted: columns weighted by their share of the snapshot's squared error
    |E_j|^2 / sum |E|^2 instead of 1 / m (converged columns, whose E is
    rounding noise, then carry no weight)."""
    dev = _Dev(d)
    S, m = d.size, snap.m
    d.lp_set_params(theta)
    loss, grad = 0.0, np.zeros(LP_NPARAM)
    for h in (0, 1):
        B4, R = snap.B4[h], snap.R[h]
        d.lp_apply_layer_device(h, R, B4, m, 0, m)
        QtW = dev.gram(FP32, B4, m, B4, m, 0, m)
        dev.combine(FP32, B4, 2 * m, np.vstack([-QtW, np.eye(m)]), B4, 0, 3 * m)
        EW = dev.gram(FP32, B4, 2 * m, B4, 2 * m, 2 * m, 2 * m)
        ee = np.maximum(np.diag(EW)[:m], 1e-300)
        ww = np.maximum(np.diag(EW)[m:], 1e-300)
        c = np.diag(EW[:m, m:])
        wt = ee / ee.sum() if weighted else np.full(m, 1.0 / m)
        loss += 0.5 * float(np.sum(wt * (1.0 - c * c / (ww * ee))))
        alpha = 0.5 * wt * (-2.0 * c / (ww * ee))
        beta = 0.5 * wt * (2.0 * c * c / (ww * ww * ee))
        dev.combine(FP32, B4, 2 * m, np.vstack([np.diag(alpha), np.diag(beta)]), B4, 2 * m, m)
        QtD = dev.gram(FP32, B4, m, B4, m, 0, m)
        dev.combine(FP32, B4, 2 * m, np.vstack([-QtD, np.eye(m)]), B4, 0, 3 * m)
        grad += d.lp_backward_layer_device(h, R, B4, m, 3 * m).astype(np.float64)
    return loss, grad


def train_fo(d, snaps, theta0=None, steps=60, lr=0.02, reg=0.0, log=None, weighted=False):
    """Adam over the first-order snapshots (one per step, cycling), pulled
    toward theta0 with weight reg.  Only the spectral-layer entries move."""
    th = np.zeros(LP_NPARAM) if theta0 is None else np.asarray(theta0, np.float64).copy()
    th0 = th.copy()
    mom, vel = np.zeros_like(th), np.zeros_like(th)
    b1, b2, eps = 0.9, 0.999, 1e-8
    hist = []
    t0 = time.perf_counter()
    for it in range(1, steps + 1):
        loss, g = loss_and_grad_fo(d, th.astype(np.float32), snaps[(it - 1) % len(snaps)],
                                   weighted)
        hist.append(loss)
        if reg > 0:
            g = g + 2.0 * reg * (th - th0)
        mom = b1 * mom + (1 - b1) * g
        vel = b2 * vel + (1 - b2) * g * g
        th -= lr * (mom / (1 - b1 ** it)) / (np.sqrt(vel / (1 - b2 ** it)) + eps)
        if log and (it % 10 == 0 or it == 1):
            log(f"step {it}: loss {loss:.5f} ({time.perf_counter() - t0:.1f} s)")
    return th.astype(np.float32), hist


def evaluate_fo(d, theta, snaps, weighted=False):
    return float(np.mean([loss_and_grad_fo(d, theta, s, weighted)[0] for s in snaps]))


def collect_fo(make_problem, nev, snapshots=(4, 8, 14, 22, 32, 48, 70, 100, 140), log=None):
    """Converged first-order solve plus truncated solves of the same problem
    (all FP64, first-order form); returns (problem, its device context,
    snapshots)."""
    ref = make_problem(maxiter=2000)
    ref.solve()
    Ustar, Vstar = ref.modes(nev)
    snaps = []
    d = ref.device_context()
    for k in snapshots:
        p = make_problem(maxiter=k)
        p.solve()
        U, V = p.modes(nev)
        snaps.append(make_snapshot_fo(d, U, V, p.eigenvalues(nev), Ustar, Vstar))
        p.close()
        if log:
            log(f"first-order snapshot after {k} iterations")
    return ref, d, snaps

