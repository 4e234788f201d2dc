"""numpy restatement of the device first-order solver (bdg_lobpcg_amd
csrc/host/lrep.c, csrc/host/lrep_impl.inc).

TEST INFRASTRUCTURE ONLY (CPU reference of the first-order iteration for
tests).  The reference hands A = diag(K, M), B = Swap, T = [T_K; T_M] to an
external, un-vendored iLOBPCG (pstuermer/BDG_LOBPCG src/solver.c:147-164,
262); its iteration is not available.  This module states the build's own
first-order iteration:

  * separate search spaces for the two halves of x = [u; v]:
    Z_U = [X_U, P_U, W_U] with images K Z_U, Z_V = [X_V, P_V, W_V] with M Z_V,
    each kept orthonormal (Euclidean);
  * Rayleigh-Ritz of the Galerkin conditions Z_U^H (K u - w v) = 0,
    Z_V^H (M v - w u) = 0 through Hermitian eigenproblems only (rr below),
    with the uncoupled directions of Z_U handled by a Schur complement;
  * W_U = T_K R_U, W_V = T_M R_V for the residuals R_U = K u - s|w| v,
    R_V = M v - |w| u (s = sign w^2; w^2 < 0 is a pair with imaginary w),
    orthonormalised against [X, P] before their images are formed;
  * X, P updated with orthonormal coefficients [Q1 | Q2]: A = Q1 R_A,
    Q2 = the part of the active Ritz coefficients (X rows zeroed) outside
    span(A); Ritz vectors u = X_U R_A, v = X_V R_B;
  * a pair converges when ||A x - w B x|| <= tol ||x|| max(1, |w|).
Storage rounding is injectable (`rnd`, e.g. to float32) to state the FP32
phase of the mixed policy.
"""
from __future__ import annotations

import numpy as np


def _h(a):
    return 0.5 * (a + a.conj().T)


def orth_coef(GQW, GWW, drop):
    """C with [Q | W] C orthonormal and orthogonal to Q (Q orthonormal), from
    GQW = Q^H W, GWW = W^H W; directions keeping < drop of their norm left out."""
    nq, nw = GQW.shape
    d = np.where(np.real(np.diag(GWW)) > 0, 1.0 / np.sqrt(np.abs(np.real(np.diag(GWW)))), 0.0)
    G = (GWW - GQW.conj().T @ GQW) * d[:, None] * d[None, :]
    lam, V = np.linalg.eigh(_h(G))
    keep = lam > drop * drop
    T = d[:, None] * (V[:, keep] / np.sqrt(lam[keep])[None, :])
    return np.concatenate([-GQW @ T, T], axis=0)


def rr(GK, GM, GW, m):
    """Ritz pairs of G_K a = w G_W b, G_M b = w G_W^H a.
    Returns (w2[m], A (ku x m), B (kv x m)), ||a||^2 + ||b||^2 = 1."""
    lam, QM = np.linalg.eigh(_h(GM))
    keep = lam > 1e-12 * lam.max()
    Li = QM[:, keep] / np.sqrt(lam[keep])[None, :]
    C = GW @ Li
    s2, Pm = np.linalg.eigh(_h(C @ C.conj().T))
    cpl = s2 > 1e-12 * s2.max()
    P, s = Pm[:, cpl], np.sqrt(s2[cpl])
    Pn = Pm[:, ~cpl]
    Q = (C.conj().T @ P) / s[None, :]
    N = P.conj().T @ GK @ P
    if Pn.shape[1]:
        G12 = P.conj().T @ GK @ Pn
        G22i = np.linalg.pinv(_h(Pn.conj().T @ GK @ Pn), rcond=1e-12, hermitian=True)
        N = N - G12 @ G22i @ G12.conj().T
    N = N / s[:, None] / s[None, :]
    w2, Y = np.linalg.eigh(_h(N))
    w2, Y = w2[:m], Y[:, :m]
    al = Y / s[:, None]
    a = P @ al
    if Pn.shape[1]:
        a = a - Pn @ (G22i @ (G12.conj().T @ al))
    b = (Li @ (Q @ Y)) * np.sqrt(np.abs(w2))[None, :]
    nrm = np.sqrt((np.abs(a) ** 2).sum(0) + (np.abs(b) ** 2).sum(0))
    return w2, a / nrm, b / nrm


def update_coef(A, act, m, drop=1e-3):
    """[Q1 | Q2] orthonormal with A = Q1 R11, Q2 spanning the part of
    A[m:, act] (first m rows zeroed) outside span(A)."""
    Q1, R11 = np.linalg.qr(A)
    A0 = A[:, act].copy()
    A0[:m] = 0
    n0 = np.linalg.norm(A0, axis=0)
    P = A0 - Q1 @ (Q1.conj().T @ A0)
    P = P - Q1 @ (Q1.conj().T @ P)
    cols = []
    for j in range(P.shape[1]):
        v = P[:, j]
        for q in cols:
            v = v - q * (q.conj() @ v)
        nv = np.linalg.norm(v)
        if n0[j] > 0 and nv > drop * n0[j]:
            cols.append(v / nv)
    Q2 = np.stack(cols, axis=1) if cols else np.zeros((A.shape[0], 0), A.dtype)
    return np.concatenate([Q1, Q2], axis=1), R11


def lrep_lobpcg(applyK, applyM, applyTK, applyTM, U0, V0, nev, tol=1e-8, maxiter=500,
                rnd=None, drop=1e-3):
    """First-order LOBPCG from the start blocks U0, V0 (S x m).  Returns the
    signed frequencies w (w^2 < 0 reported as -sqrt(-w^2)), the modes u, v
    (M v = w u), residuals and the iteration count."""
    rnd = rnd or (lambda X: X)
    f = lambda X: np.asarray(X, dtype=np.complex128 if np.iscomplexobj(X) else np.float64)
    m = U0.shape[1]

    def orth(W, Q):
        W = f(rnd(W))
        GQW = f(Q).conj().T @ W if Q is not None else np.zeros((0, W.shape[1]), W.dtype)
        C = orth_coef(GQW, W.conj().T @ W, drop)
        Z = W if Q is None else np.concatenate([f(Q), W], axis=1)
        return rnd(Z @ C)

    XU, XV = orth(U0, None), orth(V0, None)
    KXU, MXV = rnd(applyK(XU)), rnd(applyM(XV))
    ZU, ZV, KZU, MZV = XU, XV, KXU, MXV
    act = np.zeros(0, dtype=int)
    it = 0
    for it in range(maxiter + 1):
        GK, GM, GW = (f(ZU).conj().T @ f(KZU), f(ZV).conj().T @ f(MZV),
                      f(ZU).conj().T @ f(ZV))
        w2, A, Bc = rr(GK, GM, GW, m)
        CU, RA = update_coef(A, act, m)
        CV, RB = update_coef(Bc, act, m)
        XU, PU = rnd(f(ZU) @ CU[:, :m]), rnd(f(ZU) @ CU[:, m:])
        KXU, KPU = rnd(f(KZU) @ CU[:, :m]), rnd(f(KZU) @ CU[:, m:])
        XV, PV = rnd(f(ZV) @ CV[:, :m]), rnd(f(ZV) @ CV[:, m:])
        MXV, MPV = rnd(f(MZV) @ CV[:, :m]), rnd(f(MZV) @ CV[:, m:])
        wt, sg = np.sqrt(np.abs(w2)), np.sign(w2)
        u, v = f(XU) @ RA, f(XV) @ RB
        RU = f(KXU) @ RA - v * (sg * wt)[None, :]
        RV = f(MXV) @ RB - u * wt[None, :]
        xn = np.sqrt((np.abs(u) ** 2).sum(0) + (np.abs(v) ** 2).sum(0))
        res = np.sqrt((np.abs(RU) ** 2).sum(0) + (np.abs(RV) ** 2).sum(0)) / (
            xn * np.maximum(1.0, wt))
        conv = res < tol
        if np.all(conv[:nev]) or it == maxiter:
            break
        act = np.nonzero(~conv)[0]
        QU = np.concatenate([f(XU), f(PU)], axis=1)
        QV = np.concatenate([f(XV), f(PV)], axis=1)
        WU, WV = orth(applyTK(RU[:, act]), QU), orth(applyTM(RV[:, act]), QV)
        ZU, KZU = np.concatenate([QU, f(WU)], 1), np.concatenate([f(KXU), f(KPU), f(applyK(WU))], 1)
        ZV, MZV = np.concatenate([QV, f(WV)], 1), np.concatenate([f(MXV), f(MPV), f(applyM(WV))], 1)
        ZU, KZU, ZV, MZV = rnd(ZU), rnd(KZU), rnd(ZV), rnd(MZV)
    u = np.where(sg[None, :] < 0, -u, u)
    w = sg * wt
    nrm = np.sqrt((np.abs(u) ** 2).sum(0) + (np.abs(v) ** 2).sum(0))
    return dict(w=w[:nev], u=(u / nrm)[:, :nev], v=(v / nrm)[:, :nev], res=res[:nev],
                iters=it, w2=w2[:nev])
