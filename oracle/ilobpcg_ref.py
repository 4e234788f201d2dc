"""numpy restatement of the device iLOBPCG for the BdG pencil.

TEST INFRASTRUCTURE ONLY (the CPU reference solver used by tests and the
CPU-baseline leg of bench.py).  The reference calls an external, un-vendored
iLOBPCG (pstuermer/BDG_LOBPCG src/solver.c:262, 417) on
A x = lambda B x, A = diag(K, M), B = Swap, x = [u; v].  Its iteration is not
available, so this module (and the device solver it mirrors) uses the
equivalent Hermitian-definite form of the same eigenproblem:

    K u = w v,  M v = w u   =>   (M K M) v = w^2 M v,   u = M v / w,

and runs block LOBPCG (Knyazev 2001) on the pencil (MKM, M) for the v half:

  * basis V = [X, W, P] with images MV and KMV kept by linear combination;
  * Rayleigh-Ritz: the Euclidean Gram V^H V orthonormalises the basis
    (Cholesky after diagonal scaling; eigen-decomposition with rank
    truncation when the basis is numerically dependent), then
    G_M = V^H M V is Cholesky-factored (M-orthonormalisation) and w^2 are
    the eigenvalues of L^-1 (V^H M K M V) L^-H, ascending.  Directions in
    the numerical null space of G_M (only when M itself is singular, e.g. a
    non-interacting gas) are eigenvectors with w = 0, kept as zero modes;
  * preconditioner W = T_M T_K R with T_K, T_M the reference's precondK /
    precondM (src/precond_impl.inc:21-111), i.e. an approximation of
    (M K M)^-1 M applied to R = KMx - w^2 x;
  * soft locking: converged columns get no new W / P; when [X, W, P] is
    numerically dependent the P block is dropped for one iteration, and when
    [X, W] still is, W falls back to the plain residual R for that iteration
    (guards against preconditioners that blow one direction up, e.g.
    safe_val(mu + 0) = 1e-8 at k = 0 for mu = 0);
  * a pair converges when ||A x - w B x|| <= tol * ||x|| * max(1, |w|), with
    x = [M v / w; v].  The Goldstone mode (|w| below zero_tol * scale) is
    reported as a zero mode (for a non-stationary psi0 its w^2 is slightly
    negative; it is then reported as -sqrt(-w^2)); it converges when its
    absolute residual ||K M v - w^2 v|| / (||v|| scale^2) is below tol.  Only
    directions in the numerical null space of the M-Gram skip the test.
"""
from __future__ import annotations

import numpy as np

TAU_CHOL = 1e-10    # relative squared pivot floor: basis accepted by Cholesky
TAU_RANK = 1e-12    # rank truncation of G_M in the eigen fallback


def _herm(a):
    return 0.5 * (a + a.conj().T)


def _orth_transform(G, tau_chol=TAU_CHOL, tau_rank=TAU_RANK):
    """T with T^H G T = I on the numerically independent part of a Gram G."""
    dg = np.real(np.diag(G))
    dinv = np.where(dg > 0, 1.0 / np.sqrt(np.where(dg > 0, dg, 1.0)), 0.0)
    G1 = _herm(G * dinv[:, None] * dinv[None, :])
    try:
        L = np.linalg.cholesky(G1)
        if np.min(np.abs(np.diag(L))) ** 2 >= tau_chol:
            return dinv[:, None] * np.linalg.inv(L).conj().T, None
    except np.linalg.LinAlgError:
        pass
    s, U = np.linalg.eigh(G1)
    keep = s > tau_rank * s.max()
    null = dinv[:, None] * U[:, ~keep] if np.any(~keep) else None
    return dinv[:, None] * (U[:, keep] / np.sqrt(s[keep])), null


def rayleigh_ritz(GV, GM, GA, m):
    """Ritz pairs of (GA, GM) on a basis with Euclidean Gram GV.
    Returns (w2[m'], C[k, m'], zero_flags[m'])."""
    TV, _ = _orth_transform(_herm(GV))          # drop dependent directions
    GMp = _herm(TV.conj().T @ GM @ TV)
    GAp = _herm(TV.conj().T @ GA @ TV)
    T2, null = _orth_transform(GMp, tau_chol=1e-14, tau_rank=1e-12)
    C = _herm(T2.conj().T @ GAp @ T2)
    w2, Z = np.linalg.eigh(C)
    Y = TV @ (T2 @ Z)
    flags = np.zeros(len(w2), dtype=bool)
    if null is not None:
        nn = TV @ null
        nn = nn / np.linalg.norm(nn, axis=0)
        w2 = np.concatenate([np.zeros(nn.shape[1]), w2])
        Y = np.concatenate([nn, Y], axis=1)
        flags = np.concatenate([np.ones(nn.shape[1], dtype=bool), flags])
        order = np.argsort(w2, kind="stable")
        w2, Y, flags = w2[order], Y[:, order], flags[order]
    return w2[:m], Y[:, :m], flags[:m]


def gram_ok(GV):
    dg = np.real(np.diag(GV))
    if np.any(dg <= 0):
        return False
    d = 1.0 / np.sqrt(dg)
    try:
        L = np.linalg.cholesky(_herm(GV * d[:, None] * d[None, :]))
    except np.linalg.LinAlgError:
        return False
    return float(np.min(np.abs(np.diag(L)))) ** 2 >= TAU_CHOL


def solver_precond_weights(ltK, ltM, mu):
    """The solver's preconditioner weights: the reference's
    1/sqrt(max(1e-8, lt)) (src/setup.c:165-166) on the pre-mu local term,
    with the floor raised to mu (same rule as the device solve)."""
    fl = mu if mu > 1e-8 else 1e-8
    return (1.0 / np.sqrt(np.maximum(fl, np.asarray(ltK) + mu)),
            1.0 / np.sqrt(np.maximum(fl, np.asarray(ltM) + mu)))


def signed_sqrt(w2):
    return np.sign(w2) * np.sqrt(np.abs(w2))


def pencil_residual(rn, xv_norm, mv_norm, w):
    """||A x - w B x|| / (||x|| max(1, |w|)) for x = [M v / w; v], given
    rn = ||K M v - w^2 v||, ||v|| and ||M v||."""
    aw = np.maximum(np.abs(w), 1e-300)
    xnorm = np.sqrt(xv_norm ** 2 + (mv_norm / aw) ** 2)
    return (rn / aw) / (xnorm * np.maximum(1.0, np.abs(w)))


def column_residuals(rn, xv_norm, mv_norm, w, scale, zero_tol):
    """Per-column convergence measure: the pencil residual, except for soft
    modes (|w| < zero_tol * scale), whose 1/|w| scaling degenerates; those
    take the absolute product-form residual ||K M v - w^2 v|| / (||v|| scale^2)."""
    soft = np.abs(w) < zero_tol * scale
    absr = rn / (np.maximum(xv_norm, 1e-300) * scale * scale)
    return np.where(soft, absr, pencil_residual(rn, xv_norm, mv_norm, w))


def ilobpcg(applyK, applyM, applyTK, applyTM, V0, nev, tol=1e-8, maxiter=500,
            zero_tol=1e-3, callback=None):
    """LOBPCG on (MKM, M).  apply* map an S x c block to an S x c block.
    V0: S x m start block (the v half of the reference's start block).
    Returns w (signed frequencies), the u and v halves of the modes, etc."""
    X = np.array(V0, copy=True)
    m = X.shape[1]
    MX = applyM(X)
    KMX = applyK(MX)
    w2, Cc, flags = rayleigh_ritz(X.conj().T @ X, X.conj().T @ MX, MX.conj().T @ KMX, m)
    X, MX, KMX = X @ Cc, MX @ Cc, KMX @ Cc
    P = MP = KMP = None
    it = 0
    res = None
    hist = []
    restarts = 0
    precond_fallbacks = 0
    applies = m
    for it in range(1, maxiter + 1):
        R = KMX - X * w2[None, :]
        w = signed_sqrt(w2)
        scale = max(1.0, float(np.max(np.abs(w[:nev]))))
        res = column_residuals(np.linalg.norm(R, axis=0), np.linalg.norm(X, axis=0),
                               np.linalg.norm(MX, axis=0), w, scale, zero_tol)
        zero = flags | (np.abs(w) < zero_tol * scale)
        # only the numerical null space of the M-Gram bypasses the test
        conv = (res < tol) | flags
        hist.append(float(np.max(np.where(flags[:nev], 0.0, res[:nev]))))
        if callback:
            callback(it, w, res)
        if np.all(conv[:nev]):
            break
        act = np.nonzero(~conv)[0]
        W = applyTM(applyTK(R[:, act]))
        MW = applyM(W)
        KMW = applyK(MW)
        applies += len(act)
        blocks = [(X, MX, KMX), (W, MW, KMW)]
        if P is not None:
            blocks.append((P[:, act], MP[:, act], KMP[:, act]))
        fallback = False
        while True:
            V = np.concatenate([b[0] for b in blocks], axis=1)
            GV = V.conj().T @ V
            if gram_ok(GV):
                break
            if len(blocks) == 3:
                blocks = blocks[:2]
                restarts += 1
                continue
            if fallback:
                break
            # the preconditioned block is numerically dependent on X (e.g. a
            # preconditioner that amplifies one direction by 1e8): use the
            # plain residual for this iteration
            W = R[:, act]
            MW = applyM(W)
            KMW = applyK(MW)
            applies += len(act)
            blocks = [(X, MX, KMX), (W, MW, KMW)]
            fallback = True
            precond_fallbacks += 1
        MV = np.concatenate([b[1] for b in blocks], axis=1)
        KMV = np.concatenate([b[2] for b in blocks], axis=1)
        w2, Cc, flags = rayleigh_ritz(GV, V.conj().T @ MV, MV.conj().T @ KMV, m)
        if Cc.shape[1] < m:
            raise RuntimeError("Rayleigh-Ritz lost rank")
        Xn, MXn, KMXn = V @ Cc, MV @ Cc, KMV @ Cc
        P, MP, KMP = V[:, m:] @ Cc[m:], MV[:, m:] @ Cc[m:], KMV[:, m:] @ Cc[m:]
        X, MX, KMX = Xn, MXn, KMXn
    w = signed_sqrt(w2)
    scale = max(1.0, float(np.max(np.abs(w[:nev]))))
    zero = flags | (np.abs(w) < zero_tol * scale)
    safe = np.where(np.abs(w) > 0, w, 1.0)
    U = MX / safe[None, :]
    nrm = np.sqrt(np.linalg.norm(U, axis=0) ** 2 + np.linalg.norm(X, axis=0) ** 2)
    return dict(w=w[:nev], v=(X / nrm)[:, :nev], u=(U / nrm)[:, :nev], iters=it,
                res=res[:nev], hist=hist, zero=zero[:nev], restarts=restarts,
                precond_fallbacks=precond_fallbacks,
                applies=applies, w2=w2[:nev])
