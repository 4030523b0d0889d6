"""FP64 restatement of the solve's fixed point and its adjoint (test helper).

The forward of src/solve.rs:47-108 converges to irr = M (E + rho (.) irr) with
M = W P: W is the N x C link matrix of the transfer CSR
(src/formfactor.rs:244-270) and P the C x N area-weighted cluster mean
(src/solve.rs:60-72); E_i = rho_i (.) D_i with D_i = sum_l e_l g_il
(src/shade.rs:137-162).  Everything here is solved directly in float64 with
scipy, channel by channel, so it shares no code path with the device's
iterative float32 adjoint.  Its own formulas are pinned against float64
central differences in test_adjoint_math.py.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl


def build_operators(starts, idx, w, mstart, members, surf_area, cl_area, n):
    """W (N x C) from the CSR rows, P (C x N) with P[c, m] = a_m / max(A_c, 1e-12)."""
    nc = len(cl_area)
    W = sp.csr_matrix((np.asarray(w, np.float64), np.asarray(idx, np.int64), np.asarray(starts, np.int64)),
                      shape=(n, nc))
    rows = np.repeat(np.arange(nc), np.diff(np.asarray(mstart, np.int64)))
    members = np.asarray(members, np.int64)
    vals = np.asarray(surf_area, np.float64)[members] / np.maximum(np.asarray(cl_area, np.float64)[rows], 1e-12)
    P = sp.csr_matrix((vals, (rows, members)), shape=(nc, n))
    return W, P


def direct_factor(g, emit):
    """D_i = sum_l e_l g_il for g (nl x N) and emit (nl x 3)."""
    return np.asarray(g, np.float64).T @ np.asarray(emit, np.float64)


def forward(W, P, rho, D):
    """Fixed point irr of irr = M (rho (.) D + rho (.) irr), per channel."""
    M = (W @ P).tocsc()
    n = M.shape[0]
    I = sp.identity(n, format="csc")
    irr = np.zeros((n, 3))
    for c in range(3):
        r = np.asarray(rho[:, c], np.float64)
        irr[:, c] = spl.spsolve(I - M @ sp.diags(r), M @ (r * D[:, c]))
    return irr


def adjoint(W, P, rho, D, irr, ubar, mats, g, nm):
    """mu = M^T (ubar + rho (.) mu) and the parameter gradients.

    Returns (mu, d_albedo[nm, 3], d_emit[nl, 3]) for L with dL/dirr = ubar:
    dL/drho_i = mu_i (.) (D_i + irr_i), summed over the surfels of each
    material, and dL/de_l = sum_i mu_i (.) rho_i g_il."""
    M = (W @ P).tocsc()
    n = M.shape[0]
    I = sp.identity(n, format="csc")
    MT = M.T.tocsc()
    mu = np.zeros((n, 3))
    for c in range(3):
        r = np.asarray(rho[:, c], np.float64)
        mu[:, c] = spl.spsolve(I - MT @ sp.diags(r), MT @ np.asarray(ubar[:, c], np.float64))
    return (mu,) + params(mu, rho, D, irr, mats, g, nm)


def params(mu, rho, D, irr, mats, g, nm):
    """dL/drho summed per material and dL/de from the adjoint state mu."""
    drho = mu * (D + irr)
    d_alb = np.zeros((nm, 3))
    np.add.at(d_alb, np.asarray(mats, np.int64), drho)
    g = np.asarray(g, np.float64)
    d_emit = g @ (mu * np.asarray(rho, np.float64))
    return d_alb, d_emit


def _fixed_point(step, n, tol, max_iter):
    x = np.zeros((n, 3))
    for _ in range(max_iter):
        y = step(x)
        if np.abs(y - x).max() <= tol * max(np.abs(y).max(), 1e-300):
            return y
        x = y
    raise RuntimeError("float64 fixed point did not converge in %d iterations" % max_iter)


def forward_iterative(W, P, rho, D, tol=1e-12, max_iter=20000):
    """forward() by float64 Jacobi on the factors W and P (never forming
    W P, which fills in at tens of thousands of surfels)."""
    rho = np.asarray(rho, np.float64)
    E = rho * D
    return _fixed_point(lambda x: W @ (P @ (E + rho * x)), W.shape[0], tol, max_iter)


def adjoint_iterative(W, P, rho, D, irr, ubar, mats, g, nm, tol=1e-12, max_iter=20000):
    """adjoint() by float64 Jacobi on the factors: mu = P^T W^T (ubar + rho (.) mu)."""
    rho = np.asarray(rho, np.float64)
    ubar = np.asarray(ubar, np.float64)
    WT, PT = W.T.tocsr(), P.T.tocsr()
    mu = _fixed_point(lambda x: PT @ (WT @ (ubar + rho * x)), W.shape[0], tol, max_iter)
    return (mu,) + params(mu, rho, D, irr, mats, g, nm)
