"""Pins the float64 adjoint helper (tests/adjoint64.py) on the CPU.

A random small instance of the solve's structure (sparse links, clusters
that partition the surfels, materials, lights) is differentiated by float64
central differences of the directly solved fixed point; the helper's
implicit-adjoint gradients must agree.  A planted mistake (dropping the
rho (.) mu term of the adjoint recursion) must be caught."""
import numpy as np
import scipy.sparse as sp

import adjoint64 as a64


def random_instance(seed, n=60, nc=9, nm=3, nl=2):
    rng = np.random.default_rng(seed)
    starts = [0]
    idx, w = [], []
    for i in range(n):
        cols = np.sort(rng.choice(nc, size=rng.integers(1, nc), replace=False))
        wt = rng.random(len(cols))
        wt *= rng.uniform(0.3, 0.9) / wt.sum()  # row sums below one, as the transfer's
        idx += list(cols)
        w += list(wt)
        starts.append(len(idx))
    cl_of = rng.integers(0, nc, size=n)
    cl_of[:nc] = np.arange(nc)  # no empty cluster
    order = np.argsort(cl_of, kind="stable")
    mstart = np.concatenate([[0], np.cumsum(np.bincount(cl_of, minlength=nc))])
    surf_area = rng.uniform(0.5, 1.5, n)
    cl_area = np.bincount(cl_of, weights=surf_area, minlength=nc)
    mats = rng.integers(0, nm, size=n)
    alb = rng.uniform(0.1, 0.9, (nm, 3))
    g = rng.random((nl, n)) * 0.3
    emit = rng.uniform(5, 20, (nl, 3))
    W, P = a64.build_operators(np.array(starts), np.array(idx), np.array(w), mstart, order, surf_area, cl_area, n)
    return W, P, mats, alb, g, emit


def loss(W, P, mats, alb, g, emit, c):
    rho = alb[mats]
    D = a64.direct_factor(g, emit)
    irr = a64.forward(W, P, rho, D)
    return float(np.sum(c * irr) + 0.5 * np.sum(irr * irr))


def grads(W, P, mats, alb, g, emit, c, drop_rho_mu=False):
    rho = alb[mats]
    D = a64.direct_factor(g, emit)
    irr = a64.forward(W, P, rho, D)
    ubar = c + irr
    if not drop_rho_mu:
        return a64.adjoint(W, P, rho, D, irr, ubar, mats, g, alb.shape[0])[1:]
    MT = (W @ P).T.tocsr()
    mu = np.stack([MT @ ubar[:, k] for k in range(3)], 1)  # planted: mu = M^T ubar only
    d_alb = np.zeros_like(alb)
    np.add.at(d_alb, mats, mu * (D + irr))
    return d_alb, g @ (mu * rho)


def fd(W, P, mats, alb, g, emit, c, eps=1e-6):
    d_alb = np.zeros_like(alb)
    for m in range(alb.shape[0]):
        for k in range(3):
            a1, a0 = alb.copy(), alb.copy()
            a1[m, k] += eps
            a0[m, k] -= eps
            d_alb[m, k] = (loss(W, P, mats, a1, g, emit, c) - loss(W, P, mats, a0, g, emit, c)) / (2 * eps)
    d_emit = np.zeros_like(emit)
    for l in range(emit.shape[0]):
        for k in range(3):
            e1, e0 = emit.copy(), emit.copy()
            e1[l, k] += eps
            e0[l, k] -= eps
            d_emit[l, k] = (loss(W, P, mats, alb, g, e1, c) - loss(W, P, mats, alb, g, e0, c)) / (2 * eps)
    return d_alb, d_emit


def test_operators_match_definitions():
    W, P, mats, alb, g, emit = random_instance(3)
    # P rows are area-weighted means: each row sums to one
    assert np.allclose(np.asarray(P.sum(1)).ravel(), 1.0)
    rho = alb[mats]
    D = a64.direct_factor(g, emit)
    irr = a64.forward(W, P, rho, D)
    # fixed point of the Jacobi map of src/solve.rs:60-80
    B = rho * D + rho * irr
    assert np.allclose(irr, np.stack([W @ (P @ B[:, k]) for k in range(3)], 1), rtol=1e-12, atol=1e-12)
    # and the Jacobi iteration itself converges to it
    x = np.zeros_like(irr)
    for _ in range(400):
        x = np.stack([W @ (P @ (rho * D + rho * x)[:, k]) for k in range(3)], 1)
    assert np.allclose(x, irr, rtol=1e-10, atol=1e-10)


def test_adjoint_matches_central_differences():
    for seed in (1, 2):
        W, P, mats, alb, g, emit = random_instance(seed)
        c = np.random.default_rng(seed + 10).normal(size=(W.shape[0], 3))
        ga, ge = grads(W, P, mats, alb, g, emit, c)
        fa, fe = fd(W, P, mats, alb, g, emit, c)
        assert np.allclose(ga, fa, rtol=1e-6, atol=1e-6 * np.abs(fa).max())
        assert np.allclose(ge, fe, rtol=1e-6, atol=1e-6 * np.abs(fe).max())


def test_planted_mistake_is_caught():
    W, P, mats, alb, g, emit = random_instance(4)
    c = np.random.default_rng(14).normal(size=(W.shape[0], 3))
    fa, fe = fd(W, P, mats, alb, g, emit, c)
    ba, be = grads(W, P, mats, alb, g, emit, c, drop_rho_mu=True)
    assert not np.allclose(ba, fa, rtol=1e-3, atol=1e-3 * np.abs(fa).max())


def test_sparse_shapes():
    W, P, *_ = random_instance(5)
    assert sp.issparse(W) and W.shape == (60, 9) and P.shape == (9, 60)


def test_iterative_solves_match_the_direct_ones():
    """The factored float64 Jacobi forms used at sizes where W P cannot be
    formed agree with the direct sparse solves."""
    for seed in (6, 7):
        W, P, mats, alb, g, emit = random_instance(seed, n=200, nc=17)
        rho = alb[mats]
        D = a64.direct_factor(g, emit)
        irr = a64.forward(W, P, rho, D)
        irr_it = a64.forward_iterative(W, P, rho, D, tol=1e-14)
        assert np.allclose(irr_it, irr, rtol=1e-11, atol=1e-11 * np.abs(irr).max())
        ubar = np.random.default_rng(seed).normal(size=(200, 3))
        mu, da, de = a64.adjoint(W, P, rho, D, irr, ubar, mats, g, alb.shape[0])
        mu2, da2, de2 = a64.adjoint_iterative(W, P, rho, D, irr, ubar, mats, g, alb.shape[0], tol=1e-14)
        assert np.allclose(mu2, mu, rtol=1e-11, atol=1e-11 * np.abs(mu).max())
        assert np.allclose(da2, da, rtol=1e-10, atol=1e-10 * np.abs(da).max())
        assert np.allclose(de2, de, rtol=1e-10, atol=1e-10 * np.abs(de).max())
    # transposing on the wrong side (W P instead of its transpose) is caught
    W, P, mats, alb, g, emit = random_instance(8, n=200, nc=17)
    rho = alb[mats]
    ubar = np.random.default_rng(8).normal(size=(200, 3))
    D = a64.direct_factor(g, emit)
    irr = a64.forward(W, P, rho, D)
    mu = a64.adjoint(W, P, rho, D, irr, ubar, mats, g, alb.shape[0])[0]
    wrong = a64._fixed_point(lambda x: W @ (P @ (ubar + rho * x)), 200, 1e-14, 20000)
    assert not np.allclose(wrong, mu, rtol=1e-3, atol=1e-3 * np.abs(mu).max())
