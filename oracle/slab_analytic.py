"""Deterministic one-speed slab solution — TEST INFRASTRUCTURE ONLY.

Oracle for the one-speed slab decks: a slab 0 <= x <= L (Sigma_t = 1/cm),
scattering ratio c, isotropic scattering, a normally incident pencil beam of
unit current at x = 0.  The scalar flux solves the Peierls equation

    phi(x) = exp(-x) + (c/2) int_0^L E1(|x - x'|) phi(x') dx'

and the leakages are

    T = exp(-L) + (c/2) int phi(x') E2(L - x') dx'
    R =           (c/2) int phi(x') E2(x')     dx'.

The reference's track-length kernel scores w*l with no volume division
(Estimator.cpp:22-25), so flux bin k is int_{bin k} phi dx.  Balance:
T + R + (1 - c) * sum_k flux_k = 1.

Method: phi piecewise linear on a uniform node grid, collocation at the nodes,
the E1 kernel integrated exactly against each linear piece with
int E1 = -E2, int u E1 = -u E2 - E3 (product integration; no quadrature of the
log singularity).  Extrapolation over h, h/2, h/4 removes the h^2 log h and
h^2 terms of the error.  The half-space albedo 1 - sqrt(1-c) H(1; c) from Chandrasekhar's H
function (h_function below, an independent method) pins thick slabs.
"""
from __future__ import annotations

import numpy as np
from scipy.special import expn


def _piece_weights(x, a, b, n_lo):
    """int_a^b E_n(|x - x'|) f(x') dx' for linear f, as weights (w_a, w_b) on f(a), f(b).

    x must lie outside (a, b) (nodes coincide with piece ends).  n_lo is the
    kernel order (1 for E1, 2 for E2).
    """
    x = np.asarray(x, dtype=float)
    right = x >= b
    u1 = np.where(right, x - b, a - x)       # distance to the near end
    u2 = np.where(right, x - a, b - x)       # distance to the far end
    n = n_lo
    # int_{u1}^{u2} E_n(u) du = E_{n+1}(u1) - E_{n+1}(u2)
    i0 = expn(n + 1, u1) - expn(n + 1, u2)
    # int_{u1}^{u2} u E_n(u) du = [u E_{n+1} + E_{n+2}]_{u2}^{u1}
    i1 = (u1 * expn(n + 1, u1) + expn(n + 2, u1)) - (u2 * expn(n + 1, u2) + expn(n + 2, u2))
    h = b - a
    # f at distance u from x, linear: f = f_near + (f_far - f_near) (u - u1)/h
    w_far = (i1 - u1 * i0) / h
    w_near = i0 - w_far
    w_a = np.where(right, w_far, w_near)
    w_b = np.where(right, w_near, w_far)
    return w_a, w_b


def _toeplitz_weights(n, h, order):
    """Piece weights of E_order against linear pieces of a uniform grid.

    For node i and piece [x_j, x_j+1] they depend only on k = i - j:
    returns WA[k + n], WB[k + n] (weights on phi_j and phi_j+1), k in [-n, n].
    """
    m = np.arange(n + 1, dtype=float)
    u1, u2 = m * h, (m + 1) * h
    i0 = expn(order + 1, u1) - expn(order + 1, u2)
    i1 = (u1 * expn(order + 1, u1) + expn(order + 2, u1)) - (u2 * expn(order + 1, u2) + expn(order + 2, u2))
    w_far = (i1 - u1 * i0) / h
    w_near = i0 - w_far
    k = np.arange(-n, n + 1)
    mm = np.where(k >= 1, k - 1, -k)
    WA = np.where(k >= 1, w_far[np.minimum(mm, n)], w_near[np.minimum(mm, n)])
    WB = np.where(k >= 1, w_near[np.minimum(mm, n)], w_far[np.minimum(mm, n)])
    return WA, WB


def _solve_nodes(L, c, n):
    x = np.linspace(0.0, L, n + 1)
    h = L / n
    WA, WB = _toeplitz_weights(n, h, 1)
    i = np.arange(n + 1)[:, None]
    l = np.arange(n + 1)[None, :]
    K = np.where(l <= n - 1, WA[np.clip(i - l, -n, n) + n], 0.0)
    K += np.where(l >= 1, WB[np.clip(i - l + 1, -n, n) + n], 0.0)
    A = np.eye(n + 1) - 0.5 * c * K
    phi = np.linalg.solve(A, np.exp(-x))
    # leakages: E2 kernel seen from the faces x = L (node n) and x = 0 (node 0)
    WA2, WB2 = _toeplitz_weights(n, h, 2)
    j = np.arange(n)
    T = np.exp(-L) + 0.5 * c * (np.sum(WA2[n - j + n] * phi[j]) + np.sum(WB2[n - j + n] * phi[j + 1]))
    R = 0.5 * c * (np.sum(WA2[-j + n] * phi[j]) + np.sum(WB2[-j + n] * phi[j + 1]))
    return x, phi, T, R


def _bin_integrals(x, phi, edges):
    """int phi over each bin of edges for the piecewise-linear phi on nodes x."""
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (phi[1:] + phi[:-1]) * np.diff(x))])
    at = np.interp(edges, x, cum)
    # exact for linear pieces when edges fall on nodes; interpolation otherwise
    return np.diff(at)


def _extrapolate(hs, vals):
    """Limit of V(h) = V + A h^2 log h + B h^2 from three grids (the boundary
    layers of phi make the leading error h^2 log h, not h^2)."""
    vals = np.asarray(vals, dtype=float)
    A = np.array([[1.0, h * h * np.log(h), h * h] for h in hs])
    flat = vals.reshape(3, -1)
    return np.linalg.solve(A, flat)[0].reshape(vals.shape[1:])


def slab(L: float, c: float, n_bins: int = 100, nodes_per_bin: int = 4):
    """T, R and the n_bins track-length flux bins (three-grid extrapolation)."""
    n = max(n_bins * nodes_per_bin, int(np.ceil(L * 32)))
    n += (-n) % n_bins
    edges = np.linspace(0.0, L, n_bins + 1)
    hs, TR, FL = [], [], []
    for m in (n, 2 * n, 4 * n):
        x, phi, T, R = _solve_nodes(L, c, m)
        hs.append(L / m)
        TR.append([T, R])
        FL.append(_bin_integrals(x, phi, edges))
    T, R = _extrapolate(hs, TR)
    return dict(T=float(T), R=float(R), flux=_extrapolate(hs, FL))


def slab_leakage(L: float, c: float, nodes_per_mfp: int = 16):
    """T and R only (no flux bins), for the batched sweep of many slabs."""
    n = max(64, int(np.ceil(L * nodes_per_mfp)))
    hs, TR = [], []
    for m in (n, 2 * n, 4 * n):
        hs.append(L / m)
        TR.append(_solve_nodes(L, c, m)[2:])
    T, R = _extrapolate(hs, TR)
    return float(T), float(R)


def h_function(c: float, mu, n_quad: int = 200, iters: int = 500):
    """Chandrasekhar's H(mu) for isotropic scattering with albedo c.

    Solves 1/H(mu) = sqrt(1-c) + (c/2) int_0^1 mu' H(mu') / (mu + mu') dmu'
    by fixed-point iteration on Gauss-Legendre nodes.
    """
    g, w = np.polynomial.legendre.leggauss(n_quad)
    m = 0.5 * (g + 1.0)
    w = 0.5 * w
    H = np.ones_like(m)
    s = np.sqrt(1.0 - c)
    for _ in range(iters):
        Hn = 1.0 / (s + 0.5 * c * np.sum(w * m * H / (m[:, None] + m), axis=1))
        if np.max(np.abs(Hn - H)) < 1e-15:
            H = Hn
            break
        H = Hn
    mu = np.atleast_1d(np.asarray(mu, dtype=float))
    return 1.0 / (s + 0.5 * c * np.sum(w * m * H / (mu[:, None] + m), axis=1))


def half_space_albedo(c: float, mu0: float = 1.0) -> float:
    """Total reflection of a beam at direction cosine mu0 on a half space."""
    return float(1.0 - np.sqrt(1.0 - c) * h_function(c, mu0)[0])
