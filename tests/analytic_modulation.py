"""Deterministic solutions of the spherically symmetric Parker modulation problem (test infrastructure).

The backward-in-time Parker SDE integrated by the engine estimates u(r, p) = E[w(p_exit)], where w is the
weight of the outer ("hot") boundary and 0 that of the inner one.  By the Feynman-Kac correspondence u
solves the steady Parker equation with those boundary values,

    (1/r^2) d/dr (r^2 kappa du/dr) - V du/dr + (2V / 3r) p du/dp = 0,    u(R_in) = 0, u(R_out) = w(p),

for a radial wind of constant speed V and an isotropic kappa(r, T).  Here it is solved by finite
differences: xi = ln r on a uniform grid, conservative second differences, and Crank-Nicolson in
s = ln p, marching from high to low momentum (the direction in which the equation is parabolic), from a
start at momenta so high that the wind does nothing and u is w(p) times the harmonic function of the
diffusion operator.  The MC spectrum and u at r = 1 au must agree within the Monte Carlo error.

Also the closed form that the reference's modulation benchmark plots (Fisk & Axford 1969,
benchmarks/main_postprocess_modulation_cartesian_parker.cc:9-31).  Substituting x = A r^(1-b) shows that
2/Gamma(eta) x^(eta/2) K_eta(2 sqrt x) solves (1/r^2)(r^2 kappa u')' - (2 V K / 3 r) u = 0 with kappa ~ r^b,
K = 3 + alpha (mu - 1), regular at r = 0 and 1 at infinity: no convection term, and the momentum
derivative frozen at the local slope of the unmodulated spectrum.  With convection=False, a kappa that
does not depend on p and w = p^-K the solver's solution separates into p^-K times that radial problem,
which pins the solver against the Bessel function.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.linalg import solve_banded
from scipy.special import gamma, kv

C = 10.0        # speed of light, code units
MC2 = 100.0     # proton rest energy, code units


def ekin(p):
    return np.sqrt((p * C) ** 2 + MC2 ** 2) - MC2


def speed(p):
    return p * C * C / np.sqrt((p * C) ** 2 + MC2 ** 2)


def mom(T):
    return np.sqrt(T * (T + 2.0 * MC2)) / C


def fisk_axford(T, r, kap0, T0, r0, a, b, V, mu, alpha):
    """Modulation factor j(r, T) / j_LIS(T) of the reference's benchmark post-processor:
    2 / Gamma(eta) x^(eta/2) K_eta(2 sqrt x), x = 2 (1 + alpha (mu - 1) / 3) / (1 - b)^2 V r / kappa(T, r),
    eta = 1 - 2 / (1 - b)."""
    kap = kap0 * (T / T0) ** a * (r / r0) ** b
    eta = 1.0 - 2.0 / (1.0 - b)
    x = (2.0 * (1.0 + alpha * (mu - 1.0) / 3.0) / (1.0 - b) ** 2) * (V * r / kap)
    return 2.0 / gamma(eta) * x ** (eta / 2.0) * kv(eta, 2.0 * np.sqrt(x))


def pde_solution(kappa, weight, V, R_in, R_out, p_eval, r_eval, n_r=2400, ds=2.0e-3, p_top_factor=1.0e4, convection=True):
    """u(r_eval, p) for every p in p_eval (ascending or not).

    kappa(r, p) -> isotropic diffusion coefficient (vectorised in r), weight(p) -> outer-boundary value."""
    p_eval = np.asarray(p_eval, dtype=float)
    xi = np.linspace(math.log(R_in), math.log(R_out), n_r)
    h = xi[1] - xi[0]
    r = np.exp(xi)
    rh = np.exp(0.5 * (xi[1:] + xi[:-1]))             # cell faces
    s_top = math.log(p_eval.max() * p_top_factor)
    s_bot = math.log(p_eval.min())
    n_s = int(math.ceil((s_top - s_bot) / ds))
    ds = (s_top - s_bot) / n_s
    # start: w(p) times the harmonic function of (1/r^2) d/dr (r^2 kappa d/dr) with h(R_in) = 0, h(R_out) = 1,
    # from the exact integral of r^2 kappa dh/dr = const at the top momentum
    p_top = math.exp(s_top)
    kr = kappa(r, p_top)
    g = 1.0 / (r * r * kr)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (g[1:] * r[1:] + g[:-1] * r[:-1]) * h)])
    u = weight(p_top) * cum / cum[-1]

    def operator(p):
        """Banded matrix of A in du/ds = A u: (3 / 2V) [V du/dxi - r^-2 d/dxi (r kappa du/dxi)] on interior nodes."""
        kf = kappa(rh, p) * rh                          # r kappa at the faces
        lo = np.zeros(n_r); di = np.zeros(n_r); up = np.zeros(n_r)
        c = 1.5 / V
        Vc = V if convection else 0.0
        inner = slice(1, n_r - 1)
        ri2 = 1.0 / (r[inner] ** 2)
        lo[inner] = c * (-Vc / (2 * h) - ri2 * kf[:-1] / h ** 2)
        up[inner] = c * (Vc / (2 * h) - ri2 * kf[1:] / h ** 2)
        di[inner] = c * (ri2 * (kf[1:] + kf[:-1]) / h ** 2)
        return lo, di, up

    order = np.argsort(-p_eval)
    out = np.empty(len(p_eval))
    k = 0
    s = s_top
    while k < len(order) and p_eval[order[k]] >= math.exp(s) * (1 - 1e-12):
        out[order[k]] = np.interp(math.log(r_eval), xi, u); k += 1
    for n in range(n_s):
        s_new = s - ds
        lo, di, up = operator(math.exp(s - 0.5 * ds))
        # s decreases: u_new - u = -ds * A u  ->  (I + ds/2 A) u_new = (I - ds/2 A) u   (Crank-Nicolson); the
        # diffusion part of A is positive semi-definite, so this is the stable direction
        rhs = u.copy()
        rhs[1:-1] -= 0.5 * ds * (lo[1:-1] * u[:-2] + di[1:-1] * u[1:-1] + up[1:-1] * u[2:])
        ab = np.zeros((3, n_r))
        ab[0, 1:] = 0.5 * ds * up[:-1]
        ab[1, :] = 1.0 + 0.5 * ds * di
        ab[2, :-1] = 0.5 * ds * lo[1:]
        ab[1, 0] = ab[1, -1] = 1.0
        ab[0, 1] = 0.0
        ab[2, -2] = 0.0
        rhs[0] = 0.0
        rhs[-1] = weight(math.exp(s_new))
        u_new = solve_banded((1, 1), ab, rhs)
        while k < len(order) and p_eval[order[k]] >= math.exp(s_new) * (1 - 1e-12):
            # linear interpolation in s between the two levels
            pe = math.log(p_eval[order[k]])
            f = (s - pe) / ds
            ue = (1 - f) * u + f * u_new
            out[order[k]] = np.interp(math.log(r_eval), xi, ue)
            k += 1
        u, s = u_new, s_new
    while k < len(order):
        out[order[k]] = np.interp(math.log(r_eval), xi, u); k += 1
    return out
