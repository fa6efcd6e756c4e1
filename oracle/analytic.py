"""Closed forms and exact integrals the Monte Carlo is checked against.

TEST INFRASTRUCTURE ONLY — imported by tests/ as the checker, never by the
product path.

  analy_tl_iso          Huntley's (2006) tunnelling-only decay, the reference's
                        analy_TL_iso (src/est_params/functions.py:264-268):
                        n(t) = n0 exp(-rho' ln(1.8 s t)^3).
  nn_survival           the exact surviving fraction of the same model it
                        approximates: each electron recombines with its nearest
                        hole at rate s e^{-alpha r}, holes Poisson-distributed
                        at density rho = rho' 3 alpha^3 / (4 pi) in an infinite
                        medium, no hole shared or used up.
  geometry_survival     the same expectation over a given set of nearest-hole
                        factors w = e^{-alpha r} (one cluster's actual geometry).
  fill_fraction         the dose-filling law of TLTrapSim._filling_time
                        (src/class/tl_trap_lab.py:53-60) with recombination off:
                        each of the N_e - n_e empty traps fills at rate D / D0,
                        so n_e(t) / N_e = 1 - (1 - f0) e^{-(D/D0) t}.
  fill_observed_moments mean and variance of what the lab loops record at an
                        observation time t under that law: the loops record
                        after the first event at or past t (tl_trap_lab.py:90-108,
                        147-172), which with recombination off is always one
                        more fill, even with every trap full (lambda = 1e-20 still
                        draws a fill, some 1e20 s later, tl_trap_lab.py:56-60):
                        n_e(0) + X + 1 with X ~ Binomial(N_e - n_e(0),
                        1 - e^{-(D/D0) t}).

Huntley's form replaces the survival factor exp(-s t e^{-u}) (u = alpha r) by a
step at u* = ln(1.8 s t); the factor 1.8 is e^gamma (gamma = 0.5772, Euler's
constant), the shift that makes the step carry the same mean as the
Gumbel-shaped factor.  Its error against nn_survival grows with rho': below
0.004 in n / n0 for rho' <= 1e-4 over s t in [1, 1e10], about 0.017 at
rho' = 1e-3 (tests/test_oracle_tunnelling.py states and checks these).
"""
from __future__ import annotations

import math

import numpy as np


def analy_tl_iso(n0: float, rho_prime: float, s: float, t) -> np.ndarray:
    """n0 exp(-rho' ln(1.8 s t)^3) (functions.py:264-268: mc.electrons,
    mc.rho_prime, phys.s, timebin)."""
    t = np.asarray(t, np.float64)
    return n0 * np.exp(-rho_prime * np.log(1.8 * s * t) ** 3)


def nn_survival(rho_prime: float, s: float, t) -> np.ndarray:
    """Exact surviving fraction at times t of the nearest-neighbour tunnelling
    model in an infinite medium.

    With u = alpha r the nearest-hole distance has density
    3 rho' u^2 exp(-rho' u^3) (Poisson holes: P(no hole within r) =
    exp(-rho 4 pi r^3 / 3) = exp(-rho' u^3)), and an electron at u survives to t
    with probability exp(-s t e^{-u}); the fraction is the integral of their
    product over u, computed by adaptive quadrature split at the step
    u* = ln(s t)."""
    from scipy import integrate
    out = []
    u_max = (60.0 / rho_prime) ** (1.0 / 3.0)           # density below e^-60 beyond
    for tt in np.atleast_1d(np.asarray(t, np.float64)):
        st = s * tt

        def f(u):
            return 3.0 * rho_prime * u * u * math.exp(-rho_prime * u ** 3 - st * math.exp(-u))

        u_star = min(max(math.log(max(st, 1e-300)), 0.0), 0.99 * u_max)
        a, _ = integrate.quad(f, 0.0, u_star, limit=400) if u_star > 0 else (0.0, 0.0)
        b, _ = integrate.quad(f, u_star, u_max, limit=400)
        out.append(a + b)
    return np.array(out)


def geometry_survival(w: np.ndarray, s: float, t) -> np.ndarray:
    """Mean over electrons of exp(-s t w_e), w_e = e^{-alpha r_e} of each electron's
    initial nearest hole: the expected surviving fraction of the given geometry
    when no electron loses its nearest hole to another."""
    w = np.asarray(w, np.float64)
    t = np.atleast_1d(np.asarray(t, np.float64))
    return np.array([np.mean(np.exp(-s * tt * w)) for tt in t])


def fill_fraction(f0: float, D: float, D0: float, t) -> np.ndarray:
    """1 - (1 - f0) e^{-(D/D0) t}: the mean trapped fraction of a pure filling
    process started at fraction f0."""
    t = np.asarray(t, np.float64)
    return 1.0 - (1.0 - f0) * np.exp(-(D / D0) * t)


def fill_observed_moments(n0: int, N: int, D: float, D0: float, t: float):
    """(mean, variance) of the recorded trapped fraction at observation time t of a
    pure filling process from n0 of N traps: (n0 + X + 1) / N with
    X ~ Binomial(N - n0, 1 - e^{-(D/D0) t}), from the binomial pmf."""
    from scipy.stats import binom
    m = N - n0
    q = -math.expm1(-(D / D0) * t)
    x = np.arange(m + 1)
    pmf = binom.pmf(x, m, q)
    rec = (n0 + x + 1) / N
    mean = float(np.sum(pmf * rec))
    return mean, float(np.sum(pmf * (rec - mean) ** 2))
