"""Ewald and mesh parameter selection for the RPY SPME product.

The reference declares `cutoff2xi(rmax, ewald_eps, box_size, *xi)`
(old-version/spme.h:85-86) but never defines it and hard-codes xi = 0.5
(old-version/main.c:210).  This module chooses the whole set
(xi, K, p, rc) for a target relative error of u = M f:

  error model   eps^2 = eps_real^2 + eps_recip^2, relative to rms(u)
    eps_real    truncation of the real-space sum at rc, from the tail
                integral rho * int_rc^inf 4 pi r^2 (3F^2 + 2FG + G^2)/3 dr
                of the Beenakker real-space tensor F I + G rr^T/r^2 (exact
                for uncorrelated unit forces);
    eps_recip   mesh (B-spline aliasing) error, a per-order power law
                exp(a_p) (xi h)^b_p xi^c_p phi^d_p fitted to the restated
                reference SPME against the exact Ewald k-sum;
    rms(u)      sqrt(1 + 0.66 phi L) rms(f), from the FP64 direct sum on
                random suspensions (the long-range 1/r hydrodynamic field
                grows with the box);
  cost model    seconds per apply on one MI355X, linear in the mesh points,
                the stencil work and the real-space pairs, with constants
                from the measured kernel times.

`analytic_select` is the fallback of the learned selector (spme_amd/mlp.py).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SMOOTH_K = (32, 48, 64, 96, 128, 192, 256)
ORDERS = (4, 6, 8)

# log eps_recip = a + b log(xi h) + c log(xi) + d log(phi), absolute error per
# unit rms force.  Fitted on 108 runs (phi in {0.05, 0.2, 0.35}, L in {30, 45},
# xi in {0.3, 0.5, 0.7}, K in {24..64}); rms residual 0.39 / 0.26 / 0.27 in
# log for p = 4 / 6 / 8.
RECIP_FIT = {4: (0.8134, 4.251, 1.8641, 0.1683),
             6: (2.3767, 7.1157, 3.0202, 0.1786),
             8: (4.1749, 10.1193, 3.7535, 0.2002)}

# Cost model (seconds): c_mesh K^3 C + c_sten N p^3 C + c_pair N npair C
# + c_vec N C + c_fix, for C = 3 nrhs components, FP32; FP64 doubles the
# bandwidth-bound terms.  From the kernel times at N = 2e5, K = 192, p = 8,
# 16 right-hand sides.
COST = dict(mesh=1.24e-11, sten=5.1e-13, pair=8.3e-13, vec=1.35e-11, fix=1.5e-4)


def box_length(n: int, phi: float) -> float:
    return (n * (4.0 / 3.0) * math.pi / phi) ** (1.0 / 3.0)


def real_coeffs(r, xi):
    """Beenakker real-space tensor coefficients F, G (unit radius), vectorised."""
    from scipy.special import erfc
    r = np.asarray(r, dtype=np.float64)
    r2 = r * r
    e = np.exp(-(xi * r) ** 2) / math.sqrt(math.pi)
    ec = erfc(xi * r)
    x3, x5, x7 = xi ** 3, xi ** 5, xi ** 7
    F = ec * (0.75 / r + 0.5 / (r2 * r)) + e * (
        4 * x7 * r2 * r2 + 3 * x3 * r2 - 20 * x5 * r2 - 4.5 * xi + 14 * x3 + xi / r2)
    G = ec * (0.75 / r - 1.5 / (r2 * r)) + e * (
        -4 * x7 * r2 * r2 - 3 * x3 * r2 + 16 * x5 * r2 + 1.5 * xi - 2 * x3 - 3 * xi / r2)
    return F, G


_TAIL = {}


def _tail_table(xi: float):
    """Cumulative tail T(r) = int_r^inf 4 pi s^2 (3F^2 + 2FG + G^2)/3 ds on a
    grid (cached per xi)."""
    key = round(xi, 12)
    tab = _TAIL.get(key)
    if tab is None:
        r = np.linspace(0.25, 0.25 + 40.0 / xi, 20001)
        F, G = real_coeffs(r, xi)
        g = 4 * math.pi * r * r * (3 * F * F + 2 * F * G + G * G) / 3.0
        seg = 0.5 * (g[1:] + g[:-1]) * np.diff(r)
        tail = np.concatenate([np.cumsum(seg[::-1])[::-1], [0.0]])
        tab = (r, tail)
        if len(_TAIL) > 4096:
            _TAIL.clear()
        _TAIL[key] = tab
    return tab


def real_error_abs(xi: float, rc: float, phi: float) -> float:
    """sqrt(rho * int_rc^inf 4 pi r^2 (3F^2 + 2FG + G^2)/3 dr)."""
    rho = phi / (4.0 / 3.0 * math.pi)
    r, tail = _tail_table(xi)
    return math.sqrt(max(rho * float(np.interp(rc, r, tail)), 0.0))


def recip_error_abs(xi: float, h: float, p: int, phi: float) -> float:
    a, b, c, d = RECIP_FIT[p]
    return math.exp(a + b * math.log(xi * h) + c * math.log(xi) + d * math.log(phi))


def rms_u(phi: float, L: float) -> float:
    return math.sqrt(1.0 + 0.66 * phi * L)


def estimate_error(xi, K, p, rc, phi, L) -> float:
    """Estimated relative error of the full M.f."""
    er = real_error_abs(xi, rc, phi)
    ek = recip_error_abs(xi, L / K, p, phi)
    return math.sqrt(er * er + ek * ek) / rms_u(phi, L)


def pairs_per_particle(rc: float, phi: float) -> float:
    rho = phi / (4.0 / 3.0 * math.pi)
    return rho * 4.0 / 3.0 * math.pi * max(rc, 2.0) ** 3


def estimate_cost(n, K, p, rc, phi, nrhs=1, fp64=False) -> float:
    C = 3 * nrhs
    t = (COST["mesh"] * K ** 3 * C + COST["sten"] * n * p ** 3 * C
         + COST["pair"] * n * 1.25 * pairs_per_particle(rc, phi) * C
         + COST["vec"] * n * C)
    return COST["fix"] + (2.0 * t if fp64 else t)


@dataclass
class Params:
    xi: float
    K: int
    p: int
    rc: float
    est_error: float
    est_cost: float

    def as_tuple(self):
        return (self.xi, self.K, self.p, self.rc)


def rc_for(xi: float, target_abs: float, phi: float, rc_max: float):
    """Smallest rc with real_error_abs <= target, or None past rc_max."""
    rho = phi / (4.0 / 3.0 * math.pi)
    r, tail = _tail_table(xi)
    need = target_abs * target_abs / rho
    # tail is non-increasing: first grid point where it drops below need
    i = int(np.searchsorted(-tail, -need, side="left"))
    if i >= r.size:
        return None
    rc = float(r[i])
    if rc > rc_max:
        return None
    return max(rc, 0.5)


def analytic_select(n: int, phi: float, tol: float, nrhs: int = 1,
                    k_max: int = 256, orders=ORDERS, fixed_K=None, fixed_p=None,
                    fp64=False) -> Params:
    """Cheapest (xi, K, p, rc) whose estimated relative error is <= tol.

    For each mesh and order the error budget is split between the mesh and
    real-space parts (a few splits are tried); xi follows from the mesh
    budget through the fitted power law, rc from the real-space budget.
    """
    L = box_length(n, phi)
    target = tol * rms_u(phi, L)
    rc_max = L / 3.0
    best = None
    Ks = [fixed_K] if fixed_K else [k for k in SMOOTH_K if k <= k_max]
    ps = [fixed_p] if fixed_p else list(orders)
    for K in Ks:
        h = L / K
        for p in ps:
            if 4 * p >= K:
                continue
            a, b, c, d = RECIP_FIT[p]
            for s in (0.3, 0.5, 0.7, 0.85, 0.95):
                ek = s * target
                er = math.sqrt(1 - s * s) * target
                lxi = (math.log(ek) - a - b * math.log(h) - d * math.log(phi)) / (b + c)
                xi = math.exp(lxi)
                if not (0.05 < xi < 3.0):
                    continue
                rc = rc_for(xi, er, phi, rc_max)
                if rc is None:
                    continue
                cost = estimate_cost(n, K, p, rc, phi, nrhs, fp64)
                err = estimate_error(xi, K, p, rc, phi, L)
                if best is None or cost < best.est_cost:
                    best = Params(xi, K, p, rc, err, cost)
    if best is None:
        raise ValueError(f"no parameters reach tol={tol} with K <= {k_max}")
    return best
