"""ctypes binding of the CPU oracle (oracle/spme_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke test of the package
entry point and by the CPU-baseline leg of bench.py, never by the product
(spme_amd/).  Each function restates the reference routine named in its
docstring (billy-inn/SPME, old-version/...).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE,
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        d, i = ctypes.c_double, ctypes.c_int
        L.orc_tab_splines.argtypes = [i, _dp]
        L.orc_tab_W.argtypes = [i, _dp]
        L.orc_scalar_recip.argtypes = [d, d, d, d]
        L.orc_scalar_recip.restype = d
        L.orc_compute_map.argtypes = [i, d, _dp]
        L.orc_compute_b.argtypes = [i, i, d, _dp, _dp]
        L.orc_compute_lm2.argtypes = [d, i, i, d, _dp]
        L.orc_compute_P.argtypes = [i, i, _dp, i, d, _dp, _ip, _ip, _dp]
        L.orc_spread.argtypes = [i, i, i, _dp, _ip, _dp, _dp]
        L.orc_interpolate.argtypes = [i, i, i, _dp, _ip, _dp, d, d, _dp]
        L.orc_fft1d.argtypes = [i, _dp, i]
        L.orc_fft3d_r2c.argtypes = [i, _dp, _dp]
        L.orc_fft3d_c2r.argtypes = [i, _dp, _dp]
        L.orc_apply_influence.argtypes = [i, d, _dp, _dp, _dp, _dp]
        L.orc_compute_spme.argtypes = [d, i, i, i, d, _dp, i, d, _dp, i, d,
                                       _dp, i]
        L.orc_real_coeffs.argtypes = [d, d, _dp, _dp]
        L.orc_overlap_coeffs.argtypes = [d, _dp, _dp]
        L.orc_self_coeff.argtypes = [d]
        L.orc_self_coeff.restype = d
        L.orc_real_space.argtypes = [i, _dp, d, d, d, _dp, i, _ip, _dp]
        L.orc_recip_exact.argtypes = [i, _dp, d, d, d, _dp, i, _ip, _dp]
        L.orc_ewald_direct.argtypes = [i, _dp, d, d, d, d, _dp, i, _ip, _dp]
        L.orc_spme_full.argtypes = [d, i, i, d, i, d, _dp, i, d, _dp, i, d,
                                    _dp, i]
        L.orc_real_coeffs_split.argtypes = [d, d, _dp, _dp, _dp, _dp]
        L.orc_overlap_coeffs_radii.argtypes = [d, d, d, _dp, _dp]
        L.orc_self_coeff_radius.argtypes = [d, d]
        L.orc_self_coeff_radius.restype = d
        L.orc_real_space_radii.argtypes = [i, _dp, _dp, d, d, d, _dp, i, _ip, _dp]
        L.orc_recip_exact_radii.argtypes = [i, _dp, _dp, d, d, d, _dp, i, _ip, _dp]
        L.orc_ewald_direct_radii.argtypes = [i, _dp, _dp, d, d, d, d, _dp, i, _ip, _dp]
        L.orc_spme_full_radii.argtypes = [d, i, i, d, i, d, _dp, _dp, i, d, _dp, i, d,
                                          _dp, i]
        L.orc_num_threads.restype = i
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def tab_splines(p: int) -> np.ndarray:
    """spme.c:14-22 (p = 8 derived): M_p(1..p-1)."""
    out = np.zeros(8)
    n = lib().orc_tab_splines(p, _d(out))
    if n == 0:
        raise ValueError(f"unsupported B-spline order {p}")
    return out[:n].copy()


def tab_W(p: int) -> np.ndarray:
    """spme.c:25-40 (p = 8 derived): W[x, y]."""
    out = np.zeros(64)
    n = lib().orc_tab_W(p, _d(out))
    if n == 0:
        raise ValueError(f"unsupported B-spline order {p}")
    return out[:n].reshape(p, p).copy()


def scalar_recip(k, xi, aa=1.0, ab=1.0):
    """spme.c:43-60."""
    return lib().orc_scalar_recip(float(k), float(xi), float(aa), float(ab))


def compute_map(K, L):
    """spme.c:95-107."""
    out = np.zeros(K)
    lib().orc_compute_map(K, float(L), _d(out))
    return out


def compute_b(K, p, L):
    """spme.c:109-126 -> complex b[K]."""
    m = compute_map(K, L)
    out = np.zeros(2 * K)
    lib().orc_compute_b(K, p, float(L), _d(m), _d(out))
    return out[0::2] + 1j * out[1::2]


def compute_lm2(xi, K, p, L):
    """spme.c:128-150 -> lm2[z, y, x], x in [0, K/2]."""
    out = np.zeros(K * K * (K // 2 + 1))
    lib().orc_compute_lm2(float(xi), K, p, float(L), _d(out))
    return out.reshape(K, K, K // 2 + 1)


def compute_P(K, p, pos, L):
    """spme.c:158-297 -> (P[N, p^3], ind[N, p^3], base[N, 3], q[N, 3, p])."""
    pos = _c64(pos).reshape(-1)
    N = pos.size // 3
    P = np.zeros(N * p ** 3)
    ind = np.zeros(N * p ** 3, dtype=np.int32)
    base = np.zeros(N * 3, dtype=np.int32)
    q = np.zeros(N * 3 * p)
    lib().orc_compute_P(K, p, _d(pos), N, float(L), _d(P), _i(ind), _i(base),
                        _d(q))
    return (P.reshape(N, p ** 3), ind.reshape(N, p ** 3), base.reshape(N, 3),
            q.reshape(N, 3, p))


def spread(K, p, P, ind, f):
    """spme_kernel.c:231-312 (serial) -> grid[3, K, K, K]."""
    P = _c64(P)
    ind = np.ascontiguousarray(ind, dtype=np.int32)
    f = _c64(f).reshape(-1)
    N = f.size // 3
    grid = np.zeros(3 * K ** 3)
    lib().orc_spread(K, p, N, _d(P), _i(ind), _d(f), _d(grid))
    return grid.reshape(3, K, K, K)


def interpolate(K, p, P, ind, grid, alpha, beta, u):
    """spme_kernel.c:8-70: returns beta*u + alpha*P^T grid."""
    P = _c64(P)
    ind = np.ascontiguousarray(ind, dtype=np.int32)
    grid = _c64(grid).reshape(-1)
    u = _c64(u).reshape(-1).copy()
    N = u.size // 3
    lib().orc_interpolate(K, p, N, _d(P), _i(ind), _d(grid), float(alpha),
                          float(beta), _d(u))
    return u


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    buf = np.empty(2 * x.size)
    buf[0::2] = x.real
    buf[1::2] = x.imag
    lib().orc_fft1d(x.size, _d(buf), int(sign))
    return buf[0::2] + 1j * buf[1::2]


def fft3d_r2c(r):
    """Same transform as spme.c:654-656 (MKL r2c), unnormalised."""
    r = _c64(r)
    K = r.shape[0]
    out = np.zeros(2 * K * K * (K // 2 + 1))
    lib().orc_fft3d_r2c(K, _d(r.reshape(-1)), _d(out))
    return (out[0::2] + 1j * out[1::2]).reshape(K, K, K // 2 + 1)


def fft3d_c2r(c):
    """Same transform as spme.c:680-682 without the 1/K^3 scale."""
    c = np.ascontiguousarray(c, dtype=np.complex128)
    K = c.shape[0]
    buf = np.empty(2 * c.size)
    buf[0::2] = c.real.reshape(-1)
    buf[1::2] = c.imag.reshape(-1)
    out = np.zeros(K ** 3)
    lib().orc_fft3d_c2r(K, _d(buf), _d(out))
    return out.reshape(K, K, K)


def apply_influence(K, L, lm2, spec3):
    """spme_kernel.c:114-228 on spec3[3, K, K, K/2+1] (complex)."""
    bufs = []
    for c in range(3):
        b = np.empty(2 * spec3[c].size)
        b[0::2] = spec3[c].real.reshape(-1)
        b[1::2] = spec3[c].imag.reshape(-1)
        bufs.append(b)
    lm2 = _c64(lm2).reshape(-1)
    lib().orc_apply_influence(K, float(L), _d(lm2), _d(bufs[0]), _d(bufs[1]),
                              _d(bufs[2]))
    return np.stack([(b[0::2] + 1j * b[1::2]).reshape(spec3[0].shape)
                     for b in bufs])


def compute_spme(xi, K, p, L, pos, f, alpha=1.0, beta=0.0, u=None):
    """spme.c:560-726 for f of shape [nrhs, 3N] (or [3N])."""
    pos = _c64(pos).reshape(-1)
    N = pos.size // 3
    f = _c64(f).reshape(-1, 3 * N)
    nrhs = f.shape[0]
    out = np.zeros_like(f) if u is None else _c64(u).reshape(nrhs, 3 * N).copy()
    lib().orc_compute_spme(float(xi), K, p, N, float(L), _d(pos), nrhs,
                           float(alpha), _d(f), 3 * N, float(beta), _d(out),
                           3 * N)
    return out


def real_coeffs(r, xi):
    F, G = ctypes.c_double(), ctypes.c_double()
    lib().orc_real_coeffs(float(r), float(xi), ctypes.byref(F),
                          ctypes.byref(G))
    return F.value, G.value


def overlap_coeffs(r):
    F, G = ctypes.c_double(), ctypes.c_double()
    lib().orc_overlap_coeffs(float(r), ctypes.byref(F), ctypes.byref(G))
    return F.value, G.value


def self_coeff(xi):
    return lib().orc_self_coeff(float(xi))


def _rows(rows, N):
    if rows is None:
        return None, N
    r = np.ascontiguousarray(rows, dtype=np.int32)
    return r, r.size


def real_space(pos, L, xi, rc, f, rows=None):
    pos = _c64(pos).reshape(-1)
    N = pos.size // 3
    f = _c64(f).reshape(-1)
    r, n = _rows(rows, N)
    u = np.zeros(3 * n)
    lib().orc_real_space(N, _d(pos), float(L), float(xi), float(rc), _d(f), n,
                         None if r is None else _i(r), _d(u))
    return u


def recip_exact(pos, L, xi, kmax, f, rows=None):
    pos = _c64(pos).reshape(-1)
    N = pos.size // 3
    f = _c64(f).reshape(-1)
    r, n = _rows(rows, N)
    u = np.zeros(3 * n)
    lib().orc_recip_exact(N, _d(pos), float(L), float(xi), float(kmax), _d(f),
                          n, None if r is None else _i(r), _d(u))
    return u


def ewald_direct(pos, L, f, rows=None, xi=None, tol_exp=6.6):
    """FP64 direct Beenakker-Ewald M.f for the given rows.

    xi defaults to a value that balances the real and wave-space work;
    rc = tol_exp/xi and kmax = 2*xi*tol_exp put both truncation errors near
    exp(-tol_exp^2) ~ 1e-19 relative to the leading term.
    """
    pos = _c64(pos).reshape(-1)
    N = pos.size // 3
    f = _c64(f).reshape(-1)
    r, n = _rows(rows, N)
    if xi is None:
        xi = default_direct_xi(N, L, n)
    rc = tol_exp / xi
    kmax = 2.0 * xi * tol_exp
    u = np.zeros(3 * n)
    lib().orc_ewald_direct(N, _d(pos), float(L), float(xi), float(rc),
                           float(kmax), _d(f), n, None if r is None else _i(r),
                           _d(u))
    return u


def default_direct_xi(N, L, nrows):
    """Splitting parameter that balances real-space pair work (nrows * rho *
    4/3 pi rc^3) against wave-space work (N * number of k vectors)."""
    rho = N / L ** 3
    c_real = nrows * rho * (4.0 / 3.0) * np.pi * 6.6 ** 3 * 10.0
    c_rec = N * (2.0 / 3.0) * np.pi * (2 * 6.6 * L / (2 * np.pi)) ** 3
    xi = (c_real / c_rec) ** (1.0 / 6.0)
    return float(max(xi, 4.0 / L * 6.6 / 6.0))


def spme_full(xi, K, p, rc, L, pos, f, alpha=1.0, beta=0.0, u=None):
    """Full SPME M.f on the CPU: reference reciprocal part + real + self."""
    pos = _c64(pos).reshape(-1)
    N = pos.size // 3
    f = _c64(f).reshape(-1, 3 * N)
    nrhs = f.shape[0]
    out = np.zeros_like(f) if u is None else _c64(u).reshape(nrhs, 3 * N).copy()
    lib().orc_spme_full(float(xi), K, p, float(rc), N, float(L), _d(pos), nrhs,
                        float(alpha), _d(f), 3 * N, float(beta), _d(out),
                        3 * N)
    return out


# --- unequal radii (spme_oracle.c: orc_*_radii) ------------------------------

def real_coeffs_split(r, xi):
    """(F_J, G_J, F_s, G_s): the Oseen and s2 = (a_i^2 + a_j^2)/2 parts of the
    real-space tensor, F = F_J + s2 F_s, G = G_J + s2 G_s."""
    v = [ctypes.c_double() for _ in range(4)]
    lib().orc_real_coeffs_split(float(r), float(xi), *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def overlap_coeffs_radii(r, ai, aj):
    F, G = ctypes.c_double(), ctypes.c_double()
    lib().orc_overlap_coeffs_radii(float(r), float(ai), float(aj), ctypes.byref(F),
                                   ctypes.byref(G))
    return F.value, G.value


def self_coeff_radius(xi, a):
    return lib().orc_self_coeff_radius(float(xi), float(a))


def real_space_radii(pos, rad, L, xi, rc, f, rows=None):
    pos = _c64(pos).reshape(-1)
    N = pos.size // 3
    rad, f = _c64(rad).reshape(-1), _c64(f).reshape(-1)
    r, n = _rows(rows, N)
    u = np.zeros(3 * n)
    lib().orc_real_space_radii(N, _d(pos), _d(rad), float(L), float(xi), float(rc), _d(f), n,
                               None if r is None else _i(r), _d(u))
    return u


def recip_exact_radii(pos, rad, L, xi, kmax, f, rows=None):
    pos = _c64(pos).reshape(-1)
    N = pos.size // 3
    rad, f = _c64(rad).reshape(-1), _c64(f).reshape(-1)
    r, n = _rows(rows, N)
    u = np.zeros(3 * n)
    lib().orc_recip_exact_radii(N, _d(pos), _d(rad), float(L), float(xi), float(kmax), _d(f),
                                n, None if r is None else _i(r), _d(u))
    return u


def ewald_direct_radii(pos, rad, L, f, rows=None, xi=None, tol_exp=6.6):
    """FP64 direct Ewald M.f with per-particle radii (see ewald_direct)."""
    pos = _c64(pos).reshape(-1)
    N = pos.size // 3
    rad, f = _c64(rad).reshape(-1), _c64(f).reshape(-1)
    r, n = _rows(rows, N)
    if xi is None:
        xi = default_direct_xi(N, L, n)
    u = np.zeros(3 * n)
    lib().orc_ewald_direct_radii(N, _d(pos), _d(rad), float(L), float(xi), tol_exp / xi,
                                 2.0 * xi * tol_exp, _d(f), n, None if r is None else _i(r),
                                 _d(u))
    return u


def spme_full_radii(xi, K, p, rc, L, pos, rad, f, alpha=1.0, beta=0.0, u=None):
    """Full SPME M.f with per-particle radii on the CPU."""
    pos = _c64(pos).reshape(-1)
    N = pos.size // 3
    rad = _c64(rad).reshape(-1)
    f = _c64(f).reshape(-1, 3 * N)
    nrhs = f.shape[0]
    out = np.zeros_like(f) if u is None else _c64(u).reshape(nrhs, 3 * N).copy()
    lib().orc_spme_full_radii(float(xi), K, p, float(rc), N, float(L), _d(pos), _d(rad), nrhs,
                              float(alpha), _d(f), 3 * N, float(beta), _d(out), 3 * N)
    return out


def num_threads():
    return lib().orc_num_threads()
