"""Pins of the CPU oracle (oracle/spme_oracle.c) against closed forms,
brute force and invariants.  The reference has no tests or fixtures of its
own; every pin here is chosen so that a dropped term, a wrong sign or index,
or a transposed operand fails at least one check."""
import os
from fractions import Fraction
from math import comb, factorial

import numpy as np
import pytest

from spme_amd import workload as W

# ---------------------------------------------------------------- B-splines


def _M(p, t):
    """Cardinal B-spline M_p(t) on [0, p], exact (de Boor's truncated-power
    form)."""
    s = Fraction(0)
    for k in range(p + 1):
        if t - k > 0:
            s += (-1) ** k * comb(p, k) * (t - k) ** (p - 1)
    return s / factorial(p - 1)


def _W_exact(p):
    """Coefficient of g^y in M_p(g + 1/2 + p - 1 - x), solved exactly from p
    samples of the polynomial piece."""
    rows = []
    pts = [Fraction(-1, 2) + Fraction(i + 1, 2 * p + 2) for i in range(p)]
    for x in range(p):
        A = [[g ** j for j in range(p)] + [_M(p, g + Fraction(1, 2) + p - 1 - x)]
             for g in pts]
        for c in range(p):
            piv = next(r for r in range(c, p) if A[r][c] != 0)
            A[c], A[piv] = A[piv], A[c]
            A[c] = [a / A[c][c] for a in A[c]]
            for r in range(p):
                if r != c and A[r][c] != 0:
                    fac = A[r][c]
                    A[r] = [a - fac * b for a, b in zip(A[r], A[c])]
        rows.append([A[r][p] for r in range(p)])
    return rows


# Literal tables of the reference, old-version/spme.c:19,21 and :29-39.
REF_SPLINES = {4: [1 / 6, 4 / 6, 1 / 6],
               6: [1 / 120, 26 / 120, 66 / 120, 26 / 120, 1 / 120]}
REF_W4 = np.array([1, -6, 12, -8, 23, -30, -12, 24, 23, 30, -12, -24,
                   1, 6, 12, 8]) / 48.0
REF_W6 = np.array([1, -10, 40, -80, 80, -32, 237, -750, 840, -240, -240, 160,
                   1682, -1540, -880, 1120, 160, -320, 1682, 1540, -880, -1120,
                   160, 320, 237, 750, 840, 240, -240, -160,
                   1, 10, 40, 80, 80, 32]) / 3840.0


@pytest.mark.parametrize("p", [4, 6, 8])
def test_tables_match_closed_form(oracle, p):
    W_ = oracle.tab_W(p)
    exact = _W_exact(p)
    for x in range(p):
        for y in range(p):
            assert W_[x, y] == float(exact[x][y])
    spl = oracle.tab_splines(p)
    for i in range(p - 1):
        assert spl[i] == float(_M(p, Fraction(i + 1)))


def test_tables_match_reference_literals(oracle):
    np.testing.assert_array_equal(oracle.tab_W(4).ravel(), REF_W4)
    np.testing.assert_array_equal(oracle.tab_W(6).ravel(), REF_W6)
    for p, v in REF_SPLINES.items():
        np.testing.assert_array_equal(oracle.tab_splines(p), np.array(v))


@pytest.mark.parametrize("p,nyq", [(4, 9.0), (6, 56.25), (8, (5040 / 272) ** 2)])
def test_euler_factor_special_values(oracle, p, nyq):
    K, L = 32, 20.0
    b = oracle.compute_b(K, p, L)
    assert abs(abs(b[0]) ** 2 - 1.0) < 1e-14
    assert abs(abs(b[K // 2]) ** 2 - nyq) < 1e-10 * nyq
    # |b(m)| = |b(-m)|
    for m in range(1, K // 2):
        assert abs(abs(b[m]) - abs(b[K - m])) < 1e-12 * abs(b[m])


def test_map_nyquist_convention(oracle):
    K, L = 16, 10.0
    m = oracle.compute_map(K, L)
    assert m[K // 2] == pytest.approx(np.pi * K / L)   # +pi K / L
    assert m[K // 2 + 1] == pytest.approx(-(K // 2 - 1) * 2 * np.pi / L)


# ---------------------------------------------------------------- P matrix


@pytest.mark.parametrize("p", [4, 6, 8])
def test_P_partition_of_unity_and_first_moment(oracle, p):
    K, L, N = 32, 17.0, 200
    pos = W.positions(3, N, L)
    P, ind, base, q = oracle.compute_P(K, p, pos, L)
    np.testing.assert_allclose(q.sum(axis=2), 1.0, atol=1e-13)
    # first moment: sum_x q[x] * j_x = c with j_x the unwrapped grid index
    c = pos / L * K
    j0 = np.floor(c) - p // 2 + 1
    mom = (q * (j0[:, :, None] + np.arange(p))).sum(axis=2)
    np.testing.assert_allclose(mom, c, atol=1e-11)
    # P is the separable product, ind the wrapped flattened index
    i = 17
    xs, ys, zs = np.meshgrid(np.arange(p), np.arange(p), np.arange(p),
                             indexing="xy")
    Pd = np.einsum("z,y,x->zyx", q[i, 2], q[i, 1], q[i, 0]).ravel()
    np.testing.assert_allclose(P[i], Pd, rtol=1e-14)
    zz = (base[i, 2] + np.arange(p)) % K
    yy = (base[i, 1] + np.arange(p)) % K
    xx = (base[i, 0] + np.arange(p)) % K
    idx = (zz[:, None, None] * K * K + yy[None, :, None] * K
           + xx[None, None, :]).ravel()
    np.testing.assert_array_equal(ind[i], idx)


def _dense_P(K, p, pos, L):
    """Dense N x K^3 interpolation matrix from the B-spline itself (not from
    the W table): weight of grid j along one axis is M_p(c - j + p/2),
    periodic."""
    N = pos.shape[0]
    out = np.zeros((N, K, K, K))
    for i in range(N):
        w = []
        for d in range(3):
            c = pos[i, d] / L * K
            wd = np.zeros(K)
            for j in range(-K, 2 * K):
                t = c - j + p / 2
                if 0 < t < p:
                    wd[j % K] += float(_M(p, Fraction(t)))
            w.append(wd)
        out[i] = np.einsum("z,y,x->zyx", w[2], w[1], w[0])
    return out.reshape(N, -1)


@pytest.mark.parametrize("p", [4, 6])
def test_spread_interpolate_brute_force(oracle, p):
    K, L, N = 12, 7.5, 6
    pos = W.positions(9, N, L)
    f = W.forces(9, N)[0]
    P, ind, _, _ = oracle.compute_P(K, p, pos, L)
    D = _dense_P(K, p, pos, L)
    grid = oracle.spread(K, p, P, ind, f)
    for c in range(3):
        np.testing.assert_allclose(grid[c].ravel(), D.T @ f[c::3], atol=1e-12)
    g = np.random.default_rng(1).standard_normal((3, K, K, K))
    u0 = np.random.default_rng(2).standard_normal(3 * N)
    u = oracle.interpolate(K, p, P, ind, g, 0.7, 1.3, u0)
    for c in range(3):
        np.testing.assert_allclose(u[c::3], 1.3 * u0[c::3] + 0.7 * (D @ g[c].ravel()),
                                   atol=1e-12)


# ---------------------------------------------------------------- FFT


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 18, 32, 48, 64, 96, 192])
def test_fft1d_numpy(oracle, n):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(
        n + 1).standard_normal(n)
    np.testing.assert_allclose(oracle.fft1d(x, -1), np.fft.fft(x), atol=1e-12 * n)
    np.testing.assert_allclose(oracle.fft1d(x, +1), np.fft.ifft(x) * n,
                               atol=1e-12 * n)


def test_fft1d_brute_force(oracle):
    n = 12
    x = np.random.default_rng(5).standard_normal(n) + 0j
    jk = np.outer(np.arange(n), np.arange(n))
    dft = np.exp(-2j * np.pi * jk / n) @ x
    np.testing.assert_allclose(oracle.fft1d(x, -1), dft, atol=1e-12)


@pytest.mark.parametrize("K", [8, 12, 32, 48])
def test_fft3d_roundtrip_numpy(oracle, K):
    r = np.random.default_rng(K).standard_normal((K, K, K))
    c = oracle.fft3d_r2c(r)
    np.testing.assert_allclose(c, np.fft.rfftn(r), atol=1e-10)
    np.testing.assert_allclose(oracle.fft3d_c2r(c) / K ** 3, r, atol=1e-12)


def test_c2r_non_hermitian_convention(oracle):
    """c2r of a half spectrum whose x = 0 and x = K/2 planes are not
    Hermitian returns the real part of the inverse DFT of the stored planes
    completed by conjugate symmetry in x."""
    K = 8
    rng = np.random.default_rng(3)
    c = rng.standard_normal((K, K, K // 2 + 1)) + 1j * rng.standard_normal(
        (K, K, K // 2 + 1))
    full = np.zeros((K, K, K), complex)
    full[:, :, :K // 2 + 1] = c
    for x in range(K // 2 + 1, K):
        full[:, :, x] = np.conj(c[(-np.arange(K)) % K][:, (-np.arange(K)) % K,
                                                       K - x])
    ref = np.real(np.fft.ifftn(full)) * K ** 3
    np.testing.assert_allclose(oracle.fft3d_c2r(c.copy()), ref, atol=1e-10)


# ---------------------------------------------------------------- influence


def test_influence_divergence_free(oracle):
    K, L, p, xi = 16, 11.0, 6, 0.6
    lm2 = oracle.compute_lm2(xi, K, p, L)
    rng = np.random.default_rng(4)
    spec = rng.standard_normal((3, K, K, K // 2 + 1)) + 1j * rng.standard_normal(
        (3, K, K, K // 2 + 1))
    out = oracle.apply_influence(K, L, lm2, spec)
    m = oracle.compute_map(K, L)
    kz, ky, kx = np.meshgrid(m, m, m[:K // 2 + 1], indexing="ij")
    div = kx * out[0] + ky * out[1] + kz * out[2]
    assert np.abs(div).max() < 1e-12 * np.abs(out).max()
    assert np.all(out[:, 0, 0, 0] == 0)
    # B is symmetric and its trace is 2*lm2 away from k = 0
    e = np.zeros_like(spec)
    e[0] = 1.0
    B0 = oracle.apply_influence(K, L, lm2, e)
    e = np.zeros_like(spec)
    e[1] = 1.0
    B1 = oracle.apply_influence(K, L, lm2, e)
    e = np.zeros_like(spec)
    e[2] = 1.0
    B2 = oracle.apply_influence(K, L, lm2, e)
    np.testing.assert_allclose(B0[1], B1[0], atol=1e-14)
    np.testing.assert_allclose(B0[2], B2[0], atol=1e-14)
    tr = (B0[0] + B1[1] + B2[2]).real
    tr[0, 0, 0] = 2 * lm2[0, 0, 0]
    np.testing.assert_allclose(tr, 2 * lm2, rtol=1e-12, atol=1e-300)


# ---------------------------------------------------------------- operators


def test_spme_recip_symmetric_and_converges(oracle):
    N, L, xi = 60, 12.0, 0.5
    pos = W.positions(21, N, L)
    f = W.forces(21, N, 2)
    u = oracle.compute_spme(xi, 32, 8, L, pos, f)
    # symmetry g.(M f) = f.(M g)
    assert abs(f[1] @ u[0] - f[0] @ u[1]) < 1e-12 * np.abs(u).max() * N
    ex = oracle.recip_exact(pos, L, xi, 2 * xi * 6.6, f[0])
    e16 = np.linalg.norm(oracle.compute_spme(xi, 16, 6, L, pos, f[0])[0] - ex)
    e32 = np.linalg.norm(u[0] - ex)
    assert e32 < 1e-6 * np.linalg.norm(ex)
    assert e32 < 1e-2 * e16


def test_spme_alpha_beta_and_translation(oracle):
    N, L, xi, K, p = 40, 10.0, 0.6, 16, 4
    pos = W.positions(22, N, L)
    f = W.forces(22, N)[0]
    u0 = W.forces(23, N)[0]
    a = oracle.compute_spme(xi, K, p, L, pos, f)[0]
    b = oracle.compute_spme(xi, K, p, L, pos, f, alpha=2.1, beta=2.2, u=u0)[0]
    np.testing.assert_allclose(b, 2.2 * u0 + 2.1 * a, atol=1e-12)
    # translating every particle by whole mesh cells leaves u unchanged
    sh = (pos + np.array([3, 5, 7]) * L / K) % L
    c = oracle.compute_spme(xi, K, p, L, sh, f)[0]
    np.testing.assert_allclose(c, a, atol=1e-12)


def test_ewald_xi_independence(oracle):
    N = 80
    L = W.box_length(N, 0.2)
    pos = W.positions(31, N, L)
    f = W.forces(31, N)[0]
    us = [oracle.ewald_direct(pos, L, f, xi=x) for x in (0.3, 0.55, 0.9)]
    for u in us[1:]:
        assert np.abs(u - us[0]).max() < 1e-12 * np.abs(us[0]).max()
    # positive definite
    assert f @ us[0] > 0


def test_hasimoto_single_sphere(oracle):
    """Periodic self-mobility of one sphere: 1 - 2.8372974794 a/L
    + (4 pi/3)(a/L)^3 (Hasimoto 1959; Beenakker 1986)."""
    for L in (6.0, 10.0, 25.0):
        for xi in (0.4, 0.8):
            u = oracle.ewald_direct(np.array([[0.3, 1.7, 2.9]]), L,
                                    np.array([0.0, 1.0, 0.0]), xi=xi)
            exact = 1 - 2.8372974794 / L + 4 * np.pi / 3 / L ** 3
            assert abs(u[1] - exact) < 1e-10
            assert abs(u[0]) < 1e-13 and abs(u[2]) < 1e-13


def test_overlap_continuity(oracle):
    F, G = oracle.overlap_coeffs(2.0 - 1e-12)
    assert abs(F) < 1e-11 and abs(G) < 1e-11
    assert oracle.overlap_coeffs(2.5) == (0.0, 0.0)
    # two coincident-ish spheres: full RPY at r -> 0 is the identity
    r = 1e-4
    F, G = oracle.overlap_coeffs(r)
    far_F = 0.75 / r * (1 + 2 / (3 * r * r))
    far_G = 0.75 / r * (1 - 2 / (r * r))
    assert abs(F + far_F - 1.0) < 1e-3 and abs(G + far_G) < 1e-3


def test_real_space_cell_list_matches_brute_force(oracle):
    N = 300
    L = W.box_length(N, 0.2)
    pos = W.positions(41, N, L)
    f = W.forces(41, N)[0]
    # xi = 1.0 -> rc = 6.6 < L/3 uses cells; xi = 0.25 does not
    a = oracle.ewald_direct(pos, L, f, xi=1.0)
    b = oracle.ewald_direct(pos, L, f, xi=0.25)
    assert np.abs(a - b).max() < 1e-11 * np.abs(a).max()


def test_spme_full_converges_to_direct(oracle):
    N = 200
    L = W.box_length(N, 0.1)
    pos = W.positions(11, N, L)
    f = W.forces(11, N)[0]
    ref = oracle.ewald_direct(pos, L, f, xi=0.5)
    e = []
    for K, p in [(24, 6), (32, 8), (64, 8)]:
        u = oracle.spme_full(0.5, K, p, 13.2, L, pos, f)[0]
        e.append(np.linalg.norm(u - ref) / np.linalg.norm(ref))
    assert e[0] > e[1] > e[2]
    assert e[2] < 1e-7


def _asymmetry(f, u):
    """max |f_a . M f_b - f_b . M f_a| / (|f| |M f|) over pairs (a, b)."""
    G = f @ u.T
    scale = np.outer(np.linalg.norm(f, axis=1), np.linalg.norm(u, axis=1))
    return float((np.abs(G - G.T) / np.maximum(scale, scale.T)).max())


def test_symmetry_check_catches_planted_mistakes(oracle):
    """The symmetry bar the GPU applies at the full batched size (1e-6 of
    |f| |M f|) is met by the reciprocal operator composed from the oracle's
    stages, and missed by orders of magnitude when a mistake is planted in
    the composition: the interpolation stencil one mesh cell off, or two
    force components exchanged between spreading and interpolation."""
    N, K, p, xi = 300, 32, 6, 0.5
    L = W.box_length(N, 0.1)
    pos = W.positions(41, N, L)
    f = W.forces(41, N, 6)
    P, ind, _, _ = oracle.compute_P(K, p, pos, L)
    lm2 = oracle.compute_lm2(xi, K, p, L)

    def recip(fv, plant=None):
        grid = oracle.spread(K, p, P, ind, fv)
        spec = oracle.apply_influence(K, L, lm2,
                                      np.stack([oracle.fft3d_r2c(grid[c]) for c in range(3)]))
        g = np.stack([oracle.fft3d_c2r(spec[c]) for c in range(3)]) / K ** 3
        if plant == "stencil":
            g = np.roll(g, 1, axis=1)
        elif plant == "components":
            g = g[[1, 0, 2]]
        return oracle.interpolate(K, p, P, ind, g, K ** 3 / L ** 3, 0.0, np.zeros(3 * N))

    u = np.stack([recip(fv) for fv in f])
    np.testing.assert_allclose(u, oracle.compute_spme(xi, K, p, L, pos, f),
                               rtol=0, atol=1e-12 * np.abs(u).max())
    assert _asymmetry(f, u) < 1e-12
    for plant in ("stencil", "components"):
        bad = np.stack([recip(fv, plant) for fv in f])
        assert _asymmetry(f, bad) > 1e-4, plant


def test_c1_fixture(oracle):
    """The committed configuration-1 fixture (tests/golden/c1.npz): the
    oracle reproduces its recip and full rows, and a recomputed subset of
    its FP64 direct Ewald rows.  (Configuration 1's xi = 0.896, set by the
    real-space cutoff rc = 5, is beyond what its K = 32 mesh resolves,
    xi h = 0.97, so its SPME rows are far from the direct sum; they pin the
    restated algorithm, not its accuracy.)"""
    import json
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    g = np.load(os.path.join(gdir, "c1.npz"))
    meta = json.load(open(os.path.join(gdir, "c1.json")))
    s = W.make_system(meta["seed"], meta["n"], meta["phi"])
    f = W.forces(meta["seed"], meta["n"], 1).astype(np.float32).astype(np.float64)
    recip = oracle.compute_spme(meta["xi"], meta["K"], meta["p"], s.L, s.pos, f)
    full = oracle.spme_full(meta["xi"], meta["K"], meta["p"], meta["rc"], s.L, s.pos, f)
    assert np.abs(recip - g["recip"]).max() <= 1e-13 * np.abs(g["recip"]).max()
    assert np.abs(full - g["full"]).max() <= 1e-13 * np.abs(g["full"]).max()
    sub = np.arange(0, meta["n"], 97, dtype=np.int32)
    d = oracle.ewald_direct(s.pos, s.L, f[0], rows=sub)
    ref = g["direct"][0].reshape(-1, 3)[sub].reshape(-1)
    assert np.abs(d - ref).max() <= 1e-12 * np.abs(ref).max()


# --- unequal radii ------------------------------------------------------------

def test_radii_real_space_pieces(oracle):
    """The Oseen and s2 parts sum to the unit-radius coefficients; the
    unequal-sphere overlap correction reduces to Rotne-Prager's for equal
    radii, vanishes continuously at r = a_i + a_j and is continuous at the
    containment radius r = |a_i - a_j|."""
    for r in (0.3, 1.5, 3.0, 6.0):
        FJ, GJ, Fs, Gs = oracle.real_coeffs_split(r, 0.7)
        F, G = oracle.real_coeffs(r, 0.7)
        assert abs(FJ + Fs - F) < 1e-13 * max(1, abs(F)) and abs(GJ + Gs - G) < 1e-13 * max(1, abs(G))
    for r in (0.5, 1.2, 1.9):
        assert np.allclose(oracle.overlap_coeffs_radii(r, 1.0, 1.0), oracle.overlap_coeffs(r),
                           rtol=0, atol=1e-14)
    ai, aj = 1.7, 0.6
    assert np.abs(oracle.overlap_coeffs_radii(ai + aj - 1e-10, ai, aj)).max() < 1e-9
    lo = oracle.overlap_coeffs_radii(ai - aj - 1e-10, ai, aj)
    hi = oracle.overlap_coeffs_radii(ai - aj + 1e-10, ai, aj)
    assert np.abs(np.subtract(lo, hi)).max() < 1e-8
    # inside the larger sphere the pair moves as the larger sphere alone:
    # M_ov = I / a_max, so F + far field = 1 / a_max, G + far field = 0
    r = 0.5
    F, G = oracle.overlap_coeffs_radii(r, ai, aj)
    s2 = 0.5 * (ai * ai + aj * aj)
    assert abs(F + 0.75 / r + s2 / (2 * r ** 3) - 1 / ai) < 1e-13
    assert abs(G + 0.75 / r - 1.5 * s2 / r ** 3) < 1e-12


def test_radii_direct_sum(oracle):
    """Direct Ewald sum with radii: independent of the splitting parameter,
    equal to the unit-radius sum for unit radii, and Hasimoto's periodic
    self-mobility 1/a - 2.8372974794/L + 4 pi a^2/(3 L^3) for one sphere of
    radius a."""
    N, L = 60, 14.0
    pos = W.positions(5, N, L)
    f = W.forces(5, N, 1)[0]
    d1 = oracle.ewald_direct(pos, L, f, xi=0.5)
    d1r = oracle.ewald_direct_radii(pos, np.ones(N), L, f, xi=0.5)
    assert np.abs(d1 - d1r).max() <= 1e-13 * np.abs(d1).max()
    rad = 0.5 + 1.5 * W.uniform(6, 0, N)
    a = oracle.ewald_direct_radii(pos, rad, L, f, xi=0.45)
    b = oracle.ewald_direct_radii(pos, rad, L, f, xi=0.75)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    for aa in (0.5, 2.0):
        u = oracle.ewald_direct_radii(np.array([[3.0, 4.0, 5.0]]), np.array([aa]), 25.0,
                                      np.array([1.0, 0.0, 0.0]), xi=0.4)
        ref = 1 / aa - 2.8372974794 / 25.0 + 4 * np.pi * aa * aa / (3 * 25.0 ** 3)
        assert abs(u[0] - ref) < 1e-10 and abs(u[1]) < 1e-12 and abs(u[2]) < 1e-12


def test_radii_spme(oracle):
    """SPME with radii: unit radii give the unit-radius SPME; the operator is
    symmetric (the a_i^2 terms of interpolation and spreading are each other's
    transposes); it converges to the direct sum."""
    N, L = 60, 14.0
    pos = W.positions(5, N, L)
    F = W.forces(7, N, 2)
    s1 = oracle.spme_full(0.8, 32, 6, 5.0, L, pos, F)
    s1r = oracle.spme_full_radii(0.8, 32, 6, 5.0, L, pos, np.ones(N), F)
    assert np.abs(s1 - s1r).max() <= 1e-13 * np.abs(s1).max()
    rad = 0.5 + 1.5 * W.uniform(6, 0, N)
    u = oracle.spme_full_radii(0.7, 48, 8, 6.9, L, pos, rad, F)
    asym = abs(F[0] @ u[1] - F[1] @ u[0]) / (np.linalg.norm(F[0]) * np.linalg.norm(u[1]))
    assert asym < 1e-13, asym
    d = oracle.ewald_direct_radii(pos, rad, L, F[0], xi=0.5)
    err = np.linalg.norm(u[0] - d) / np.linalg.norm(d)
    assert err < 2e-5, err


@pytest.mark.parametrize("K,p", [(128, 6), (48, 8), (24, 6), (96, 8), (64, 6)])
def test_coloured_spread_equals_serial_sum(oracle, K, p):
    """The coloured parallel spread (blocks of one parity colour spread
    concurrently) equals the plain particle-by-particle sum, including
    meshes that p does not divide (K = 128, p = 6: a stencil that wraps from
    the last block into block 0) and densely packed particles."""
    N = 3000
    pos = W.positions(17, N, 7.0)   # 3000 particles in a box of side 7: many per block
    f = W.forces(17, N, 1)[0]
    P, ind = oracle.compute_P(K, p, pos, 7.0)[:2]
    g = oracle.spread(K, p, P, ind, f)
    ref = np.zeros((3, K ** 3))
    for c in range(3):
        np.add.at(ref[c], ind.reshape(-1), (P * f[c::3][:, None]).reshape(-1))
    assert np.abs(g.reshape(3, -1) - ref).max() <= 1e-12 * np.abs(ref).max()
