"""Edge cases of the engine against the CPU oracle and closed forms.

* One sphere: the periodic self-mobility 1 - 2.8372974794 a/L + (4 pi/3)
  (a/L)^3 (Hasimoto 1959), with no real-space neighbours (empty rows);
* two and five particles (mostly empty mesh tiles and sort buckets);
* clustered particles: 2,000 spheres on a jittered lattice in a cube of
  side 2.5 inside a box of side 55, so that a few mesh cells and one sort
  bucket per plane hold hundreds of particles (every staging path of
  spreading and interpolation runs in many batches) and each real-space row
  holds all 1,999 others, in the middle of the box and around the corner
  (coordinates below 0 and wrapped).

* extreme clustering: 50,000 particles in a unit cube, so that single sort
  buckets hold tens of thousands (sorted by a stable pass over the keys,
  not per thread), and a real-space block list that would need more than
  2^31 entries is refused with an error.

FP64 bar: relative max-norm difference <= 1e-12; FP32 <= 2e-5 (as the
other parity tests).
"""
import numpy as np
import pytest

from spme_amd import workload as W
from spme_amd.lib import SpmeError

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-5


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


@pytest.fixture(scope="module")
def E():
    from spme_amd import engine
    return engine


def _apply(E, torch, xi, K, p, L, rc, pos, f, prec, parts):
    dt = torch.float64 if prec == E.SPME_FP64 else torch.float32
    eng = E.Engine(xi, K, p, pos.shape[0], L, rc=rc, precision=prec, max_rhs=f.shape[0])
    eng.set_positions(torch.tensor(pos, device="cuda"))
    u = eng.apply(torch.tensor(f, device="cuda", dtype=dt), parts=parts)
    out = u.double().cpu().numpy()
    eng.close()
    return out


def _check(oracle, E, torch, xi, K, p, L, rc, pos, f):
    f = f.astype(np.float32).astype(np.float64)   # FP32 and FP64 runs share inputs
    ref = {E.SPME_RECIP: oracle.compute_spme(xi, K, p, L, pos, f),
           E.SPME_FULL: oracle.spme_full(xi, K, p, rc, L, pos, f)}
    for prec, tol in [(E.SPME_FP64, FP64_TOL), (E.SPME_FP32, FP32_TOL)]:
        for parts in (E.SPME_RECIP, E.SPME_FULL):
            got = _apply(E, torch, xi, K, p, L, rc, pos, f, prec, parts)
            d = _rel(got, ref[parts])
            _say(f"  precision {prec} parts {parts}: {d:.2e}")
            assert d < tol, (prec, parts)


def test_single_sphere_hasimoto(oracle, E, torch):
    L, xi, K, p, rc = 20.0, 0.5, 48, 8, 6.0
    _say(f"case: one sphere L={L} K={K} p={p}")
    pos = np.array([[3.3, 7.1, 12.9]])
    f = np.eye(3)                                  # three right-hand sides: e_x, e_y, e_z
    _check(oracle, E, torch, xi, K, p, L, rc, pos, f)
    u = _apply(E, torch, xi, K, p, L, rc, pos, f, E.SPME_FP64, E.SPME_FULL)
    hasimoto = 1 - 2.8372974794 / L + 4 * np.pi / 3 / L ** 3
    np.testing.assert_allclose(u, hasimoto * np.eye(3), rtol=0, atol=1e-7)


@pytest.mark.parametrize("N", [2, 5])
def test_few_particles(oracle, E, torch, N):
    L, xi, K, p, rc = 16.0, 0.6, 32, 6, 5.0
    _say(f"case: N={N} L={L} K={K} p={p}")
    pos = W.positions(50 + N, N, L)
    pos[0] = pos[1] + np.array([1.2, 0.3, -0.4])   # one overlapping pair (r < 2)
    _check(oracle, E, torch, xi, K, p, L, rc, pos, W.forces(60 + N, N, 2))


def _cluster(n, side, seed):
    """n points in a cube of the given side centred on 0: a jittered cubic
    lattice (spacing side/m, jitter below a fifth of it), so that no pair is
    closer than 0.6 spacings.  Nearer pairs are valid input, but the pair
    coefficients then cancel terms of size 1/r^3 (erfc(xi r)/(2r^3) against
    the overlap correction's 1/(2r^3)), and one ulp of erfc moves the result
    by 1e-16/r^3: the FP64 bar of 1e-12 needs r >~ 0.05."""
    m = int(np.ceil(n ** (1 / 3)))
    h = side / m
    g = (np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
         + 0.5) * h - side / 2
    rng = np.random.default_rng(seed)
    return g + rng.uniform(-0.2 * h, 0.2 * h, g.shape)


@pytest.mark.parametrize("where,p,K", [("centre", 6, 32), ("corner", 8, 48)])
def test_clustered_particles(oracle, E, torch, where, p, K):
    N = 2000
    L = W.box_length(N, 0.05)
    c = L / 2 if where == "centre" else 0.0
    pos = c + _cluster(N, 2.5, seed=7)
    _say(f"case: {N} particles clustered at the {where}, L={L:.1f} K={K} p={p}")
    _check(oracle, E, torch, 0.5, K, p, L, 5.0, pos, W.forces(70, N, 3))


def test_extreme_clustering(oracle, E, torch):
    N, L, K, p, xi = 50_000, 40.0, 32, 6, 0.5
    pos = 20.0 + np.random.default_rng(11).uniform(-0.5, 0.5, (N, 3))
    f = W.forces(80, N, 1)
    _say(f"case: {N} particles in a unit cube, L={L} K={K} p={p}, reciprocal part")
    ref = oracle.compute_spme(xi, K, p, L, pos, f)
    got = _apply(E, torch, xi, K, p, L, 0.0, pos, f, E.SPME_FP64, E.SPME_RECIP)
    d = _rel(got, ref)
    _say(f"  fp64 vs oracle: {d:.2e}")
    assert d < FP64_TOL
    _say("case: the same with rc = 5: N(N-1) block entries, refused")
    eng = E.Engine(xi, K, p, N, L, rc=5.0, precision=E.SPME_FP64)
    with pytest.raises(SpmeError, match="too clustered"):
        eng.set_positions(torch.tensor(pos, device="cuda"))
    eng.close()


def test_failed_set_positions_leaves_no_positions(E, torch):
    """A set_positions that fails after an earlier success leaves the engine
    without positions (not with half-rebuilt P data and block list): good
    positions, then a set refused as too clustered, then one with a NaN;
    each refusal is followed by an apply, which must report that positions
    have not been set, and good positions make it usable again."""
    N, L, K, p, xi = 50_000, 40.0, 32, 6, 0.5
    good = W.positions(81, N, L)
    clustered = 20.0 + np.random.default_rng(12).uniform(-0.5, 0.5, (N, 3))
    f = torch.tensor(W.forces(82, N, 1), device="cuda")
    _say("case: good positions, then refused ones, then apply")
    eng = E.Engine(xi, K, p, N, L, rc=5.0, precision=E.SPME_FP64)
    eng.set_positions(torch.tensor(good, device="cuda"))
    u_good = eng.apply(f, parts=E.SPME_FULL).cpu().numpy()
    with pytest.raises(SpmeError, match="too clustered"):
        eng.set_positions(torch.tensor(clustered, device="cuda"))
    for parts in (E.SPME_FULL, E.SPME_RECIP):
        with pytest.raises(SpmeError, match="positions have not been set"):
            eng.apply(f, parts=parts)
    bad = good.copy()
    bad[17, 2] = np.nan
    eng.set_positions(torch.tensor(good, device="cuda"))
    with pytest.raises(SpmeError, match="finite"):
        eng.set_positions(torch.tensor(bad, device="cuda"))
    with pytest.raises(SpmeError, match="positions have not been set"):
        eng.apply(f, parts=E.SPME_FULL)
    eng.set_positions(torch.tensor(good, device="cuda"))
    np.testing.assert_array_equal(eng.apply(f, parts=E.SPME_FULL).cpu().numpy(), u_good)
    eng.close()


def test_close_overlapping_pairs(oracle, E, torch):
    """Overlapping pairs at r = 0.01 .. 0.05.  There the real-space and
    overlap coefficients each hold terms of size 1/(2r^3) (erfc(xi r)/(2r^3)
    and the far-field subtraction's 1/(2r^3); G three times that) that cancel
    to O(1): one rounding of erfc or exp, in the engine or in the oracle,
    moves F and G by about eps/(2r^3).  The FP64 bar of a row is therefore
    1e-12 of the output scale plus 64 eps |f|_max / (2 r^3) for its closest
    overlapping pair (2e-9 at r = 0.01, 1e-11 at r = 0.05); a wrong or missing
    1/r^3 term would be off by about 1/(2 r^3) |f| (5e5 at r = 0.01).  FP32:
    the coefficients are formed in FP64 after the cancellation, so the usual
    2e-5 holds."""
    eps = np.finfo(np.float64).eps
    N0, L, K, p, xi, rc = 400, W.box_length(400, 0.1), 32, 6, 0.5, 5.0
    base = W.positions(41, N0, L)
    rng = np.random.default_rng(41)
    seps = np.array([0.01, 0.02, 0.03, 0.05])
    extra = []
    for r in seps:
        c = rng.uniform(0, L, 3)
        d = rng.normal(size=3)
        d *= r / np.linalg.norm(d)
        extra += [c, (c + d) % L]
    pos = np.concatenate([base, np.array(extra)])
    N = pos.shape[0]
    f = W.forces(42, N, 2).astype(np.float32).astype(np.float64)
    rmin = np.full(N, np.inf)
    for k, r in enumerate(seps):
        rmin[N0 + 2 * k] = rmin[N0 + 2 * k + 1] = r
    _say(f"case: {N} particles with overlapping pairs at r = {seps.tolist()}, full M.f")
    ref = oracle.spme_full(xi, K, p, rc, L, pos, f)
    got = _apply(E, torch, xi, K, p, L, rc, pos, f, E.SPME_FP64, E.SPME_FULL)
    scale = np.abs(ref).max()
    fmax = np.abs(f).max()
    bound = 1e-12 * scale + np.where(np.isfinite(rmin), 64 * eps * fmax / (2 * rmin ** 3), 0.0)
    diff = np.abs(got - ref).reshape(2, N, 3).max(axis=(0, 2))
    worst = float((diff / bound).max())
    _say(f"  fp64: worst row at {worst:.2e} of its bar; close rows differ by "
         f"{diff[N0:].max():.2e}, other rows by {diff[:N0].max() / scale:.2e} relative")
    assert worst <= 1.0
    got32 = _apply(E, torch, xi, K, p, L, rc, pos, f, E.SPME_FP32, E.SPME_FULL)
    assert _rel(got32, ref) < FP32_TOL
