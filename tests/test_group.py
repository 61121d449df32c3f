"""Grouped multi-system applies, tile shapes and the autotuner.

A group runs the same device bodies as the separate engines (for meshes up
to K = 96; larger single applies fuse the x and y FFT passes), and no tile
shape changes the arithmetic, so the comparisons are bitwise:
  - grouped apply (plain launches and one captured hipGraph, replayed with
    eager work and synchronisation between replays) == the systems'
    separate spme_engine_apply results, for the reciprocal part and the
    full M.f, FP32 and FP64;
  - every tile shape, persistent or not, and the autotuned shapes == the
    default shapes;
  - after a member's positions change, the group follows.
Error paths: mixed precision or order, nrhs beyond max_rhs, bad tiles.
"""
import numpy as np
import pytest

from spme_amd import workload as W

pytestmark = pytest.mark.gpu

NRHS = 4


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _members(kset=(32, 48, 64), nsys=7, seed=77):
    """Small uneven systems (N in [300, 5000]) with meshes from kset."""
    u = W.uniform(seed, 3, 2 * nsys)
    out = []
    for s in range(nsys):
        n = int(300 * (5000 / 300) ** u[2 * s])
        phi = 0.05 + 0.25 * float(u[2 * s + 1])
        L = W.box_length(n, phi)
        K = kset[s % len(kset)]
        rc = min(6.0, L / 3.05)
        out.append(W.GroupMember(seed * 1000 + s, n, phi, L, K, 6, 0.5, rc))
    return out


def _build(E, members, prec):
    torch = _torch()
    engines, fs = [], []
    dt = torch.float32 if prec == E.SPME_FP32 else torch.float64
    for m in members:
        e = E.Engine(m.xi, m.K, m.p, m.n, m.L, rc=m.rc, precision=prec, max_rhs=NRHS)
        e.set_positions(W.positions(m.seed, m.n, m.L))
        engines.append(e)
        fs.append(torch.tensor(W.forces(m.seed, m.n, NRHS), device="cuda", dtype=dt))
    return engines, fs


def _separate(engines, fs, parts, alpha=1.0, beta=0.0, u0=None):
    torch = _torch()
    outs = []
    for k, (e, f) in enumerate(zip(engines, fs)):
        u = torch.zeros_like(f) if u0 is None else u0[k].clone()
        e.apply(f, u, alpha=alpha, beta=beta, parts=parts)
        outs.append(u)
    torch.cuda.synchronize()
    return outs


def _equal(a, b):
    torch = _torch()
    return all(torch.equal(x, y) for x, y in zip(a, b))


def _close(a, b, engines, tol=2e-5):
    """Bitwise for members with K <= 96; within tol (relative max norm) for
    larger ones, whose single applies fuse the x and y FFT passes."""
    torch = _torch()
    for x, y, e in zip(a, b, engines):
        if e.info().K <= 96:
            if not torch.equal(x, y):
                return False
        elif float((x - y).abs().max() / y.abs().max()) > tol:
            return False
    return True


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("parts", ["recip", "full"])
def test_group_equals_separate(prec, parts):
    torch = _torch()
    from spme_amd import engine as E
    P = E.SPME_FP32 if prec == "fp32" else E.SPME_FP64
    pm = E.SPME_RECIP if parts == "recip" else E.SPME_FULL
    members = _members()
    engines, fs = _build(E, members, P)
    ref = _separate(engines, fs, pm)
    g = E.Group(engines, NRHS)
    us = [torch.full_like(f, float("nan")) for f in fs]
    g.apply(fs, us, parts=pm, graph=False)
    torch.cuda.synchronize()
    assert _equal(us, ref)
    # graph: capture once, replay with eager work and syncs in between
    for rep in range(3):
        us = [torch.full_like(f, float("nan")) for f in fs]
        g.apply(fs, us, parts=pm, graph=True)
        torch.cuda.synchronize()
        assert _equal(us, ref), f"replay {rep}"
        _ = (fs[0] * 2.0).sum().item()  # unrelated eager work
    # alpha/beta and new output buffers: arguments are re-uploaded
    u0 = [torch.randn_like(f) for f in fs]
    ref2 = _separate(engines, fs, pm, alpha=0.7, beta=1.3, u0=u0)
    us = [u.clone() for u in u0]
    g.apply(fs, us, alpha=0.7, beta=1.3, parts=pm, graph=True)
    torch.cuda.synchronize()
    assert _equal(us, ref2)
    g.close()


def test_group_tiles_and_autotune_are_bitwise_neutral():
    torch = _torch()
    from spme_amd import engine as E
    members = _members()
    engines, fs = _build(E, members, E.SPME_FP32)
    ref = _separate(engines, fs, E.SPME_FULL)
    g = E.Group(engines, NRHS)
    variants = [
        dict(spread_cc=3, spread_zc=8, interp_zc=8),
        dict(spread_cc=6, spread_zc=64, interp_zc=64),
        dict(spread_cc=12, spread_zc=24, interp_zc=16, spread_persist=1, interp_persist=1,
             real_persist=1),
        dict(spread_persist=2, interp_persist=4, real_persist=8),
    ]
    for v in variants:
        g.set_tiles(v)
        for graph in (False, True):
            us = [torch.full_like(f, float("nan")) for f in fs]
            g.apply(fs, us, parts=E.SPME_FULL, graph=graph)
            torch.cuda.synchronize()
            assert _equal(us, ref), (v, graph)
    best = g.autotune(parts=E.SPME_FULL, reps=2)
    assert best == g.tiles()
    print("group autotune:", best, flush=True)
    us = [torch.full_like(f, float("nan")) for f in fs]
    g.apply(fs, us, parts=E.SPME_FULL, graph=True)
    torch.cuda.synchronize()
    assert _equal(us, ref)
    g.close()


@pytest.mark.parametrize("n,K", [(6000, 64), (20000, 128)])
def test_engine_tiles_and_autotune_are_bitwise_neutral(n, K):
    """Every tile shape gives the default's bits; at K = 128 a z chunk of 256
    (clamped to 128) is spread in two sweeps (96 + 32 planes)."""
    torch = _torch()
    from spme_amd import engine as E
    m = W.GroupMember(5151 + K, n, 0.2, W.box_length(n, 0.2), K, 8, 0.5, 6.0)
    for prec, dt in ((E.SPME_FP32, torch.float32), (E.SPME_FP64, torch.float64)):
        e = E.Engine(m.xi, m.K, m.p, m.n, m.L, rc=m.rc, precision=prec, max_rhs=NRHS)
        e.set_positions(W.positions(m.seed, m.n, m.L))
        f = torch.tensor(W.forces(m.seed, m.n, NRHS), device="cuda", dtype=dt)
        ref = e.apply(f, parts=E.SPME_FULL)
        torch.cuda.synchronize()
        for v in [dict(spread_cc=3, spread_zc=9, interp_zc=5), dict(spread_cc=6, spread_zc=64),
                  dict(spread_cc=12, spread_zc=256, interp_zc=64)]:
            e.set_tiles(v)
            u = e.apply(f, parts=E.SPME_FULL)
            torch.cuda.synchronize()
            assert torch.equal(u, ref), v
        best = e.autotune(nrhs=NRHS, reps=2)
        assert best == e.tiles(NRHS)
        u = e.apply(f, parts=E.SPME_FULL)
        torch.cuda.synchronize()
        assert torch.equal(u, ref)
        e.set_tiles(None)
        e.close()


def test_group_follows_new_positions_and_large_meshes():
    torch = _torch()
    from spme_amd import engine as E
    # one member on a 128 mesh takes the large-K instantiation
    members = _members(kset=(32, 128, 48), nsys=4, seed=91)
    engines, fs = _build(E, members, E.SPME_FP32)
    g = E.Group(engines, NRHS)
    us = [torch.zeros_like(f) for f in fs]
    g.apply(fs, us, parts=E.SPME_FULL, graph=True)
    torch.cuda.synchronize()
    assert _close(us, _separate(engines, fs, E.SPME_FULL), engines)
    m = members[2]
    engines[2].set_positions(W.positions(m.seed + 1, m.n, m.L))
    ref = _separate(engines, fs, E.SPME_FULL)
    us = [torch.zeros_like(f) for f in fs]
    g.apply(fs, us, parts=E.SPME_FULL, graph=True)
    torch.cuda.synchronize()
    assert _close(us, ref, engines)
    g.close()


def test_group_errors():
    _torch()
    from spme_amd import engine as E
    from spme_amd.lib import SpmeError
    members = _members(nsys=2)
    a, _ = _build(E, members[:1], E.SPME_FP32)
    b, _ = _build(E, members[1:], E.SPME_FP64)
    with pytest.raises(SpmeError, match="precision"):
        E.Group(a + b, NRHS)
    with pytest.raises(SpmeError, match="nrhs"):
        E.Group(a, NRHS + 1)
    with pytest.raises(SpmeError, match="twice"):
        E.Group(a + a, 1)
    g = E.Group(a, NRHS)
    with pytest.raises(SpmeError, match="spread_cc"):
        g.set_tiles(dict(spread_cc=5))
    with pytest.raises(SpmeError, match="interp_zc"):
        g.set_tiles(dict(interp_zc=257))
    g.close()
