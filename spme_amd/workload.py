"""Seeded synthetic workloads: positions, forces and row subsets.

Random numbers come from a counter-based SplitMix64 (Steele, Lea & Flood,
OOPSLA 2014): value i of stream s under seed S is mix(key + (i+1)*G) with
key = mix(S*G ^ (s << 56)), G = 0x9E3779B97F4A7C15.  Uniforms use the top
53 bits; normals use Box-Muller.  Everything is vectorised numpy, so the same
seed gives the same arrays on every machine with this numpy.

The workloads are those the reference's driver sets up (old-version/main.c:
130-260: positions wrapped into [0, L), box from the volume fraction) at the
sizes of the benchmark configurations.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

STREAM_POS, STREAM_FORCE, STREAM_ROWS = 0, 1, 2


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix_u64(seed: int, stream: int, n: int, offset: int = 0):
    with np.errstate(over="ignore"):
        key = _mix(np.uint64(seed) * _G ^ (np.uint64(stream) << np.uint64(56)))
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        return _mix(key + i * _G)


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n uniforms in [0, 1) with 53-bit resolution."""
    u = splitmix_u64(seed, stream, n, offset) >> np.uint64(11)
    return u.astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """n standard normals by Box-Muller on pairs of uniforms."""
    m = (n + 1) // 2
    u = uniform(seed, stream, 2 * m)
    u1 = 1.0 - u[0::2]          # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * np.pi * u2)
    z[1::2] = r * np.sin(2.0 * np.pi * u2)
    return z[:n]


def box_length(n: int, phi: float, a: float = 1.0) -> float:
    """L = (N (4/3) pi a^3 / phi)^(1/3)."""
    return (n * (4.0 / 3.0) * math.pi * a ** 3 / phi) ** (1.0 / 3.0)


def positions(seed: int, n: int, L: float) -> np.ndarray:
    """Uniform positions in [0, L)^3, AoS [N, 3] FP64."""
    p = uniform(seed, STREAM_POS, 3 * n).reshape(n, 3) * L
    # the product L*u can round up to L; wrap as the reference driver does
    # (old-version/main.c:107-121), then clamp the rounding case.
    p = np.where(p >= L, p - L, p)
    return p


def forces(seed: int, n: int, nrhs: int = 1) -> np.ndarray:
    """i.i.d. N(0,1) forces [nrhs, 3N] FP64."""
    return normal(seed, STREAM_FORCE, 3 * n * nrhs).reshape(nrhs, 3 * n)


def rows(seed: int, n: int, count: int) -> np.ndarray:
    """`count` distinct row indices in [0, n), sorted."""
    count = min(count, n)
    key = splitmix_u64(seed, STREAM_ROWS, n)
    return np.sort(np.argsort(key, kind="stable")[:count]).astype(np.int32)


@dataclass
class System:
    """One periodic system: N unit spheres in a cube of side L."""
    n: int
    phi: float
    L: float
    pos: np.ndarray
    seed: int
    extra: dict = field(default_factory=dict)


def make_system(seed: int, n: int, phi: float) -> System:
    L = box_length(n, phi)
    return System(n=n, phi=phi, L=L, pos=positions(seed, n, L), seed=seed)


# Benchmark configurations (sizes only; xi comes from parameter selection).
CONFIGS = {
    # plain FP64, CPU-runnable: N=1000, phi=0.1, K=32, p=6, rc=5
    "c1": dict(seed=1001, n=1000, phi=0.1, K=32, p=6, rc=5.0, nrhs=1),
    # plain vs FP32 single RHS: N=100k, phi=0.2, K=128, p=6, rc=4
    "c2": dict(seed=1002, n=100_000, phi=0.2, K=128, p=6, rc=4.0, nrhs=1),
    # batched: N=200k, phi=0.2, K=192, p=8, rc=4, 16 RHS
    "c3": dict(seed=1003, n=200_000, phi=0.2, K=192, p=8, rc=4.0, nrhs=16),
    # uneven systems: 64 boxes, N in [2e3, 5e4], phi in [0.05, 0.3], K in
    # {32, 48, 64, 96}, p=6, 4 RHS, FP32 (multi_systems below)
    "c4": dict(seed=1004, nsys=64, p=6, nrhs=4),
}


@dataclass
class GroupMember:
    """One system of an uneven multi-system workload, with its parameters."""
    seed: int
    n: int
    phi: float
    L: float
    K: int
    p: int
    xi: float
    rc: float


def multi_systems(seed: int = 1004, nsys: int = 64, n_range=(2000, 50000),
                  phi_range=(0.05, 0.3), meshes=(32, 48, 64, 96), h_max: float = 1.2,
                  p: int = 6, tol: float = 1e-4, nrhs: int = 4):
    """Independent periodic boxes: N log-uniform in n_range, phi uniform in
    phi_range, mesh the smallest of `meshes` with L/K <= h_max (else the
    largest), B-spline order p; xi and rc from the analytic estimate for the
    relative error tol at that mesh.  System s draws its sizes from stream 3
    of `seed` and its positions and forces from seed*1000 + s."""
    from . import params as PR
    u = uniform(seed, 3, 2 * nsys)
    lo, hi = math.log(n_range[0]), math.log(n_range[1])
    out = []
    for s in range(nsys):
        n = int(math.exp(lo + u[2 * s] * (hi - lo)))
        phi = phi_range[0] + (phi_range[1] - phi_range[0]) * float(u[2 * s + 1])
        L = box_length(n, phi)
        K = next((k for k in meshes if L / k <= h_max), meshes[-1])
        a = PR.analytic_select(n, phi, tol, nrhs=nrhs, fixed_K=K, fixed_p=p)
        out.append(GroupMember(seed * 1000 + s, n, phi, L, K, p, a.xi, a.rc))
    return out
