"""The reference driver's particle file format (old-version/main.c:45-127).

    # comment lines start with '#'; blank lines are skipped
    N L                      first other line: particle count, box side
    name x y z a             one line per particle (a, the radius, is unused)

Positions are wrapped into [0, L) with fmod (old-version/main.c:112-123).
The same rules as spme_amd/cli/pme.c, which is the driver that reads them
on the GPU side.
"""
from __future__ import annotations

import math

import numpy as np


def read_xyz(path: str):
    """Return (pos[N, 3] float64 wrapped into [0, L), L).  A particle count
    different from N is reported on stderr as the reference does; fewer rows
    than N, malformed rows or a bad header raise ValueError."""
    import sys
    with open(path) as fh:
        lines = fh.readlines()
    k = 0
    n, L = 0, 0.0
    while k < len(lines):
        tok = lines[k].split()
        k += 1
        if not tok or tok[0].startswith("#"):
            continue
        try:
            n, L = int(tok[0]), float(tok[1])
        except (ValueError, IndexError):
            n = 0
        break
    if n <= 0:
        raise ValueError("Invalid number of particles")
    if not L > 0.0:
        raise ValueError("Invalid box size")
    rows = []
    for line in lines[k:]:
        tok = line.split()
        if not tok or tok[0].startswith("#"):
            continue
        if len(tok) < 4:
            raise ValueError(f"Malformed line in {path}: {line!r}")
        rows.append([float(tok[1]), float(tok[2]), float(tok[3])])
    if len(rows) != n:
        print(f"{len(rows)} particles are found in {path}, but num-particles = {n}",
              file=sys.stderr)
        if len(rows) < n:
            raise ValueError("too few particle rows")
    pos = np.array(rows[:n], dtype=np.float64).reshape(n, 3)
    w = np.vectorize(math.fmod)(pos, L) if pos.size else pos
    return np.where(w >= 0.0, w, L + w), L


def write_xyz(path: str, pos, L: float, radius: float = 1.0, name: str = "H",
              comment: str | None = None) -> None:
    """Write positions in the format read_xyz and the pme driver read
    (coordinates with 17 significant digits, so they read back exactly)."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as fh:
        if comment:
            fh.write(f"# {comment}\n")
        fh.write(f"{pos.shape[0]} {L:.17g}\n")
        for x, y, z in pos:
            fh.write(f"{name} {x:.17g} {y:.17g} {z:.17g} {radius:g}\n")
