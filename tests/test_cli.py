"""The reference driver's input format and command line (old-version/main.c):
spme_amd.xyz (Python) and the pme driver (spme_amd/cli/pme.c).

CPU: the xyz rules (comments, blank lines, header, wrap into [0, L), count
mismatch), the driver reading exactly what the Python reader reads, and the
driver's option and error handling (none of which touches the GPU).
GPU: the driver's run (srand(1234) inputs, v <- 2.2 v + 2.1 M_recip f with
positions passed on every call) against the oracle on the driver's own
dumped inputs, <= 1e-12 relative.
"""
import os
import subprocess

import numpy as np
import pytest

from spme_amd import workload as W
from spme_amd.xyz import read_xyz, write_xyz

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PME = os.path.join(ROOT, "spme_amd", "pme")


def _pme(*args):
    if not os.path.exists(PME):
        pytest.skip("pme driver not built (python -m spme_amd.build)")
    return subprocess.run([PME, *args], capture_output=True, text=True, timeout=120)


XYZ_TEXT = """# configuration with comments
#   and a second comment line

3 10.0
# a comment between rows
H  1.5  -2.0  10.0  1.0

O  25.25  3.0  -12.5  1.0
H  -0.0  9.999  0.0  1.0
"""


def test_xyz_rules(tmp_path):
    p = tmp_path / "a.xyz"
    p.write_text(XYZ_TEXT)
    pos, L = read_xyz(str(p))
    assert L == 10.0
    # fmod wrap: -2 -> 8, 10 -> 0, 25.25 -> 5.25, -12.5 -> 7.5
    expect = np.array([[1.5, 8.0, 0.0], [5.25, 3.0, 7.5], [0.0, 9.999, 0.0]])
    np.testing.assert_array_equal(pos, expect)
    assert np.all((pos >= 0.0) & (pos < L))


def test_xyz_roundtrip_exact(tmp_path):
    s = W.make_system(7, 257, 0.15)
    p = tmp_path / "b.xyz"
    write_xyz(str(p), s.pos, s.L, comment="seeded")
    pos, L = read_xyz(str(p))
    assert L == s.L
    np.testing.assert_array_equal(pos, s.pos)


def test_xyz_errors(tmp_path):
    bad = tmp_path / "bad.xyz"
    bad.write_text("# only a comment\n0 5.0\n")
    with pytest.raises(ValueError, match="number of particles"):
        read_xyz(str(bad))
    bad.write_text("2 -1.0\nH 0 0 0 1\nH 1 1 1 1\n")
    with pytest.raises(ValueError, match="box size"):
        read_xyz(str(bad))
    bad.write_text("3 5.0\nH 0 0 0 1\nH 1 1 1 1\n")
    with pytest.raises(ValueError, match="too few"):
        read_xyz(str(bad))
    # more rows than announced: reported, the first N kept
    bad.write_text("1 5.0\nH 1 2 3 1\nH 4 4 4 1\n")
    pos, _ = read_xyz(str(bad))
    np.testing.assert_array_equal(pos, [[1.0, 2.0, 3.0]])


def test_driver_reads_like_python(tmp_path):
    p = tmp_path / "a.xyz"
    p.write_text(XYZ_TEXT)
    s = W.make_system(9, 500, 0.2)
    q = tmp_path / "c.xyz"
    write_xyz(str(q), s.pos - 0.37 * s.L, s.L)   # negative coordinates to wrap
    for path in (p, q):
        r = _pme("--xyz", str(path), "--check-xyz")
        assert r.returncode == 0, r.stderr
        rows = [list(map(float, ln.split())) for ln in r.stdout.strip().splitlines()]
        n, L = int(rows[0][0]), rows[0][1]
        pos, Lp = read_xyz(str(path))
        assert n == pos.shape[0] and L == Lp
        np.testing.assert_array_equal(np.array(rows[1:]), pos)


def test_driver_options_and_errors(tmp_path):
    r = _pme("--version")
    assert r.returncode == 0 and "version 0.1.0" in r.stdout
    r = _pme("-h")
    assert r.returncode == 0 and "--porder" in r.stderr and "--xyz" in r.stderr
    r = _pme("--xyz", str(tmp_path / "missing.xyz"))
    assert r.returncode != 0 and "Can not open file" in r.stderr
    p = tmp_path / "a.xyz"
    p.write_text(XYZ_TEXT)
    # old-version/main.c:205: dim must be larger than 4 * porder
    r = _pme("--xyz", str(p), "--dim", "16", "--porder", "4")
    assert r.returncode != 0 and "dim must be larger than 4 * porder" in r.stderr
    r = _pme("--bogus")
    assert r.returncode != 0
    short = tmp_path / "short.xyz"
    short.write_text("4 5.0\nH 0 0 0 1\n")
    r = _pme("--xyz", str(short), "--check-xyz")
    assert r.returncode != 0 and "but num-particles = 4" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("dim,porder,nrhs", [(32, 6, 2), (48, 8, 1)])
def test_driver_run_matches_oracle(oracle, tmp_path, dim, porder, nrhs):
    """The reference's run (N = 1000, phi = 0.1, xi = 0.5, alpha = 2.1,
    beta = 2.2) through the driver, checked on its own dumped inputs."""
    print(f"case: pme --dim {dim} --porder {porder} --nrhs {nrhs}, N=1000", flush=True)
    s = W.make_system(1001, 1000, 0.1)
    p = tmp_path / "N1000_Phi0.1.xyz"
    write_xyz(str(p), s.pos, s.L)
    out = str(tmp_path / "run")
    r = _pme("--xyz", str(p), "--dim", str(dim), "--porder", str(porder),
             "--nrhs", str(nrhs), "--out", out)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "imported 1000 particles" in r.stdout
    n3 = 3000
    f = np.fromfile(out + ".f", dtype="<f8").reshape(nrhs, n3)
    v0 = np.fromfile(out + ".v0", dtype="<f8").reshape(nrhs, n3)
    v = np.fromfile(out + ".v", dtype="<f8").reshape(nrhs, n3)
    assert f.max() > 1e6          # rand() * 3.0, as the reference fills them
    pos, L = read_xyz(str(p))
    ref = oracle.compute_spme(0.5, dim, porder, L, pos, f, alpha=2.1, beta=2.2, u=v0)
    rel = float(np.abs(v - ref).max() / np.abs(ref).max())
    assert rel < 1e-12, rel
    # the operator's share alone (beta v0 dominates the magnitudes)
    mf, mf_ref = v - 2.2 * v0, ref - 2.2 * v0
    rel = float(np.abs(mf - mf_ref).max() / np.abs(mf_ref).max())
    assert rel < 1e-10, rel
