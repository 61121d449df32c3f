"""AddressSanitizer and UndefinedBehaviorSanitizer builds of the host C code
(oracle/Makefile, target sanitize), run on the CPU:
  - the oracle (oracle/spme_oracle.c), every function driven once at small
    sizes by oracle/check_main.c, with leak detection on;
  - the pme driver (spme_amd/cli/pme.c): options, xyz reading and its error
    paths, which return before any GPU work (leak detection off: the HIP
    runtime the driver links keeps allocations of its own until exit).
Any sanitizer report aborts the program (-fno-sanitize-recover=all) and
fails the test."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "oracle", "_san")
LIB = os.path.join(ROOT, "spme_amd", "libspme_hip.so")


def _clean(r):
    err = r.stderr + r.stdout
    assert "AddressSanitizer" not in err and "runtime error:" not in err, err[-3000:]
    assert "LeakSanitizer" not in err, err[-3000:]


@pytest.fixture(scope="module")
def built():
    target = ["_san/check_oracle"] + (["_san/pme"] if os.path.exists(LIB) else [])
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), *target],
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + r.stdout + r.stderr)
    return True


def test_oracle_under_sanitizers(built):
    r = subprocess.run([os.path.join(SAN, "check_oracle")], capture_output=True, text=True,
                       timeout=300, env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    _clean(r)
    assert r.returncode == 0 and "ok" in r.stdout


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_pme_driver_under_sanitizers(built, tmp_path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    pme = os.path.join(SAN, "pme")
    good = tmp_path / "good.xyz"
    good.write_text("# comment\n3 10.0\nA 1.0 2.0 3.0 1.0\n# inside\nB -1.0 11.0 5.0 1.0\n"
                    "C 9.99 0.5 20.5 1.0\n")
    short = tmp_path / "short.xyz"
    short.write_text("4 10.0\nA 1 2 3 1\nB 1 2 3 1\n")
    bad = tmp_path / "bad.xyz"
    bad.write_text("2 10.0\nA 1 2 3 1\nB 1 x 3 1\n")
    empty = tmp_path / "empty.xyz"
    empty.write_text("")
    cases = [(["-h"], None), (["-v"], 0), (["--xyz", str(good), "--check-xyz"], 0),
             (["--xyz", str(short), "--check-xyz"], "nonzero"),
             (["--xyz", str(bad), "--check-xyz"], "nonzero"),
             (["--xyz", str(empty), "--check-xyz"], "nonzero"),
             (["--xyz", str(tmp_path / "missing.xyz"), "--check-xyz"], "nonzero"),
             (["--bogus"], "nonzero")]
    for args, want in cases:
        r = subprocess.run([pme, *args], capture_output=True, text=True, timeout=60, env=env)
        _clean(r)
        if want == 0:
            assert r.returncode == 0, (args, r.stderr)
        elif want == "nonzero":
            assert r.returncode != 0, args
    r = subprocess.run([pme, "--xyz", str(good), "--check-xyz"], capture_output=True,
                       text=True, timeout=60, env=env)
    assert r.stdout.split()[:2] == ["3", "10"]
