"""Benchmark: RPY mobility products u = M.f per second at N = 200,000,
K = 192^3, p = 8, 16 right-hand sides (block-Lanczos width) in the FP32
precision form: spreading, FFTs, interpolation and the real-space pair sums
in FP32, tables and setup in FP64, and the final real + reciprocal sum in
FP64 (SPME_FP32_SUM64: float forces in, double velocities out).  The same
engine with FP32 outputs is timed beside it.  Beside
two FP64 paths on the same positions and parameters, one right-hand side
per apply as the reference's compute_spme loops them
(old-version/spme.c:631-702):
  - plain: the reference's data structures (dense p^3 weights and int32
    offsets per particle, atomic spreading, separate influence and 1/K^3
    scale passes), SPME_FP64_PLAIN;
  - compact: this engine's FP64 path (separable weights, sorted particles,
    fused influence), SPME_FP64.

A step is one batched apply of the full operator (wave-space SPME + real
space + overlap + self) to 16 force vectors = 16 M.f.  Inputs are resident
in HBM before the timed region.  Per-kernel durations come from HIP events
the engine records on the apply stream during the timed region.

A second leg times the uneven multi-system workload (64 boxes, N in
[2e3, 5e4], K in {32, 48, 64, 96}, p = 6, 4 right-hand sides, FP32): 64
separate applies against one grouped launch per stage and against one
captured hipGraph.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-multi]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads: one per core, packed (read when the
# OpenMP runtime starts, i.e. when the CPU leg first loads its library)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
TOL = 1e-4             # target relative error of the batched form
MARGIN = 1.5           # the selection is asked for TOL / MARGIN
# rocprof kernel name of each timed stage (FP32 mesh path with the FP64
# final sum, K = 192, p = 8, 16 RHS)
ROCPROF_NAME = {
    "permute": "k_permute_in<float>",
    "spread": "k_spread<float, 8, 12, 8, 8>",
    "fft_x_fwd": "k_fft_xy_fwd<float, 192, 768, true>",
    "fft_z_infl": "k_fft_z_infl<float, 192, 16, 512, false>",
    "fft_x_inv": "k_fft_xy_inv<float, 192, 768, true>",
    "interp": "k_interp<float, 8, 4, 16, 1024, double>",
    "permute_real": "k_permute_in<float>",
    "real": "k_real_spmv<float, 48, double>",
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_c3.md")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline leg")
    ap.add_argument("--no-multi", action="store_true",
                    help="skip the uneven multi-system leg")
    ap.add_argument("--no-brownian", action="store_true",
                    help="skip the block-Lanczos Brownian displacement leg")
    ap.add_argument("--cpu-rhs", type=int, default=24,
                    help="right-hand sides of the CPU baseline sample (about 0.45 s each at "
                         "p=8 on 16 host threads)")
    ap.add_argument("--pmc-run", action="store_true",
                    help="short run for hardware-counter passes: a plain copy of "
                         "known size (calibration), then 2 batched applies")
    return ap.parse_args()


def c3_params():
    """(xi, rc, note) for the batched workload, whose mesh and order are
    fixed (K = 192, p = 8): the fixed-mesh learned selector asked for the
    target relative error TOL (its labels keep a margin of 1.5 below the
    tolerance); the analytic estimate for TOL / MARGIN if it falls back."""
    from spme_amd import mlp
    from spme_amd import params as PR
    from spme_amd import workload as W
    cfg = W.CONFIGS["c3"]
    xi, rc, how = mlp.select_mesh(cfg["n"], cfg["phi"], TOL, cfg["K"], cfg["p"])
    if how == "learned":
        note = (f"learned (fixed-mesh selector network) for rel. error {TOL:g} at fixed "
                f"K={cfg['K']}, p={cfg['p']}")
        return xi, rc, note
    a = PR.analytic_select(cfg["n"], cfg["phi"], TOL / MARGIN, nrhs=cfg["nrhs"],
                           fixed_K=cfg["K"], fixed_p=cfg["p"])
    note = (f"analytic estimate for rel. error {TOL:g}/{MARGIN:g} at fixed K={cfg['K']}, "
            f"p={cfg['p']} (the learned selector fell back)")
    return a.xi, a.rc, note


def pmc_run():
    """Workload for rocprofv3 --pmc passes.  The fill writes and the copy
    moves exactly 2^28 bytes each way, so FETCH_SIZE/WRITE_SIZE can be
    scaled per byte."""
    import torch
    from spme_amd import engine as E
    from spme_amd import workload as W
    cfg = W.CONFIGS["c3"]
    a = torch.empty(2 ** 26, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)  # calibration: 256 MiB read, 256 MiB written
    torch.cuda.synchronize()
    del a, b
    s = W.make_system(cfg["seed"], cfg["n"], cfg["phi"])
    f = torch.tensor(W.forces(cfg["seed"], cfg["n"], cfg["nrhs"]), device="cuda:0",
                     dtype=torch.float32)
    xi, rc, _ = c3_params()
    eng = E.Engine(xi, cfg["K"], cfg["p"], cfg["n"], s.L, rc=rc,
                   precision=E.SPME_FP32_SUM64, max_rhs=cfg["nrhs"])
    eng.set_positions(torch.tensor(s.pos, device="cuda:0"))
    u = torch.zeros(f.shape, dtype=torch.float64, device="cuda:0")
    for _ in range(2):
        eng.apply(f, u, parts=E.SPME_FULL)
    torch.cuda.synchronize()
    eng.close()
    print("[pmc] done", file=sys.stderr, flush=True)


def algorithmic_bytes(N, K, p, R, s, real_entries, so=None):
    """Compulsory HBM bytes per launch of each kernel: b_P = 3p*s + 12 per
    particle, Kh = (K/2+1)K^2, C = 3R components, s bytes per mesh value and
    input value, so per output value (8 when the final sum is FP64); the
    real-space SpMV reads each stored pair entry once (2 coefficients + a
    column index), each position and force once, and reads and writes the
    output the reciprocal part left."""
    so = s if so is None else so
    C = 3 * R
    K3 = K ** 3
    Kh = (K // 2 + 1) * K * K
    bP = 3 * p * s + 12
    vec = 3 * N * R * s
    vout = 3 * N * R * so
    return {
        # input read once, written in both sorted orders (the full operator)
        "permute": 3 * vec + 8 * N,
        "spread": vec + N * bP + K3 * C * s,
        # fused x + y passes: real mesh in, half spectrum out (and back)
        "fft_x_fwd": K3 * C * s + Kh * C * 2 * s,
        # the influence factors (O(K^2) values) stay in cache
        "fft_z_infl": 2 * Kh * C * 2 * s,
        "fft_x_inv": Kh * C * 2 * s + K3 * C * s,
        "interp": K3 * C * s + N * bP + vout,
        "permute_real": 2 * vec + 4 * N,
        # per stored entry: column index and the coefficients F, G/r^2; the
        # FP64 positions once per particle
        "real": real_entries * (2 * s + 4) + vec + 2 * vout + 24 * N + 4 * N,
        # the whole forward + inverse transform as one unit
        "fft_total": 2 * (K3 * C * s + Kh * C * 2 * s),
    }


def traffic_per_launch(kernel):
    """HBM bytes per launch of a kernel from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE scaled by the copy calibration), or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        m = re.match(r"\|\s*`([^`]+)`\s*\|.*\|\s*([0-9.]+)\s*\|\s*$", line)
        if m and m.group(1) == kernel:
            return float(m.group(2)) * 1e6  # the table gives MB
    return None


def brownian_leg(eng, F, R, stream, dt=1e-3, kT=1.0, tol=1e-4):
    """dX = dt M F + sqrt(2 kT dt) M^{1/2} Z for R realisations at the bench's
    workload: block Lanczos on the batched FP32 apply (tolerance 1e-4)."""
    import torch
    from spme_amd import workload as W
    from spme_amd.lanczos import brownian_displacements
    N = F.shape[1] // 3
    Z = torch.tensor(W.forces(2024, N, R), device=F.device, dtype=F.dtype)
    brownian_displacements(eng, F, Z, dt, kT, tol=tol)   # warm-up
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    dX, info = brownian_displacements(eng, F, Z, dt, kT, tol=tol)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    applies = info.iterations + 1
    return {"what": f"{R} displacements dX = dt M F + sqrt(2 kT dt) M^1/2 Z, block Lanczos "
                    f"(full reorthogonalisation) on the batched FP32 apply, tol {tol}",
            "iterations": info.iterations, "krylov_dimension": info.dimension,
            "change": info.change, "ms": round(wall, 2),
            "ms_gpu_span": round(e0.elapsed_time(e1), 2), "batched_applies": applies,
            "ritz_min": info.min_eig, "ritz_max": info.max_eig}


def hbm_copy_ceiling(gib=2, reps=10):
    """Device-to-device copy of a 2 GiB buffer (8x the 256 MB Infinity Cache):
    read + write bytes per second, the practical ceiling of a streaming
    kernel next to the published 8 TB/s."""
    import torch
    n = gib * (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * n * 4 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return {"GBs": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
            "what": f"{gib} GiB device copy, read + write bytes"}


def hbm_stream_ceilings(gib=2, reps=5):
    """The hand-written streaming kernels (read, write, copy, in-place scale
    of 2 GiB buffers, 8x the 256 MB Infinity Cache) over a sweep of grids
    (4-32 blocks of 256 threads per CU) with and without non-temporal
    accesses: the best HBM rate of each pattern, in bytes moved per second."""
    import torch
    from spme_amd import engine as E
    n = gib * (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.ones_like(a)
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    out = {}
    for mode in ("read", "write", "copy", "scale"):
        best = None
        for nt in (False, True):
            for bpc in (4, 8, 16, 32):
                ms = E.stream_probe(mode, None if mode in ("write", "scale") else a, b,
                                    nontemporal=nt, blocks=bpc * ncu, reps=reps)
                gbs = E.STREAM_TRAFFIC[mode] * 4.0 * n / (ms * 1e-3) / 1e9
                if best is None or gbs > best["GBs"]:
                    best = {"GBs": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
                            "blocks_per_cu": bpc, "nontemporal": nt, "ms": round(ms, 4)}
        out[mode] = best
    del a, b
    torch.cuda.empty_cache()
    out["what"] = (f"hand-written HIP streaming kernels over {gib} GiB buffers, bytes moved "
                   f"(copy and in-place scale: read + write); best of 8 launch shapes")
    return out


def multi_system(steps, warmup):
    """64 uneven systems: separate applies vs grouped launches vs one graph."""
    import torch
    from spme_amd import engine as E
    from spme_amd import workload as W
    members = W.multi_systems()
    R = W.CONFIGS["c4"]["nrhs"]
    engines, fs, us = [], [], []
    for m in members:
        e = E.Engine(m.xi, m.K, m.p, m.n, m.L, rc=m.rc, precision=E.SPME_FP32, max_rhs=R)
        e.set_positions(W.positions(m.seed, m.n, m.L))
        engines.append(e)
        fs.append(torch.tensor(W.forces(m.seed, m.n, R), device="cuda:0", dtype=torch.float32))
        us.append(torch.zeros_like(fs[-1]))
    lib = E.load()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    calls = [(e._h, ctypes.c_void_p(f.data_ptr()), ctypes.c_long(f.stride(0)),
              ctypes.c_void_p(u.data_ptr()), ctypes.c_long(u.stride(0)))
             for e, f, u in zip(engines, fs, us)]
    apply_fn = lib.spme_engine_apply

    def separate():
        for h, fi, li, uo, lo in calls:
            apply_fn(h, R, 1.0, fi, li, 0.0, uo, lo, E.SPME_FULL, st)

    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(steps):
            fn()
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / steps

    ms_sep = timed(separate)
    ref = [u.clone() for u in us]
    g = E.Group(engines, R)
    tiles = g.autotune(parts=E.SPME_FULL, reps=3)
    ms_grp = timed(lambda: g.apply(fs, us, parts=E.SPME_FULL, graph=False))
    ms_graph = timed(lambda: g.apply(fs, us, parts=E.SPME_FULL, graph=True))
    bitwise = all(torch.equal(a, b) for a, b in zip(us, ref))
    g.close()
    for e in engines:
        e.close()
    mf = len(members) * R
    return {
        "workload": f"{len(members)} uneven periodic boxes, N in [2000, 50000] "
                    f"(total {sum(m.n for m in members)}), phi in [0.05, 0.3], "
                    f"K in {{32, 48, 64, 96}}, p=6, {R} RHS each, FP32, full operator",
        "meshes": {str(k): sum(1 for m in members if m.K == k) for k in (32, 48, 64, 96)},
        "separate_ms": round(ms_sep, 4),
        "grouped_ms": round(ms_grp, 4),
        "graph_ms": round(ms_graph, 4),
        "separate_Mf_per_s": round(mf / ms_sep * 1e3, 1),
        "graph_Mf_per_s": round(mf / ms_graph * 1e3, 1),
        "speedup_graph_vs_separate": round(ms_sep / ms_graph, 3),
        "bitwise_equal_to_separate": bitwise,
        "tiles": tiles,
    }


def main():
    args = parse()
    if args.pmc_run:
        pmc_run()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU "
                          "sharding is a design note only, so --gpus must be 1"}))
        sys.exit(2)
    import torch
    from spme_amd import engine as E
    from spme_amd import workload as W

    cfg = W.CONFIGS["c3"]
    N, K, p, R = cfg["n"], cfg["K"], cfg["p"], cfg["nrhs"]
    sysm = W.make_system(cfg["seed"], N, cfg["phi"])
    L = sysm.L
    xi, rc, params_note = c3_params()
    dev = "cuda:0"
    torch.cuda.set_device(0)
    f64 = W.forces(cfg["seed"], N, R).astype(np.float32).astype(np.float64)
    pos_d = torch.tensor(sysm.pos, device=dev)
    f32_d = torch.tensor(f64, device=dev, dtype=torch.float32)
    u64_out = torch.zeros(f32_d.shape, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    # ---- batched FP32 form, final sum in FP64 (the headline) ---------------
    eng = E.Engine(xi, K, p, N, L, rc=rc, precision=E.SPME_FP32_SUM64, max_rhs=R)
    t0 = time.perf_counter()
    eng.set_positions(pos_d)
    torch.cuda.synchronize()
    t_pbuild_first = time.perf_counter() - t0
    # the same build again: the first call of a process also loads the code
    # objects of the build kernels
    t0 = time.perf_counter()
    eng.set_positions(pos_d)
    torch.cuda.synchronize()
    t_pbuild = time.perf_counter() - t0
    tiles = eng.autotune(nrhs=R, reps=3)
    real_entries = eng.info().real_entries
    for _ in range(args.warmup):
        eng.apply(f32_d, u64_out, parts=E.SPME_FULL)
    torch.cuda.synchronize()
    print(f"[bench] warmup done ({args.warmup} steps), tiles {tiles}", file=sys.stderr,
          flush=True)
    eng.set_timing(True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        eng.apply(f32_d, u64_out, parts=E.SPME_FULL)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_total = ev0.elapsed_time(ev1)
    stages = eng.timing()
    eng.set_timing(False)
    ms_step = ms_total / args.steps
    value = R * args.steps / (ms_total / 1e3)
    print(f"[bench] fp32 batched, fp64 sum: {ms_step:.3f} ms/step, {value:.1f} M.f/s "
          f"(wall {wall:.3f}s)", file=sys.stderr, flush=True)
    u_full = u64_out.cpu().numpy()
    eng.close()

    # ---- the same with FP32 outputs; Brownian displacements on it ----------
    u32_d = torch.zeros_like(f32_d)
    eng = E.Engine(xi, K, p, N, L, rc=rc, precision=E.SPME_FP32, max_rhs=R)
    eng.set_positions(pos_d)
    eng.set_tiles(tiles)
    for _ in range(max(1, args.warmup)):
        eng.apply(f32_d, u32_d, parts=E.SPME_FULL)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        eng.apply(f32_d, u32_d, parts=E.SPME_FULL)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_f32out = ev0.elapsed_time(ev1) / args.steps
    u_f32out = u32_d.double().cpu().numpy()
    print(f"[bench] fp32 batched, fp32 outputs: {ms_f32out:.3f} ms/step", file=sys.stderr,
          flush=True)
    brownian = None
    if not args.no_brownian:
        brownian = brownian_leg(eng, f32_d, R, stream)
        print(f"[bench] brownian: {brownian['iterations']} Lanczos steps, "
              f"{brownian['ms']:.1f} ms per block of {R}", file=sys.stderr, flush=True)
    eng.close()

    # ---- FP64 paths, one right-hand side per apply ---------------------------
    plain_steps = args.plain_steps or max(2, args.steps // 2)
    f64_d = torch.tensor(f64, device=dev)
    u64_d = torch.zeros_like(f64_d)

    def fp64_leg(precision):
        e = E.Engine(xi, K, p, N, L, rc=rc, precision=precision, max_rhs=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.set_positions(pos_d)
        torch.cuda.synchronize()
        pb = time.perf_counter() - t0
        for _ in range(max(1, args.warmup)):
            e.apply(f64_d[0], u64_d[0], parts=E.SPME_FULL)
        torch.cuda.synchronize()
        ev0.record(stream)
        for k in range(plain_steps):
            r = k % R
            e.apply(f64_d[r], u64_d[r], parts=E.SPME_FULL)
        ev1.record(stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / plain_steps
        for r in range(2):
            e.apply(f64_d[r], u64_d[r], parts=E.SPME_FULL)
        torch.cuda.synchronize()
        u = u64_d[:2].cpu().numpy()
        e.close()
        return ms, pb, u

    ms_plain, pb_plain, u_plain = fp64_leg(E.SPME_FP64_PLAIN)
    plain_value = 1e3 / ms_plain
    print(f"[bench] fp64 plain (reference structures): {ms_plain:.3f} ms per M.f, "
          f"{plain_value:.1f} M.f/s", file=sys.stderr, flush=True)
    ms_compact, pb_compact, u_compact = fp64_leg(E.SPME_FP64)
    compact_value = 1e3 / ms_compact
    print(f"[bench] fp64 compact: {ms_compact:.3f} ms per M.f, {compact_value:.1f} M.f/s",
          file=sys.stderr, flush=True)

    # ---- error against the FP64 direct Ewald sum (committed fixture) -------
    errors = {}
    rc4_leg = None
    gpath = os.path.join(ROOT, "tests", "golden", "c3.npz")
    if os.path.exists(gpath):
        g = np.load(gpath)
        rows, direct = g["rows"], g["direct"]

        def rel(u, r):
            mine = u[r].reshape(N, 3)[rows].reshape(-1)
            return float(np.linalg.norm(mine - direct[r]) / np.linalg.norm(direct[r]))
        errors = {"fp32_batched_fp64_sum": max(rel(u_full, r) for r in range(direct.shape[0])),
                  "fp32_batched_fp32_out": max(rel(u_f32out, r) for r in range(direct.shape[0])),
                  "fp64_plain": max(rel(u_plain, r) for r in range(direct.shape[0])),
                  "fp64_compact": max(rel(u_compact, r) for r in range(direct.shape[0])),
                  "rows": int(rows.size), "rhs_checked": int(direct.shape[0])}

        # The workload at the cutoff the configuration names (rc = 4a), with
        # the xi that minimises the error estimate there: at K=192, p=8 the
        # real-space truncation keeps it near 1e-2, which is why the
        # headline runs at the selector's larger cutoff.
        from spme_amd import params as PR
        rc4 = cfg["rc"]
        xi4 = min((0.3 + 0.005 * k for k in range(200)),
                  key=lambda x: PR.estimate_error(x, K, p, rc4, cfg["phi"], L))
        e4 = E.Engine(xi4, K, p, N, L, rc=rc4, precision=E.SPME_FP32_SUM64, max_rhs=R)
        e4.set_positions(pos_d)
        e4.set_tiles(tiles)
        u4 = torch.zeros_like(u64_out)
        for _ in range(max(1, args.warmup)):
            e4.apply(f32_d, u4, parts=E.SPME_FULL)
        torch.cuda.synchronize()
        n4 = max(3, args.steps // 2)
        ev0.record(stream)
        for _ in range(n4):
            e4.apply(f32_d, u4, parts=E.SPME_FULL)
        ev1.record(stream)
        torch.cuda.synchronize()
        ms4 = ev0.elapsed_time(ev1) / n4
        u4h = u4.cpu().numpy()
        e4.close()
        rc4_leg = {"rc": rc4, "xi": round(xi4, 4), "ms_per_step": round(ms4, 4),
                   "value": round(R / (ms4 / 1e3), 2), "unit": "M.f/s",
                   "rel_error_vs_fp64_direct_ewald":
                       max(rel(u4h, r) for r in range(direct.shape[0])),
                   "what": "the same batched FP32 apply at rc = 4a with the xi minimising "
                           "the error estimate there (K=192, p=8 fixed)"}
        print(f"[bench] rc=4a: xi {xi4:.3f}, {ms4:.3f} ms/step, error "
              f"{rc4_leg['rel_error_vs_fp64_direct_ewald']:.2e}", file=sys.stderr, flush=True)

    # ---- roofline of the dominant kernel -------------------------------------
    ab = algorithmic_bytes(N, K, p, R, 4, real_entries, so=8)
    per = {}
    for name, (ms, n) in stages.items():
        if n:
            avg = ms / n
            per[name] = {"avg_ms": round(avg, 4),
                         "GBs": round(ab[name] / (avg * 1e-3) / 1e9, 1)}
    dom = max(per, key=lambda k: per[k]["avg_ms"])
    dom_gbs = ab[dom] / (per[dom]["avg_ms"] * 1e-3) / 1e9
    traffic = traffic_per_launch(ROCPROF_NAME.get(dom, ""))
    fft_ms = sum(per[k]["avg_ms"] for k in per if k.startswith("fft_"))
    fft_gbs = ab["fft_total"] / (fft_ms * 1e-3) / 1e9
    mesh_ms = fft_ms + per["spread"]["avg_ms"] + per["interp"]["avg_ms"]
    mesh_bytes = ab["fft_total"] + ab["spread"] + ab["interp"]

    # ---- uneven multi-system leg -------------------------------------------
    multi = None
    if not args.no_multi:
        multi = multi_system(max(3, args.steps // 4), 2)
        print(f"[bench] multi-system: separate {multi['separate_ms']:.3f} ms, graph "
              f"{multi['graph_ms']:.3f} ms", file=sys.stderr, flush=True)

    copy_ceiling = hbm_copy_ceiling()
    print(f"[bench] HBM copy ceiling: {copy_ceiling['GBs']:.0f} GB/s", file=sys.stderr,
          flush=True)
    stream_ceilings = hbm_stream_ceilings()
    print("[bench] HBM streaming kernels: " + ", ".join(
        f"{k} {v['GBs']:.0f} GB/s" for k, v in stream_ceilings.items() if k != "what"),
        file=sys.stderr, flush=True)

    # ---- CPU baseline: restated reference on the host cores ----------------
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        nr = max(1, args.cpu_rhs)
        fcpu = W.forces(cfg["seed"], N, nr).astype(np.float32).astype(np.float64)
        t0 = time.perf_counter()
        O.spme_full(xi, K, p, rc, L, sysm.pos, fcpu)
        tc = time.perf_counter() - t0
        cpu = {"value": round(nr / tc, 4), "unit": "M.f/s", "cores": O.num_threads(),
               "kind": "port",
               "sample": f"{nr} right-hand side(s) of the same N=200k, K=192, p=8, "
                         f"rc={rc:.3f} FP64 full M.f by the restated reference algorithm "
                         f"(P build included; the reference's coloured OpenMP spread with an "
                         f"even block count, tabulated FFT twiddles, OMP_PLACES="
                         f"{os.environ.get('OMP_PLACES')}), {tc:.1f} s"}
        print(f"[bench] cpu: {tc:.2f}s for {nr} M.f on {O.num_threads()} threads",
              file=sys.stderr, flush=True)
        # p = 8 is beyond the reference (old-version/spme.c:86 asserts p in
        # {4, 6}): the same workload at its largest order, same K, xi, rc
        t0 = time.perf_counter()
        O.spme_full(xi, K, 6, rc, L, sysm.pos, fcpu)
        tc6 = time.perf_counter() - t0
        cpu["reference_order_p6"] = {
            "value": round(nr / tc6, 4), "unit": "M.f/s",
            "sample": f"{nr} right-hand side(s), same workload at p=6 (B-spline tables equal "
                      f"the reference's literals), {tc6:.1f} s"}
        print(f"[bench] cpu p=6: {tc6:.2f}s for {nr} M.f", file=sys.stderr, flush=True)

    out = {
        "metric": "RPY M.f products/sec at N=200k, K=192^3 (FP32 batched vs FP64 plain); % HBM",
        "value": round(value, 2),
        "unit": "M.f/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (mesh path; final real + recip sum f64)",
        "data": "synthetic (seeded uniform positions, N(0,1) forces)",
        "config": {"workload": f"RPY M.f, N=200000 unit spheres, phi=0.2, "
                               f"L=161.199, K=192, p=8, 16 RHS batched, full operator, "
                               f"rc={rc:.2f} (the selector's cutoff for error 1e-4; the "
                               f"configuration's rc=4a is timed beside it in at_rc_4a)",
                   "xi": round(xi, 6), "rc": round(rc, 4), "nrhs": R,
                   "params": params_note},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(dom_gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic)},
        "mesh_roofline": {"kernels": "spread + FFT passes + interpolation",
                          "achieved": round(mesh_bytes / (mesh_ms * 1e-3) / 1e9, 1),
                          "frac": round(mesh_bytes / (mesh_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "ms": round(mesh_ms, 4)},
        "fft_roofline": {"achieved": round(fft_gbs, 1),
                         "frac": round(fft_gbs / HBM_PEAK_GBS, 4),
                         "ms": round(fft_ms, 4)},
        "hbm_copy_ceiling": copy_ceiling,
        "hbm_stream_ceilings": stream_ceilings,
        "kernels": per,
        "tiles": tiles,
        "fp32_outputs": {"ms_per_step": round(ms_f32out, 4),
                         "value": round(R / (ms_f32out / 1e3), 2), "unit": "M.f/s",
                         "what": "the same engine and apply with the final sum and outputs "
                                 "in FP32 (SPME_FP32)"},
        "plain_fp64": {"value": round(plain_value, 2), "unit": "M.f/s",
                       "structures": "reference: dense p^3 P + int32 offsets, atomic "
                                     "spreading, separate influence and scale passes",
                       "ms_per_mf": round(ms_plain, 4),
                       "p_build_ms": round(pb_plain * 1e3, 2),
                       "speedup_of_batched_fp32": round(value / plain_value, 2)},
        "compact_fp64": {"value": round(compact_value, 2), "unit": "M.f/s",
                         "ms_per_mf": round(ms_compact, 4),
                         "p_build_ms": round(pb_compact * 1e3, 2),
                         "speedup_of_batched_fp32": round(value / compact_value, 2)},
        "p_build_ms": round(t_pbuild * 1e3, 2),
        "p_build_first_call_ms": round(t_pbuild_first * 1e3, 2),
        "rel_error_vs_fp64_direct_ewald": errors,
        "at_rc_4a": rc4_leg,
        "multi_system": multi,
        "brownian": brownian,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
