"""Throughput of the hot path on one MI355X: particle-steps/s at N = 2,097,152.

Workload (the large memory-bound configuration): N = 2,097,152 colloids
(radius 0.5) at volume fraction 0.30 in a periodic cube of edge
L = 154.112, screened Yukawa repulsion (inverse Debye length 5, contact
strength 20 kT -> yukawaStrength 14841.3 with debyeLength 0.2) plus the LJ
2n-n core (ljNum 18, wellDepth 1), cutoff 3, 51^3 cells of edge 3.022,
G&B integrator dt = 1e-3, gamma = 1, kT = 1, in-kernel Philox noise.
One step = forces + integration + cell rebuild (the system::run loop body).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--no-small] [--no-ensemble] [--no-adjoint]

Prints one JSON line.  The state (positions, velocities, noise generator)
is resident in HBM before the timed region; the timed region is K steps
bracketed by stream synchronisations.  Beside the headline the line
carries: the whole-step roofline (algorithmic bytes per particle-step
against the HBM peak) with the pair kernel's VALU issue fraction, the
pair-work counters, the N = 4096 Calibration + LJ and N = 262,144 AO
workloads, the 512-replica batched ensemble against one system per
launch, the trajectory adjoint's step time, and the reference's own CPU
code timed on the host cores (median of 3 runs).
"""
from __future__ import annotations

import argparse
import contextlib
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 2097152
PHI = 0.30
CUTOFF = 3.0
DEBYE = 0.2
YUK_A = 20.0 / (DEBYE * math.exp(-1.0 / DEBYE))
HBM_PEAK_GBS = 8000.0
VALU_PEAK_LANE_OPS = 39.3e12   # gfx950: 256 CUs x 64 FP64/non-packed FP32 lanes x 2.4 GHz
# Algorithmic HBM bytes per particle (neighbour re-reads from LDS/L2 not counted):
#   pair kernel (k_step_dense): read x (24 B), write F (24 B)
#   G&B kernel (k_integrate): read x, x0, F_prev, memX, F, id (124 B), write
#     x', x0', memX', id, key, rank (84 B)
PAIR_KERNEL_BYTES = 48.0
INTEGRATE_KERNEL_BYTES = 208.0


def step_bytes(ncell: int, n: int = N) -> float:
    """Algorithmic bytes per particle-step of the whole step, in-kernel noise:
    FP64 integrator state r/w 192 B, cell build 16 B, sorted position copy
    48 B, cell-start table 8 B per cell."""
    return 192.0 + 16.0 + 48.0 + 8.0 * ncell / n


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-small", action="store_true",
                    help="skip the N=4096 Calibration+LJ and N=262,144 AO workloads")
    ap.add_argument("--path", default="auto", choices=["auto", "sparse", "dense"])
    ap.add_argument("--graph", type=int, default=0,
                    help="replay the headline steps as hipGraphs of this many (even) steps; "
                         "the kernel split then comes from a separate profiled pass")
    ap.add_argument("--eager", action="store_true",
                    help="launch every step of the smaller workloads eagerly instead of as "
                         "hipGraph replays (for tracing under a profiler)")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 512-replica batched-ensemble measurement")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiable AO fit (adjoint step time) measurement")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 3 device-to-device copies of 512 MiB (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def calibrate_copy():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 512 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(nbytes)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipMemset(b, 0, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:   # device to device
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes_per_launch": nbytes, "launches": 3}))


@contextlib.contextmanager
def stdout_to_stderr():
    """Send what native code prints to stdout (the reference's console
    messages) to stderr, so that stdout carries only the JSON line."""
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    libc.fflush(None)
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        yield
    finally:
        libc.fflush(None)
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)


def cpu_model() -> str:
    """The host CPU's model name (first entry of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pos_exact, box_exact, steps, runs=3):
    """The reference's own code (oracle/_ref, compiled from its sources) on
    the host cores, in the reference's integer geometry for this workload
    (boxSize 153, cellSize 3, scale 51): ``steps`` steady-state steps after
    one untimed first step, median of ``runs`` runs; the oracle restatement
    if the reference build is absent."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import reference_available
    box_i, cell_i, scale_i = 153, 3, 51
    pos = np.ascontiguousarray(pos_exact * (box_i / box_exact))
    pos = np.minimum(pos, np.nextafter(box_i, 0))
    vel = np.random.default_rng(7).standard_normal((N, 3))
    cfg = {"cutOff": CUTOFF, "kT": 1.0, "wellDepth": 1.0, "yukawaStrength": YUK_A,
           "ljNum": 18, "debyeLength": DEBYE, "nParticles": N, "gamma": 1.0, "mass": 1.0,
           "timeStep": 0.001, "velFreq": 1000, "seed": 1}
    secs = []
    if reference_available():
        from oracle import Reference
        ref = Reference()
        for _ in range(runs):
            secs.append(ref.run("LJ", cfg, pos, vel, 0.5, 1.0, box_i, cell_i, scale_i, 0.001,
                                steps + 1, nthreads=threads, record=False, skip=1))
        kind = "reference"
    else:
        from oracle import ForceParams, load_oracle
        o = load_oracle()
        fp = ForceParams.make(2, CUTOFF, kT=1.0, lj_well_depth=1.0, yukawa_strength=YUK_A,
                              debye_length=DEBYE, lj_num=18)
        c = o.gb_coeffs(1.0, 1.0, 1.0, 1e-3)
        for _ in range(runs):
            st = dict(pos=pos.copy(), pos0=np.zeros((N, 3)), vel=vel.copy(),
                      frc=np.zeros((N, 3)), frc0=np.zeros((N, 3)), memx=np.zeros((N, 3)))
            o.run(fp, c, st, 1, float(box_i), scale_i, float(cell_i), seed=1, nthreads=threads)
            t0 = time.perf_counter()
            o.run(fp, c, st, steps, float(box_i), scale_i, float(cell_i), seed=1,
                  nthreads=threads)
            secs.append(time.perf_counter() - t0)
        kind = "port"
    med = float(np.median(secs))
    sample = (f"N={N}, reference integer geometry L={box_i} cell={cell_i} S={scale_i}: "
              f"{steps} steady-state steps after one untimed step, median of {runs} runs "
              f"({', '.join(f'{x:.2f}' for x in secs)} s), OpenMP threads={threads}")
    return {"value": N * steps / med, "unit": "particle-steps/s", "cores": threads,
            "kind": kind, "sample": sample, "seconds": med, "cpu_model": cpu_model(),
            "runs_seconds": secs}


def cpu_adjoint_baseline(steps=20):
    """The oracle's FP64 CPU trajectory adjoint (forward with stored states,
    then backward) on the adjoint workload's start, a bounded sample of the
    K = 200 trajectory.  The reference has no derivatives, so the CPU side
    is the oracle's restatement ("port")."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    import proteinsim2015_amd as ps
    from proteinsim2015_amd import fit
    from oracle import ForceParams, load_oracle
    from oracle.oracle import PO_AO
    o = load_oracle()
    n = 32768
    geom = ps.SystemGeometry.exact(n, 0.15, 1.2)
    pos = ps.rsa_positions(n, geom.box, 1234)
    vel = np.random.default_rng(1235).standard_normal((n, 3))
    L = geom.box
    pos0 = np.where(np.abs(pos) > L / 2, np.where(pos < 0, -L, L), 0.0)
    phi_p, cut, _, _ = fit.ao_params(4.0, 0.10)
    fp = ForceParams.make(PO_AO, cut, kT=1.0, ao_well_depth=phi_p)
    c = o.gb_coeffs(1.0, 1.0, 1.0, 1e-3)
    gx = np.random.default_rng(3).standard_normal((n, 3))
    t0 = time.perf_counter()
    st, _, _, _ = o.adjoint(fp, c, dict(pos=pos, pos0=pos0, vel=vel), steps, L,
                            geom.cells_per_dim, geom.cell_size, gx, vel_freq=1000, seed=90210,
                            nthreads=threads)
    secs = time.perf_counter() - t0
    if st:
        raise RuntimeError(f"CPU adjoint failed ({st})")
    return {"ms_per_step": secs / steps * 1e3, "unit": "ms per (forward + backward) step",
            "cores": threads, "kind": "port",
            "sample": f"N={n}, {steps} steps of the adjoint workload's trajectory, "
                      f"OpenMP threads={threads}", "seconds": secs}


def c3_profile() -> dict:
    """The committed rocprofv3 figures for this workload (HBM bytes from
    FETCH_SIZE / WRITE_SIZE scaled by a copy kernel, VALU instructions of
    the pair kernel); empty when absent."""
    path = os.path.join(ROOT, "profiles", "k_step_c3_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def valu_from_profiles(prof, pair_ms):
    """VALU issue of the pair kernel: SQ_INSTS_VALU (committed rocprofv3
    --pmc run) over the live launch time, against one wave64 instruction per
    SIMD every 4 cycles (the FP64 FMA rate, 39.3 T lane-ops/s)."""
    insts = prof.get("valu_insts_per_launch")
    if not insts or not pair_ms:
        return None
    rate = insts * 64 / (pair_ms * 1e-3)
    return {"kernel": prof.get("kernel"), "insts_per_launch": insts,
            "insts_per_particle_step_lanes": insts * 64 / N,
            "achieved_lane_ops_per_s": rate, "peak": VALU_PEAK_LANE_OPS,
            "frac": rate / VALU_PEAK_LANE_OPS,
            "note": "SQ_INSTS_VALU x 64 (issued lanes incl. masked ones) from the committed "
                    "pmc run over this run's launch time, against 1024 SIMDs x 16 lanes x "
                    "2.4 GHz; the profiled launches (slower under the profiler) put the "
                    "kernel at 94-98% of that model, so a frac near or slightly above 1 reads as "
                    "VALU-issue bound, with the waves' waits binding about as tightly (a cull "
                    "with 3.6% fewer VALU instructions ran in the same time)"}


def small_leg(name: str, steps: int, warmup: int = 10, eager: bool = False) -> dict:
    """One of the smaller plain-engine workloads, timed like the headline."""
    import proteinsim2015_amd as ps
    if name == "c1":
        n, phi = 4096, 0.20
        geom = ps.SystemGeometry.exact(n, phi, 2.5)
        mods = [ps.Calibration(cutoff=2.5),
                ps.LennardJones(kT=1.0, well_depth=1.0, yukawa_strength=0.0, lj_num=6,
                                cutoff=2.5, debye_length=0.5)]
        dt = 1e-3
        wl = (f"N=4096 Calibration + LJ 12-6 (eps 1, cutoff 2.5), phi=0.20, L={geom.box:.3f}, "
              f"{geom.cells_per_dim}^3 cells, dt=1e-3, {steps} steps")
    else:
        n, phi = 262144, 0.15
        geom = ps.SystemGeometry.exact(n, phi, 1.1)
        mods = [ps.AOPotential(kT=1.0, well_depth=0.25, cutoff=1.1)]
        dt = 5e-4
        wl = (f"N=262,144 AO q=0.1 (phi_p 0.25, 4 kT at contact, cutoff 1.1) + r^-38 core, "
              f"phi=0.15, L={geom.box:.3f}, {geom.cells_per_dim}^3 cells, dt=5e-4, "
              f"{steps} steps")
    integ = ps.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=dt, vel_freq=1000, seed=90210)
    pos = ps.rsa_positions(n, geom.box, 1234)
    vel = np.random.default_rng(4321).standard_normal((n, 3))
    with ps.GpuSystem(geom, mods, integ) as sysg:
        sysg.set_state(pos, vel=vel)
        # steps of a few tens of microseconds: replayed as hipGraphs of 50
        if not eager:
            sysg.set_graph(50)
        sysg.step_device_noise(max(warmup, 50), 0)
        sysg.sync()
        t0 = time.perf_counter()
        sysg.step_device_noise(steps, 0)
        sysg.sync()
        el = time.perf_counter() - t0
        if sysg.last_error()[0]:
            raise SystemExit(f"bench: physics error {sysg.last_error()[0]} in {name}")
        # kernel split from a separate profiled pass (events per step would
        # weigh on steps of a few microseconds)
        sysg.set_profiling(True)
        sysg.kernel_times(reset=True)
        sysg.step_device_noise(min(steps, 200), 0)
        kt = sysg.kernel_times(reset=True)
        sysg.set_profiling(False)
        cnt, _ = sysg.pair_stats()
        info = sysg.kernel_info()
    value = n * steps / el
    b_alg = step_bytes(geom.ncell, n)
    pair_ms = kt["force_ms"] / max(1, kt["force_launches"])
    return {"workload": wl, "n_particles": n, "steps": steps, "kernel_path": info["path"],
            "ms_per_step": el / steps * 1e3, "particle_steps_per_s": value,
            "launch": "eager launches" if eager else "hipGraph replays of 50 steps",
            "pair_kernel_ms": pair_ms,
            "g_and_b_kernel_ms": kt["integrate_ms"] / max(1, kt["integrate_launches"])
            if kt["integrate_launches"] else 0.0,
            "cell_rebuild_ms": kt["build_ms"] / max(1, kt["build_launches"]),
            "bytes_per_particle_step": b_alg,
            "hbm_frac": b_alg * value / (HBM_PEAK_GBS * 1e9),
            "candidates_per_particle_step": 27.0 * n / geom.ncell,
            "in_range_pairs_per_particle_step": float(cnt.mean())}


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)
    import proteinsim2015_amd as ps
    geom = ps.SystemGeometry.exact(N, PHI, CUTOFF)
    t_setup = time.perf_counter()
    pos = ps.rsa_positions(N, geom.box, 1234)
    vel = np.random.default_rng(4321).standard_normal((N, 3))
    lj = ps.LennardJones(kT=1.0, well_depth=1.0, yukawa_strength=YUK_A, lj_num=18,
                         cutoff=CUTOFF, debye_length=DEBYE)
    integ = ps.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=1e-3, vel_freq=1000,
                                  seed=90210)
    sysg = ps.GpuSystem(geom, lj, integ, kernel_path=args.path)
    sysg.set_state(pos, vel=vel)
    setup_s = time.perf_counter() - t_setup
    print(f"bench: setup {setup_s:.1f}s (RSA + upload), warmup {args.warmup}", file=sys.stderr,
          flush=True)
    if args.graph:
        sysg.set_graph(args.graph)
    if args.warmup:
        sysg.step_device_noise(max(args.warmup, args.graph), 0)
    sysg.sync()
    code = sysg.last_error()[0]
    if code:
        raise SystemExit(f"bench: physics error {code} during warm-up")
    if not args.graph:
        # per-kernel HIP events on the context's stream inside the timed region
        sysg.set_profiling(True)
        sysg.kernel_times(reset=True)
    sysg.sync()
    t0 = time.perf_counter()
    sysg.step_device_noise(args.steps, 0)
    sysg.sync()
    t1 = time.perf_counter()
    if args.graph:
        sysg.set_profiling(True)
        sysg.kernel_times(reset=True)
        sysg.step_device_noise(min(args.steps, 50), 0)
    kt = sysg.kernel_times(reset=True)
    sysg.set_profiling(False)
    code = sysg.last_error()[0]
    if code:
        raise SystemExit(f"bench: physics error {code} in the timed region")
    elapsed = t1 - t0
    value = N * args.steps / elapsed
    pair_ms = kt["force_ms"] / max(1, kt["force_launches"])
    integ_ms = kt["integrate_ms"] / max(1, kt["integrate_launches"]) \
        if kt["integrate_launches"] else 0.0
    build_ms = kt["build_ms"] / max(1, kt["build_launches"])
    pair_bytes = PAIR_KERNEL_BYTES if kt["integrate_launches"] else \
        PAIR_KERNEL_BYTES + INTEGRATE_KERNEL_BYTES - 48.0
    pair_gbs = pair_bytes * N / (pair_ms * 1e-3) / 1e9
    b_alg = step_bytes(geom.ncell)
    achieved = b_alg * value / 1e9
    print(f"bench: {args.steps} steps in {elapsed:.3f}s; pair kernel {pair_ms:.3f} ms, "
          f"G&B kernel {integ_ms:.3f} ms, cell rebuild {build_ms:.3f} ms", file=sys.stderr,
          flush=True)
    cnt, _ = sysg.pair_stats()
    in_range = float(cnt.mean())
    info = sysg.kernel_info()
    sysg.close()
    prof = c3_profile()
    small = {}
    if not args.no_small:
        for name, steps in (("c1", 1000), ("c2", 10000)):
            small[name] = small_leg(name, steps, eager=args.eager)
            print(f"bench: {name} {small[name]['particle_steps_per_s']:.3g} particle-steps/s "
                  f"({small[name]['ms_per_step']:.4f} ms per step)", file=sys.stderr,
                  flush=True)
    sysg.close()
    ensemble = None
    if not args.no_ensemble:
        from proteinsim2015_amd import bench_ensemble
        ensemble = bench_ensemble.run(512, steps=500, baseline_steps=20)
        print(f"bench: ensemble speedup {ensemble['speedup']:.2f}x over eager launches, "
              f"{ensemble['speedup_vs_graph_baseline']:.2f}x over per-replica graph replay",
              file=sys.stderr, flush=True)
    adjoint = None
    if not args.no_adjoint:
        from proteinsim2015_amd import bench_adjoint
        adjoint = bench_adjoint.run()
        print(f"bench: adjoint {adjoint['adjoint_ms_per_step']:.3f} ms per step",
              file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        with stdout_to_stderr():
            cpu = cpu_baseline(pos, geom.box, args.cpu_steps, args.cpu_runs)
        if adjoint is not None:
            acpu = cpu_adjoint_baseline()
            adjoint["cpu_baseline"] = acpu
            adjoint["gpu_ms_per_forward_plus_backward_step"] = \
                adjoint["forward_ms_per_step"] + adjoint["adjoint_ms_per_step"]
            print(f"bench: CPU adjoint {acpu['ms_per_step']:.1f} ms per step "
                  f"({acpu['cores']} threads)", file=sys.stderr, flush=True)
    line = {
        "metric": "particle-steps/s at N=2M",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded random sequential addition positions, Gaussian velocities",
        "config": {
            "workload": "N=2,097,152 Yukawa+LJ core, phi=0.30, L=154.112, cutoff 3, 51^3 cells, "
                        "dt=1e-3, gamma=1, kT=1, in-kernel Philox noise",
            "n_particles": N, "volume_fraction": PHI, "box": geom.box,
            "cells_per_dim": geom.cells_per_dim, "cutoff": CUTOFF,
            "potential": "LJ 36-18 (wellDepth 1) + Yukawa (debyeLength 0.2, "
                         "yukawaStrength 14841.3)",
            "parallelism": "single GPU",
        },
        "roofline": {
            # the whole step against the HBM peak on the algorithmic bytes per
            # particle-step; the step's dominant kernel is bound by VALU issue
            "bound": "valu",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": prof.get("step_bytes_per_step"),
            "bytes_per_particle_step": b_alg,
            "traffic_note": "achieved = B_alg x particle-steps/s with B_alg = 192 + 16 + 48 + "
                            "8 ncell/N bytes per particle-step; traffic = HBM bytes per step "
                            "of the step's kernels from rocprofv3 FETCH_SIZE + WRITE_SIZE "
                            "scaled by a copy kernel (profiles/)",
            "valu": valu_from_profiles(prof, pair_ms),
            "pair_kernel": {
                "kernel": f"k_step_dense<2,18> ({info['path']} path)" if kt["integrate_launches"]
                          else f"k_step ({info['path']} path, fused G&B)",
                "avg_launch_ms": pair_ms,
                "bytes_per_particle": pair_bytes,
                "achieved_gbs": pair_gbs,
                "hbm_frac": pair_gbs / HBM_PEAK_GBS,
                "traffic": prof.get("bytes_per_launch"),
            },
            "g_and_b_kernel_ms": integ_ms,
            "cell_rebuild_ms": build_ms,
        },
        "pair_work": {
            "candidates_per_particle_step": 27.0 * N / geom.ncell,
            "in_range_pairs_per_particle_step": in_range,
            "note": "candidates = particles in the 27-cell stencil (incl. self); in-range = "
                    "pairs the kernel's exact test accepts (psim_gpu_pair_stats)",
        },
        "other_workloads": small or None,
        "cpu_baseline": cpu,
        "batched_ensemble": ensemble,
        "adjoint": adjoint,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
