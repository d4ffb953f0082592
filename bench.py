"""Throughput of the hot path on one MI355X: particle-steps/s at N = 2,097,152.

Workload (the large memory-bound configuration): N = 2,097,152 colloids
(radius 0.5) at volume fraction 0.30 in a periodic cube of edge
L = 154.112, screened Yukawa repulsion (inverse Debye length 5, contact
strength 20 kT -> yukawaStrength 14841.3 with debyeLength 0.2) plus the LJ
2n-n core (ljNum 18, wellDepth 1), cutoff 3, 51^3 cells of edge 3.022,
G&B integrator dt = 1e-3, gamma = 1, kT = 1, in-kernel Philox noise.
One step = forces + integration + cell rebuild (the system::run loop body).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Prints one JSON line.  The state (positions, velocities, noise generator)
is resident in HBM before the timed region; the timed region is K steps
bracketed by stream synchronisations.
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


def step_bytes(ncell: int) -> float:
    """Algorithmic bytes per particle-step of the whole step, in-kernel noise:
    FP64 integrator state r/w 192 B, cell build 16 B, sorted position copy
    48 B, cell-start table 8 B per cell."""
    return 192.0 + 16.0 + 48.0 + 8.0 * ncell / N


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=25)
    ap.add_argument("--path", default="auto", choices=["auto", "sparse", "dense"])
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


def cpu_baseline(pos_exact, box_exact, steps):
    """The reference's own code (oracle/_ref, compiled from its sources) on
    the host cores, in the reference's integer geometry for this workload
    (boxSize 153, cellSize 3, scale 51); the oracle restatement if the
    reference build is absent."""
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
    sample = (f"N={N}, reference integer geometry L={box_i} cell={cell_i} S={scale_i}, "
              f"{steps} steps incl. the first, OpenMP threads={threads}")
    if reference_available():
        from oracle import Reference
        secs = Reference().run("LJ", cfg, pos, vel, 0.5, 1.0, box_i, cell_i, scale_i, 0.001,
                               steps, nthreads=threads, record=False)
        kind = "reference"
    else:
        from oracle import ForceParams, load_oracle
        o = load_oracle()
        fp = ForceParams.make(2, CUTOFF, kT=1.0, lj_well_depth=1.0, yukawa_strength=YUK_A,
                              debye_length=DEBYE, lj_num=18)
        c = o.gb_coeffs(1.0, 1.0, 1.0, 1e-3)
        st = dict(pos=pos, pos0=np.zeros((N, 3)), vel=vel, frc=np.zeros((N, 3)),
                  frc0=np.zeros((N, 3)), memx=np.zeros((N, 3)))
        t0 = time.perf_counter()
        o.run(fp, c, st, steps, float(box_i), scale_i, float(cell_i), seed=1,
              nthreads=threads)
        secs = time.perf_counter() - t0
        kind = "port"
    return {"value": N * steps / secs, "unit": "particle-steps/s", "cores": threads,
            "kind": kind, "sample": sample, "seconds": secs}


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


def traffic_from_profiles():
    """HBM bytes per k_step launch measured by rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE) and calibrated against a copy kernel; null when absent."""
    path = os.path.join(ROOT, "profiles", "k_step_c3_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def valu_from_profiles(pair_ms):
    """VALU lane-op rate of the pair kernel from a committed rocprofv3 --pmc
    run (SQ_INSTS_VALU x 64 lanes per launch / measured launch time)."""
    path = os.path.join(ROOT, "profiles", "k_step_c3_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        insts = d["valu_insts_per_launch"]
    except (OSError, ValueError, KeyError):
        return None
    rate = insts * 64 / (pair_ms * 1e-3)
    return {"achieved_lane_ops_per_s": rate, "peak": VALU_PEAK_LANE_OPS,
            "frac": rate / VALU_PEAK_LANE_OPS,
            "note": "SQ_INSTS_VALU x 64 counts issued lanes incl. masked ones; the peak "
                    "issues one wave64 instruction per SIMD every 4 cycles (the FP64 FMA "
                    "rate), so a frac near 1 means VALU-issue bound"}


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
    if args.warmup:
        sysg.step_device_noise(args.warmup, 0)
    sysg.sync()
    code = sysg.last_error()[0]
    if code:
        raise SystemExit(f"bench: physics error {code} during warm-up")
    sysg.set_profiling(True)
    sysg.kernel_times(reset=True)
    sysg.sync()
    t0 = time.perf_counter()
    sysg.step_device_noise(args.steps, 0)
    sysg.sync()
    t1 = time.perf_counter()
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
    achieved = pair_bytes * N / (pair_ms * 1e-3) / 1e9
    wall_step_bytes = step_bytes(geom.ncell)
    print(f"bench: {args.steps} steps in {elapsed:.3f}s; pair kernel {pair_ms:.3f} ms, "
          f"G&B kernel {integ_ms:.3f} ms, cell rebuild {build_ms:.3f} ms", file=sys.stderr,
          flush=True)
    sysg.close()
    ensemble = None
    if not args.no_ensemble:
        from proteinsim2015_amd import bench_ensemble
        ensemble = bench_ensemble.run(512, steps=500, baseline_steps=20)
        print(f"bench: ensemble speedup {ensemble['speedup']:.2f}x", file=sys.stderr, flush=True)
    adjoint = None
    if not args.no_adjoint:
        from proteinsim2015_amd import bench_adjoint
        adjoint = bench_adjoint.run()
        print(f"bench: adjoint {adjoint['adjoint_ms_per_step']:.3f} ms per step",
              file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        with stdout_to_stderr():
            cpu = cpu_baseline(pos, geom.box, args.cpu_steps)
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
            "bound": "hbm",
            "kernel": "k_step_dense (27-cell LDS-staged pair forces)" if kt["integrate_launches"]
                      else "k_step (27-cell pair forces + fused G&B)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(),
            "bytes_per_particle": pair_bytes,
            "avg_launch_ms": pair_ms,
            "g_and_b_kernel_ms": integ_ms,
            "cell_rebuild_ms": build_ms,
            "whole_step_bytes_per_particle": wall_step_bytes,
            "whole_step_achieved_gbs": wall_step_bytes * value / 1e9,
            "whole_step_frac": wall_step_bytes * value / (HBM_PEAK_GBS * 1e9),
            "valu": valu_from_profiles(pair_ms),
        },
        "cpu_baseline": cpu,
        "batched_ensemble": ensemble,
        "adjoint": adjoint,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
