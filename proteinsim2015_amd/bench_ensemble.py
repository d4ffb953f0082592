"""Batched-ensemble throughput: 512 independent AO replicas advanced together
versus the same replicas advanced one system per launch.

Replica parameters are seeded: N uniform in [500, 20000], volume fraction
uniform in [0.05, 0.35], AO contact depth uniform in [2, 8] kT, q uniform in
[0.05, 0.2] (cutoff 1 + q); radius 0.5, G&B dt 5e-4, gamma 1, kT 1,
in-kernel noise.  Both runs take steps from identical initial states.

    python -m proteinsim2015_amd.bench_ensemble [--replicas 512] [--steps 500]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import engine as E


def make_replicas(nrep: int, seed: int = 512, nmin: int = 500, nmax: int = 20000):
    rng = np.random.default_rng(seed)
    reps = []
    for _ in range(nrep):
        n = int(rng.integers(nmin, nmax + 1))
        phi = rng.uniform(0.05, 0.35)
        u = rng.uniform(2.0, 8.0)
        q = rng.uniform(0.05, 0.2)
        g = E.SystemGeometry.exact(n, phi, 1.0 + q)
        ao = E.AOPotential(kT=1.0, well_depth=E.AOPotential.well_depth_for_contact(u, q),
                           cutoff=1.0 + q)
        reps.append(E.Replica(g, ao))
    return reps


def initial_states(reps, seed: int = 1000):
    pos = [E.rsa_positions(r.geometry.n, r.geometry.box, seed + k) for k, r in enumerate(reps)]
    rng = np.random.default_rng(seed)
    vel = [rng.standard_normal((r.geometry.n, 3)) for r in reps]
    return pos, vel


def run(nrep: int = 512, steps: int = 500, baseline_steps: int = 20, graph_steps: int = 50,
        warmup: int = 5, verbose: bool = False, eager: bool = False) -> dict:
    """eager: no hipGraph replays anywhere (the batched steps launched one by
    one, the per-replica graph baseline skipped), e.g. under a profiler."""
    integ = E.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=5e-4, vel_freq=1000, seed=90210)
    t0 = time.perf_counter()
    reps = make_replicas(nrep)
    pos, vel = initial_states(reps)
    ntot = sum(r.geometry.n for r in reps)
    if verbose:
        print(f"ensemble: {nrep} replicas, {ntot} particles, setup "
              f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    # ---- batched: one launch per kernel stage for all replicas, graph replay
    with E.Ensemble(reps, integ) as ens:
        ens.set_state(np.concatenate(pos), np.concatenate(vel))
        kernel, tile, _ = ens.autotune(10)
        classes = ens.class_choices()
        merged = ens.merged
        if not eager:
            ens.set_graph(graph_steps)
        ens.step(warmup)
        ens.sync()
        t1 = time.perf_counter()
        ens.step(steps)
        ens.sync()
        t2 = time.perf_counter()
        code = ens.last_error()[0]
    if code:
        raise RuntimeError(f"ensemble physics error {code}")
    batched = ntot * steps / (t2 - t1)
    # ---- one system per launch: each replica its own context and stream,
    # launched eagerly (4 launches per step) and as one hipGraph replay of
    # baseline_steps steps per replica (the single-system engine's best mode)
    systems = []
    for r, p, v in zip(reps, pos, vel):
        s = E.GpuSystem(r.geometry, r.forces, integ, kernel_path="auto")
        s.set_state(p, vel=v)
        systems.append(s)
    for s in systems:
        s.step_device_noise(warmup, 0)
    for s in systems:
        s.sync()
    t3 = time.perf_counter()
    for s in systems:
        s.step_device_noise(baseline_steps, 0)
    for s in systems:
        s.sync()
    t4 = time.perf_counter()
    t5 = t6 = None
    if not eager:
        for s in systems:
            s.set_graph(baseline_steps)
            s.step_device_noise(baseline_steps, 0)    # captures, then replays once
        for s in systems:
            s.sync()
        t5 = time.perf_counter()
        for s in systems:
            s.step_device_noise(baseline_steps, 0)
        for s in systems:
            s.sync()
        t6 = time.perf_counter()
    codes = [s.last_error()[0] for s in systems]
    for s in systems:
        s.close()
    if any(codes):
        raise RuntimeError(f"per-replica physics error {max(codes)}")
    single = ntot * baseline_steps / (t4 - t3)
    single_graph = ntot * baseline_steps / (t6 - t5) if t6 else None
    return {
        "replicas": nrep, "particles": ntot, "steps": steps,
        "batched_particle_steps_per_s": batched,
        "batched_ms_per_step": (t2 - t1) / steps * 1e3,
        "per_replica_particle_steps_per_s": single,
        "per_replica_ms_per_step": (t4 - t3) / baseline_steps * 1e3,
        "per_replica_steps_timed": baseline_steps,
        "per_replica_graph_particle_steps_per_s": single_graph,
        "per_replica_graph_ms_per_step": (t6 - t5) / baseline_steps * 1e3 if t6 else None,
        "speedup": batched / single,
        "speedup_vs_graph_baseline": batched / single_graph if single_graph else None,
        "autotuned_kernel": kernel,
        "autotuned_tile": tile,
        "one_pair_launch_for_all": merged,
        "occupancy_classes": [{"particles_per_cell": lim, "replicas": n, "kernel": k, "tile": t}
                              for lim, (n, k, t) in zip(("<0.25", "<0.5", "<1", ">=1"),
                                                        classes)],
        "graph_steps": None if eager else graph_steps,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--baseline-steps", type=int, default=20)
    ap.add_argument("--eager", action="store_true",
                    help="no hipGraph replays (for tracing under a profiler)")
    args = ap.parse_args()
    print(json.dumps(run(args.replicas, args.steps, args.baseline_steps, verbose=True,
                         eager=args.eager)))


if __name__ == "__main__":
    main()
