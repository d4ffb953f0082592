"""Write the golden fixtures under tests/golden/ from the reference's own code.

Runs oracle/_ref/libpsimref.so (the reference's plugins, force manager,
particle, integrator and fastRandom compiled from its sources by
oracle/ref/Makefile) on seeded inputs and stores inputs + outputs as .npz:

  fastrandom.npz        fastRandom::g250 streams, threadSeed = seed*(i+1)
  forces_<case>.npz     positions, the reference's cell table (sorted order,
                        start/end) and particleForce from the plugin
  traj_<case>.npz       initial state and the particle state after every
                        step of the system::run loop (forces, push,
                        nextSystem), for 4-5 steps

Usage (needs the reference checkout, i.e. only where oracle/_ref builds):
    make -C oracle/ref REF=/root/reference && python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import Reference  # noqa: E402
from oracle.oracle import load_oracle  # noqa: E402
from cases import FORCE_CASES, NOISE_SEED, TRAJ_CASES  # noqa: E402
from helpers import maxwell_velocities, rsa_positions  # noqa: E402

FASTRANDOM_SEEDS = (90210, 1, 12345, 7, 31337)


def main():
    ref = Reference()
    o = load_oracle()
    # -- RNG streams --------------------------------------------------------
    streams = {f"seed{s}": ref.fastrandom(s, 48, 45) for s in FASTRANDOM_SEEDS}
    np.savez_compressed(os.path.join(HERE, "fastrandom.npz"),
                        seeds=np.array(FASTRANDOM_SEEDS), **streams)
    # -- forces -------------------------------------------------------------
    for name, (plugin, mask, cfg, n, conc, scale) in FORCE_CASES.items():
        box, cell, phi = o.geometry(n, conc, scale)
        pos = rsa_positions(n, box, seed=1000 + n)
        f, perm, cs, ce, s4 = ref.forces(plugin, cfg, pos, 0.5, box, cell, scale)
        np.savez_compressed(os.path.join(HERE, f"forces_{name}.npz"), pos=pos, perm=perm,
                            cstart=cs, cend=ce, force=f, box=box, cell=cell, scale=scale,
                            radius=0.5)
        print(f"forces_{name}: n={n} box={box} cell={cell} S={scale} phi={phi:.4f} "
              f"|F|max={np.abs(f).max():.3g}")
    # -- trajectories ---------------------------------------------------------
    for name, (plugin, mask, cfg, n, conc, scale, nsteps) in TRAJ_CASES.items():
        box, cell, phi = o.geometry(n, conc, scale)
        pos = rsa_positions(n, box, seed=2000 + n)
        vel = maxwell_velocities(n, cfg["kT"], seed=3000 + n)
        full = dict(cfg, nParticles=n, seed=NOISE_SEED)
        tr = ref.run(plugin, full, pos, vel, 0.5, cfg["mass"], box, cell, scale,
                     cfg["timeStep"], nsteps)
        np.savez_compressed(os.path.join(HERE, f"traj_{name}.npz"), pos_in=pos, vel_in=vel,
                            pos=tr["pos"], pos0=tr["pos0"], vel=tr["vel"], frc=tr["frc"],
                            box=box, cell=cell, scale=scale, nsteps=nsteps)
        print(f"traj_{name}: n={n} box={box} S={scale} steps={nsteps}")


if __name__ == "__main__":
    main()
