"""The oracle restatement against golden fixtures produced by the reference's
own compiled code (tests/golden/make_golden.py).  These run everywhere (no
reference checkout needed) and are bit-exact where the reference's operation
order is reproducible: RNG streams, forces given the reference's cell table,
and integrator steps given the reference's forces and replayed noise."""
import os

import numpy as np
import pytest

from cases import FORCE_CASES, NOISE_SEED, TRAJ_CASES, force_params, noise_from_streams

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_fastrandom_streams_bit_exact(oracle):
    g = load("fastrandom.npz")
    for s in g["seeds"]:
        ref = g[f"seed{s}"]
        st, mine = oracle.fastrandom_streams(int(s), ref.shape[0], ref.shape[1])
        assert st == 0
        assert np.array_equal(mine, ref), f"seed {s}"
    # the second deviate of each pair is stored as float (fastRandom.h:16)
    ref = g["seed90210"]
    assert np.array_equal(ref[:, 1::2].astype(np.float32).astype(np.float64), ref[:, 1::2])


@pytest.mark.parametrize("case", sorted(FORCE_CASES))
def test_forces_bit_exact_on_reference_table(oracle, case):
    plugin, mask, cfg, n, conc, scale = FORCE_CASES[case]
    g = load(f"forces_{case}.npz")
    box, cell = int(g["box"]), int(g["cell"])
    assert oracle.geometry(n, conc, scale)[:2] == (box, cell)
    sorted4 = oracle.reorder(g["perm"], g["pos"], float(g["radius"]))
    st, f, err = oracle.forces_table(force_params(mask, cfg), sorted4, g["perm"], g["cstart"],
                                     g["cend"], scale, cell, box)
    assert st == 0
    assert np.array_equal(f, g["force"])


@pytest.mark.parametrize("case", sorted(FORCE_CASES))
def test_cell_assignment_matches_reference(oracle, case):
    """Same cell membership as the reference's table (its within-cell order
    comes from an unstable sort and is not pinned)."""
    plugin, mask, cfg, n, conc, scale = FORCE_CASES[case]
    g = load(f"forces_{case}.npz")
    key = oracle.hash(g["pos"], int(g["cell"]), scale)
    perm, cs, ce = oracle.build_cells(key, scale ** 3)
    assert np.array_equal(cs == -1, g["cstart"] == -1)
    assert np.array_equal(cs, g["cstart"]) and np.array_equal(ce, g["cend"])
    for c in np.nonzero(cs >= 0)[0]:
        a = np.sort(perm[cs[c]:ce[c]])
        b = np.sort(g["perm"][g["cstart"][c]:g["cend"][c]])
        assert np.array_equal(a, b)
    # stable: within a cell ascending original index
    for c in np.nonzero(cs >= 0)[0][:200]:
        seg = perm[cs[c]:ce[c]]
        assert np.all(np.diff(seg) > 0)


@pytest.mark.parametrize("case", sorted(FORCE_CASES))
def test_forces_own_table_close(oracle, case):
    """With the oracle's own (stable) cell order only the summation order
    differs from the reference."""
    plugin, mask, cfg, n, conc, scale = FORCE_CASES[case]
    g = load(f"forces_{case}.npz")
    st, f, _ = oracle.forces(force_params(mask, cfg), g["pos"], scale, int(g["cell"]),
                             int(g["box"]))
    assert st == 0
    scale_f = np.abs(g["force"]).max()
    assert np.abs(f - g["force"]).max() <= 1e-13 * scale_f


@pytest.mark.parametrize("case", sorted(TRAJ_CASES))
def test_integrator_steps_bit_exact(oracle, case):
    """Each step of the reference's nextSystem reproduced from the
    reference's own state, forces and replayed fastRandom noise."""
    plugin, mask, cfg, n, conc, scale, nsteps = TRAJ_CASES[case]
    g = load(f"traj_{case}.npz")
    box = float(g["box"])
    c = oracle.gb_coeffs(cfg["kT"], cfg["mass"], cfg["gamma"], cfg["timeStep"])
    st, streams = oracle.fastrandom_streams(NOISE_SEED, n, 3 + 6 * (nsteps - 1))
    noise = noise_from_streams(streams, nsteps)
    pos, pos0, vel, frc = g["pos"], g["pos0"], g["vel"], g["frc"]
    st, p, p0, memx = oracle.gb_first_step(c, box, pos[0], pos0[0], vel[0], frc[1],
                                           noise[0, 3:6].T)
    assert st == 0
    assert np.array_equal(p, pos[1]) and np.array_equal(p0, pos0[1])
    vf, vc = cfg["velFreq"], 0
    for s in range(1, nsteps):
        do_vel = vf == 0 or vc == vf
        st, p, p0, v, memx, _ = oracle.gb_normal_step(c, box, pos[s], pos0[s], vel[s],
                                                      frc[s + 1], frc[s], memx,
                                                      noise[s, 0:3].T, noise[s, 3:6].T,
                                                      do_vel)
        vc = 0 if vc == vf else vc + 1
        assert st == 0
        assert np.array_equal(p, pos[s + 1]), f"step {s} pos"
        assert np.array_equal(p0, pos0[s + 1]), f"step {s} pos0"
        assert np.array_equal(v, vel[s + 1]), f"step {s} vel"


@pytest.mark.parametrize("case", sorted(TRAJ_CASES))
def test_full_run_matches_reference_trajectory(oracle, case):
    plugin, mask, cfg, n, conc, scale, nsteps = TRAJ_CASES[case]
    g = load(f"traj_{case}.npz")
    box, cell = float(g["box"]), float(g["cell"])
    c = oracle.gb_coeffs(cfg["kT"], cfg["mass"], cfg["gamma"], cfg["timeStep"])
    st, streams = oracle.fastrandom_streams(NOISE_SEED, n, 3 + 6 * (nsteps - 1))
    noise = noise_from_streams(streams, nsteps)
    state = dict(pos=g["pos"][0].copy(), pos0=g["pos0"][0].copy(), vel=g["vel"][0].copy(),
                 frc=np.zeros((n, 3)), frc0=np.zeros((n, 3)), memx=np.zeros((n, 3)))
    st, err = oracle.run(force_params(mask, cfg), c, state, nsteps, box, scale, cell,
                         vel_freq=cfg["velFreq"], noise=noise)
    assert st == 0
    assert np.abs(state["pos"] - g["pos"][-1]).max() < 1e-12
    assert np.abs(state["pos0"] - g["pos0"][-1]).max() < 1e-12
    assert np.abs(state["vel"] - g["vel"][-1]).max() < 1e-9
    fs = np.abs(g["frc"][-1]).max()
    assert np.abs(state["frc"] - g["frc"][-1]).max() <= 1e-12 * fs
