"""GPU force plugins driven by a PSim host (needs a GPU).

The strongest check: the reference's own compiled force manager and
integrator (oracle/_ref, built from the reference sources) dlopen the GPU
plugin exactly as the reference host does and run with it."""
import os
import subprocess

import numpy as np
import pytest

from cases import FORCE_CASES, NOISE_SEED, TRAJ_CASES, force_params
from helpers import force_rel_err, pair_abs_sum
from test_gpu_parity import load, pair_fmag

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUGIN_DIR = os.path.join(ROOT, "proteinsim2015_amd", "plugins")


def plugin_path(plugin):
    return os.path.join(PLUGIN_DIR, f"{plugin}.so")


@pytest.mark.parametrize("case", ["cal", "lj6", "ljyuk", "ljdefault", "ao", "lj_s3"])
def test_reference_host_with_gpu_plugin_forces(reference, case):
    plugin, mask, cfg, n, conc, scale = FORCE_CASES[case]
    g = load(f"forces_{case}.npz")
    pos, box, cell = g["pos"], int(g["box"]), int(g["cell"])
    print(f"case={case}", flush=True)
    f, perm, cs, ce, s4 = reference.forces(plugin_path(plugin), cfg, pos, 0.5, box, cell, scale)
    fp = force_params(mask, cfg)
    denom = pair_abs_sum(pos, box, fp.cutoff, pair_fmag(fp))
    assert force_rel_err(f, g["force"], denom).max() <= 1e-6


@pytest.mark.parametrize("case", sorted(TRAJ_CASES))
def test_reference_integrator_with_gpu_plugin(reference, case):
    """system::run loop of the reference (its integrator, particle and RNG)
    with the GPU plugin supplying forces: same trajectory as with the
    reference plugin."""
    plugin, mask, cfg, n, conc, scale, nsteps = TRAJ_CASES[case]
    g = load(f"traj_{case}.npz")
    full = dict(cfg, nParticles=n, seed=NOISE_SEED)
    tr = reference.run(plugin_path(plugin), full, g["pos_in"], g["vel_in"], 0.5, cfg["mass"],
                       int(g["box"]), int(g["cell"]), scale, cfg["timeStep"], nsteps)
    assert np.abs(tr["pos"][-1] - g["pos"][-1]).max() <= 1e-11
    assert np.abs(tr["vel"][-1] - g["vel"][-1]).max() <= 1e-8


@pytest.mark.parametrize("mode", ["each", "batch"])
def test_demo_host_per_particle_and_batch(oracle, tmp_path, mode):
    plugin, mask, cfg, n, conc, scale = FORCE_CASES["ljyuk"]
    g = load("forces_ljyuk.npz")
    settings = dict(cfg, nParticles=n, conc=conc, scale=scale, radius=0.5)
    (tmp_path / "settings.cfg").write_text("".join(f"{k} = {v!r}\n" if isinstance(v, float)
                                                   else f"{k} = {v}\n"
                                                   for k, v in settings.items()))
    g["pos"].astype(np.float64).tofile(tmp_path / "pos.f64")
    r = subprocess.run([os.path.join(PLUGIN_DIR, "psim_host_demo"), plugin_path(plugin),
                        str(tmp_path / "settings.cfg"), str(tmp_path / "pos.f64"),
                        str(tmp_path / "f.f64"), mode], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    f = np.fromfile(tmp_path / "f.f64").reshape(-1, 3)
    fp = force_params(mask, cfg)
    denom = pair_abs_sum(g["pos"], int(g["box"]), fp.cutoff, pair_fmag(fp))
    assert force_rel_err(f, g["force"], denom).max() <= 1e-6


def test_per_particle_calls_cost_one_evaluation_per_step(tmp_path):
    """N = 262,144 AO through the unchanged per-particle contract (one
    getAcceleration call per sorted index, from OpenMP threads): each step
    costs one GPU evaluation plus O(1) per call, in the same range as the
    one-call batch entry, and gives the same forces bit for bit."""
    import proteinsim2015_amd as ps
    from cases import PHI_P_C2
    n, conc, scale = 262144, 0.15, 48          # reference geometry: box 96, cell 2
    settings = {"nParticles": n, "conc": conc, "scale": scale, "radius": 0.5,
                "cutOff": 1.1, "kT": 1.0, "wellDepth": PHI_P_C2}
    (tmp_path / "settings.cfg").write_text("".join(f"{k} = {v!r}\n" if isinstance(v, float)
                                                   else f"{k} = {v}\n"
                                                   for k, v in settings.items()))
    ps.rsa_positions(n, 96.0, 1234).astype(np.float64).tofile(tmp_path / "pos.f64")
    secs, forces = {}, {}
    for mode in ("batch", "each"):
        print(f"plugin 262144: mode={mode}", flush=True)
        r = subprocess.run([os.path.join(PLUGIN_DIR, "psim_host_demo"), plugin_path("AO"),
                            str(tmp_path / "settings.cfg"), str(tmp_path / "pos.f64"),
                            str(tmp_path / f"f_{mode}.f64"), mode, "6"],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        line = [x for x in r.stdout.splitlines() if x.startswith("seconds per evaluation")][0]
        secs[mode] = float(line.split(":")[1])
        forces[mode] = np.fromfile(tmp_path / f"f_{mode}.f64")
    print(f"plugin 262144: {secs}", flush=True)
    assert np.array_equal(forces["each"], forces["batch"])
    assert secs["each"] <= 10 * secs["batch"] + 0.05
