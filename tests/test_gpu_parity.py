"""HIP path vs the oracle and the reference's golden fixtures (needs a GPU).

Tolerances: cell assignment and neighbour membership
bit-exact; per-particle forces within 1e-6 relative of FP64 with the metric
|dF_i| / max(|F_i|, sum_j |f_ij|); one-step positions within 1e-12 absolute."""
import os

import numpy as np
import pytest

import proteinsim2015_amd as ps
from cases import (FORCE_CASES, NOISE_SEED, TRAJ_CASES, force_params, noise_from_streams)
from helpers import force_rel_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FORCE_TOL = 1e-6
POS_TOL = 1e-12


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def modules_for(plugin, cfg):
    cls = ps.PLUGINS[plugin]
    return cls.from_config(ps.Config(cfg))


PATHS = ["sparse", "dense", "brick"]


def system_for(plugin, cfg, n, box, cell, scale, integ=None, seed=NOISE_SEED, path="auto"):
    geom = ps.SystemGeometry(n=n, box=float(box), cells_per_dim=int(scale),
                             cell_size=float(cell))
    if integ is None:
        integ = ps.BrownianIntegrator(kT=cfg.get("kT", 1.0), gamma=1.0, dt=1e-3, seed=seed)
    mods = modules_for(plugin, cfg) if isinstance(plugin, str) else plugin
    return ps.GpuSystem(geom, mods, integ, noise_seed=seed, kernel_path=path)


def pair_fmag(fp):
    """|fNet(r)| of the module from its energy (independent of the kernels)."""
    from test_oracle_closed_form import fd, u_ao, u_cal, u_lj_yukawa
    rc = fp.cutoff
    if fp.mask & 2:
        fn = lambda r: fd(lambda x: u_lj_yukawa(x, fp.kT, fp.lj_well_depth, fp.lj_num,  # noqa
                                                fp.yukawa_strength, fp.debye_length), r)
    elif fp.mask & 4:
        fn = lambda r: fd(lambda x: u_ao(x, fp.kT, fp.ao_well_depth, rc - 1), r, 1e-7) / r \
            - 36 * r ** -38  # noqa
    else:
        fn = lambda r: fd(u_cal, r, 1e-7)  # noqa
    return fn


def pair_abs_sum_from_oracle(oracle, fp, pos, scale, cell, box):
    """sum_j |f_ij| per particle, the scale of the per-particle force metric."""
    from helpers import pair_abs_sum
    return pair_abs_sum(pos, box, fp.cutoff, pair_fmag(fp))


@pytest.mark.parametrize("case", sorted(FORCE_CASES))
def test_cell_assignment_bit_exact(oracle, case):
    plugin, mask, cfg, n, conc, scale = FORCE_CASES[case]
    g = load(f"forces_{case}.npz")
    pos, box, cell = g["pos"], int(g["box"]), int(g["cell"])
    print(f"case={case} n={n} box={box} S={scale}", flush=True)
    with system_for(plugin, cfg, n, box, cell, scale) as sysg:
        sysg.set_state(pos)
        key, cs, ob = sysg.cells()
    assert np.array_equal(key, oracle.hash(pos, cell, scale))
    counts = np.bincount(key, minlength=scale ** 3)
    assert np.array_equal(cs, np.concatenate([[0], np.cumsum(counts)]))
    # slot order: sorted by cell; from an identity start, by original index
    assert np.array_equal(np.sort(key[ob], kind="stable"), key[ob])
    for c in np.nonzero(counts)[0]:
        seg = ob[cs[c]:cs[c + 1]]
        assert np.all(np.diff(seg) > 0)
        rs, re_ = g["cstart"][c], g["cend"][c]
        assert np.array_equal(np.sort(seg), np.sort(g["perm"][rs:re_]))


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("case", sorted(FORCE_CASES))
def test_forces_match_reference(oracle, case, path):
    plugin, mask, cfg, n, conc, scale = FORCE_CASES[case]
    g = load(f"forces_{case}.npz")
    pos, box, cell = g["pos"], int(g["box"]), int(g["cell"])
    print(f"case={case} n={n} path={path}", flush=True)
    with system_for(plugin, cfg, n, box, cell, scale, path=path) as sysg:
        sysg.set_state(pos)
        st, f = sysg.forces_now()
    assert st == 0
    fp = force_params(mask, cfg)
    denom = pair_abs_sum_from_oracle(oracle, fp, pos, scale, cell, box)
    err = force_rel_err(f, g["force"], denom)
    assert err.max() <= FORCE_TOL, f"max rel err {err.max():.3g}"


@pytest.mark.parametrize("path", PATHS)
def test_composite_calibration_plus_lj_is_sum_of_reference_modules(oracle, path):
    """Config 1: Calibration + LJ 12-6; the reference evaluates one plugin per
    run (defaultForceManager.cpp:56), so the oracle is the sum of two runs."""
    gc, gl = load("forces_cal.npz"), load("forces_lj6.npz")
    assert np.array_equal(gc["pos"], gl["pos"])
    pos, box, cell, scale = gc["pos"], int(gc["box"]), int(gc["cell"]), int(gc["scale"])
    cfg = FORCE_CASES["lj6"][2]
    mods = [ps.Calibration(cutoff=2.5), ps.LennardJones.from_config(ps.Config(cfg))]
    with system_for(mods, cfg, len(pos), box, cell, scale, path=path) as sysg:
        sysg.set_state(pos)
        st, f = sysg.forces_now()
    assert st == 0
    ref = gc["force"] + gl["force"]
    denom = (pair_abs_sum_from_oracle(oracle, force_params(1, {"cutOff": 2.5}), pos, scale,
                                      cell, box)
             + pair_abs_sum_from_oracle(oracle, force_params(2, cfg), pos, scale, cell, box))
    assert force_rel_err(f, ref, denom).max() <= FORCE_TOL


@pytest.mark.parametrize("case", ["cal", "ljyuk", "ao", "lj_s3"])
def test_force_table_entry_on_reference_arrays(oracle, case):
    """psim_gpu_force_table consumes the reference's own sortedParticles /
    particleHashIndex / cellStartEnd (the plugin adaptor's batch path)."""
    plugin, mask, cfg, n, conc, scale = FORCE_CASES[case]
    g = load(f"forces_{case}.npz")
    pos, box, cell = g["pos"], int(g["box"]), int(g["cell"])
    sorted4 = oracle.reorder(g["perm"], pos, 0.5)
    key = oracle.hash(pos, cell, scale)
    hi = np.stack([key[g["perm"]], g["perm"]], axis=1).astype(np.int32)
    cse = np.stack([g["cstart"], g["cend"]], axis=1).astype(np.int32)
    with system_for(plugin, cfg, n, box, cell, scale) as sysg:
        st, f = sysg.force_table(sorted4, hi, cse)
    assert st == 0
    denom = pair_abs_sum_from_oracle(oracle, force_params(mask, cfg), pos, scale, cell, box)
    assert force_rel_err(f, g["force"], denom).max() <= FORCE_TOL


def traj_setup(case):
    plugin, mask, cfg, n, conc, scale, nsteps = TRAJ_CASES[case]
    g = load(f"traj_{case}.npz")
    integ = ps.BrownianIntegrator(kT=cfg["kT"], mass=cfg["mass"], gamma=cfg["gamma"],
                                  dt=cfg["timeStep"], vel_freq=cfg["velFreq"], seed=NOISE_SEED)
    return plugin, mask, cfg, n, scale, nsteps, g, integ


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("case", sorted(TRAJ_CASES))
def test_single_steps_match_reference(oracle, case, path):
    """Each reference step reproduced from the reference's state: positions
    within 1e-12 absolute (injected, replayed fastRandom noise)."""
    plugin, mask, cfg, n, scale, nsteps, g, integ = traj_setup(case)
    st, streams = oracle.fastrandom_streams(NOISE_SEED, n, 3 + 6 * (nsteps - 1))
    noise = noise_from_streams(streams, nsteps)
    box, cell = float(g["box"]), float(g["cell"])
    vf = cfg["velFreq"]
    with system_for(plugin, cfg, n, box, cell, scale, integ=integ, path=path) as sysg:
        for s in range(nsteps):
            print(f"case={case} step={s}", flush=True)
            memx = noise[s - 1, 3:6].T if s > 0 else None
            vc = (s - 1) % (vf + 1) if s > 0 else 0
            sysg.set_state(g["pos"][s], g["pos0"][s], g["vel"][s], g["frc"][s], memx,
                           time=s * cfg["timeStep"], vel_counter=vc, step_index=s)
            assert sysg.step(1, noise[s:s + 1]) == 0
            out = sysg.get_state()
            assert np.abs(out["pos"] - g["pos"][s + 1]).max() <= POS_TOL
            assert np.abs(out["pos0"] - g["pos0"][s + 1]).max() <= POS_TOL
            assert np.abs(out["vel"] - g["vel"][s + 1]).max() <= 1e-9
            fs = np.abs(g["frc"][s + 1]).max()
            assert np.abs(out["frc"] - g["frc"][s + 1]).max() <= 1e-9 * fs


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("case", sorted(TRAJ_CASES))
def test_trajectory_matches_reference(oracle, case, path):
    plugin, mask, cfg, n, scale, nsteps, g, integ = traj_setup(case)
    st, streams = oracle.fastrandom_streams(NOISE_SEED, n, 3 + 6 * (nsteps - 1))
    noise = noise_from_streams(streams, nsteps)
    with system_for(plugin, cfg, n, float(g["box"]), float(g["cell"]), scale,
                    integ=integ, path=path) as sysg:
        sysg.set_state(g["pos"][0], g["pos0"][0], g["vel"][0])
        assert sysg.step(nsteps, noise) == 0
        out = sysg.get_state()
    assert np.abs(out["pos"] - g["pos"][-1]).max() <= 1e-11
    assert np.abs(out["vel"] - g["vel"][-1]).max() <= 1e-8


@pytest.mark.parametrize("path", PATHS)
def test_in_kernel_noise_matches_oracle_philox(oracle, path):
    """Philox4x32-10 normals generated in the kernel = the oracle's, so a
    free-running trajectory follows the oracle's to rounding."""
    plugin, mask, cfg, n, scale, nsteps, g, integ = traj_setup("ljyuk_vel1")
    nsteps = 20
    box, cell = float(g["box"]), float(g["cell"])
    seed = 777
    with system_for(plugin, cfg, n, box, cell, scale, integ=integ, seed=seed,
                    path=path) as sysg:
        sysg.set_state(g["pos"][0], g["pos0"][0], g["vel"][0])
        assert sysg.step(nsteps) == 0
        out = sysg.get_state()
    c = oracle.gb_coeffs(cfg["kT"], cfg["mass"], cfg["gamma"], cfg["timeStep"])
    state = dict(pos=g["pos"][0].copy(), pos0=g["pos0"][0].copy(), vel=g["vel"][0].copy(),
                 frc=np.zeros((n, 3)), frc0=np.zeros((n, 3)), memx=np.zeros((n, 3)))
    st, _ = oracle.run(force_params(mask, cfg), c, state, nsteps, box, scale, cell,
                       vel_freq=cfg["velFreq"], seed=seed)
    assert st == 0
    assert np.abs(out["pos"] - state["pos"]).max() <= 1e-10
    assert out["step_index"] == nsteps


def test_error_codes():
    cfg = {"cutOff": 2.0}
    # 7704: a pair closer than 0.8 (r_i + r_j)
    pos = np.array([[5.0, 5.0, 5.0], [5.79, 5.0, 5.0], [1.0, 1.0, 1.0]])
    with system_for("Cal", cfg, 3, 10.0, 2.0, 5) as sysg:
        sysg.set_state(pos)
        assert sysg.forces_now()[0] == 7704
        code, i, j = sysg.last_error()
        assert code == 7704 and {i, j} == {0, 1}
        sysg.clear_error()
        assert sysg.last_error()[0] == 0
    # 7703: out of the box after the single wrap of particle::setX
    pos = np.array([[5.0, 5.0, 5.0], [1.0, 1.0, 1.0], [8.0, 8.0, 8.0]])
    vel = np.array([[3e4, 0.0, 0.0], [0, 0, 0], [0, 0, 0]])
    with system_for("Cal", cfg, 3, 10.0, 2.0, 5) as sysg:
        sysg.set_state(pos, vel=vel)
        assert sysg.step(1) == 7703
        assert sysg.last_error()[:2] == (7703, 0)
    # 7705: NaN AO force (AOPotential.cpp:105-109)
    pos = np.array([[5.0, 5.0, 5.0], [6.0, 5.0, 5.0], [1.0, 1.0, 1.0]])
    with system_for(ps.AOPotential(well_depth=float("nan"), cutoff=1.1), cfg, 3, 10.0, 2.0,
                    5) as sysg:
        sysg.set_state(pos)
        assert sysg.forces_now()[0] == 7705


_C3_SCALE = {}


@pytest.mark.parametrize("path", PATHS)
def test_full_size_c3_properties(oracle, path):
    """N = 2,097,152 Yukawa + LJ at phi 0.30: a few steps, then size-independent
    properties and a full force comparison with the oracle."""
    n = 2097152
    geom = ps.SystemGeometry.exact(n, 0.30, 3.0)
    pos = ps.rsa_positions(n, geom.box, 1234)
    cfg = FORCE_CASES["ljyuk"][2]
    lj = ps.LennardJones.from_config(ps.Config(cfg))
    integ = ps.BrownianIntegrator(kT=1.0, gamma=1.0, dt=1e-3, seed=90210)
    print("c3: building", flush=True)
    with ps.GpuSystem(geom, lj, integ, kernel_path=path) as sysg:
        rng = np.random.default_rng(3)
        sysg.set_state(pos, vel=rng.standard_normal((n, 3)))
        assert sysg.step(3) == 0
        key, cs, ob = sysg.cells()
        st, f = sysg.forces_now()
        out = sysg.get_state()
    assert st == 0
    assert cs[-1] == n and np.all(np.diff(cs) >= 0)
    assert np.array_equal(np.sort(ob), np.arange(n))
    assert np.array_equal(key, oracle.hash(out["pos"], geom.cell_size, geom.cells_per_dim))
    assert np.abs(f.sum(axis=0)).max() <= 1e-9 * np.abs(f).sum()
    fp = force_params(2, cfg)
    sto, fo, _ = oracle.forces(fp, out["pos"], geom.cells_per_dim, geom.cell_size, geom.box,
                               half_uv=geom.box / 2)
    assert sto == 0
    # every particle, relative to max(|F_i|, sum_j |f_ij|): a pair dropped or
    # added at the cutoff (|f| ~ 1.6e-3 kT/sigma at r = 3) shows
    # The scale is computed once and reused by the other kernel paths: their
    # positions agree to round-off, far below what moves a pair sum.
    from helpers import pair_abs_sum_all
    if _C3_SCALE.get("pos") is None or np.abs(_C3_SCALE["pos"] - out["pos"]).max() > 1e-9:
        print("c3: pair sums", flush=True)
        _C3_SCALE["pos"] = out["pos"]
        _C3_SCALE["denom"] = pair_abs_sum_all(out["pos"], geom.box, fp.cutoff, pair_fmag(fp))
    denom = _C3_SCALE["denom"]
    err = force_rel_err(f, fo, denom)
    assert err.max() <= FORCE_TOL, f"max {err.max():.3g} at {int(err.argmax())}"
    assert np.abs(f - fo).max() <= FORCE_TOL * np.abs(fo).max()


def test_c1_thousand_steps_follow_the_oracle(oracle):
    """C1 at its own size and length: N = 4096, Calibration + LJ 12-6 (eps 1,
    cutoff 2.5), phi 0.20 (exact L, 8^3 cells), dt 1e-3, 1000 steps with
    in-kernel Philox noise, against the oracle's trajectory with the same
    noise.  The two differ by rounding only, which the dynamics amplifies
    slowly: positions within 1e-8 after 100, 500 and 1000 steps, and the
    neighbour sets equal at the end."""
    from oracle.oracle import PO_CAL, PO_LJ, ForceParams
    n = 4096
    geom = ps.SystemGeometry.exact(n, 0.20, 2.5)
    mods = [ps.Calibration(cutoff=2.5),
            ps.LennardJones(kT=1.0, well_depth=1.0, yukawa_strength=0.0, lj_num=6, cutoff=2.5,
                            debye_length=0.5)]
    integ = ps.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=1e-3, vel_freq=1000,
                                  seed=90210)
    pos = ps.rsa_positions(n, geom.box, 1234)
    vel = np.random.default_rng(1235).standard_normal((n, 3))
    marks = (100, 500, 1000)
    print("c1: 1000 steps", flush=True)
    gpu = []
    with ps.GpuSystem(geom, mods, integ) as sysg:
        sysg.set_state(pos, vel=vel)
        done = 0
        for m in marks:
            assert sysg.step(m - done) == 0
            done = m
            gpu.append(sysg.get_state()["pos"])
        cnt, hs = sysg.pair_stats()
    L, S, cell = geom.box, geom.cells_per_dim, geom.cell_size
    fp = ForceParams.make(PO_CAL | PO_LJ, 2.5, kT=1.0, lj_well_depth=1.0, yukawa_strength=0.0,
                          debye_length=0.5, lj_num=6)
    c = oracle.gb_coeffs(1.0, 1.0, 1.0, 1e-3)
    state = dict(pos=pos.copy(), pos0=np.where(np.abs(pos) > L / 2, np.where(pos < 0, -L, L), 0.0),
                 vel=vel.copy(), frc=np.zeros((n, 3)), frc0=np.zeros((n, 3)),
                 memx=np.zeros((n, 3)))
    done = 0
    for m, gp in zip(marks, gpu):
        print(f"c1: oracle to step {m}", flush=True)
        st, _ = oracle.run(fp, c, state, m - done, L, S, cell, vel_freq=1000, seed=90210)
        assert st == 0
        done = m
        d = gp - state["pos"]
        err = np.abs(d - L * np.round(d / L)).max()   # a wrap on one side only
        print(f"c1: step {m}: max |dx| {err:.3g}", flush=True)
        assert err <= 1e-8, (m, err)
    ocnt, ohs = oracle.pair_stats(fp, state["pos"], S, cell, L)
    assert np.array_equal(cnt, ocnt) and np.array_equal(hs, ohs)
