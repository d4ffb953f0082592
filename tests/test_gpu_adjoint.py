"""Trajectory adjoint on the GPU vs the FP64 CPU adjoint of the oracle and
vs central differences of GPU trajectories (needs a GPU).

The loss is linear in the final positions, L = sum w . x_K, so dL/dx_K = w.
GPU and oracle trajectories agree to rounding, so their adjoints agree to
1e-8 relative.  The finite differences use a relative parameter step small
enough that no pair crosses the cutoff between the two runs (the adjoint is
pathwise) and are compared at 1e-4 relative, the bar for adjoint gradients.
The cutoff window's crossing term is compared with the oracle's (whose own
test pins it against central differences).
"""
import numpy as np
import pytest

import proteinsim2015_amd as ps
from helpers import rsa_positions
from oracle.oracle import PO_AO, PO_CAL, PO_LJ, ForceParams

pytestmark = pytest.mark.gpu

CASES = {
    # name: (forces kwargs, n, L, S, K, every, seed, parameters checked)
    "ao": (dict(ao=dict(well_depth=0.4, cutoff=1.12)), 600, 16.0, 13, 25, 7, 3,
           ["ao_well_depth", "cutoff"]),
    "cal_lj_yukawa_s3": (dict(cal=dict(cutoff=2.5),
                              lj=dict(well_depth=1.0, yukawa_strength=30.0, debye_length=0.3,
                                      lj_num=6, cutoff=2.5)),
                         150, 11.0, 3, 12, 5, 4,
                         ["lj_well_depth", "yukawa_strength", "debye_length"]),
}


def modules(kw):
    mods = []
    if "cal" in kw:
        mods.append(ps.Calibration(**kw["cal"]))
    if "lj" in kw:
        mods.append(ps.LennardJones(kT=1.0, **kw["lj"]))
    if "ao" in kw:
        mods.append(ps.AOPotential(kT=1.0, **kw["ao"]))
    return mods


def oracle_params(kw):
    mask, extra, cut = 0, {}, None
    if "cal" in kw:
        mask |= PO_CAL
        cut = kw["cal"]["cutoff"]
    if "lj" in kw:
        mask |= PO_LJ
        lj = kw["lj"]
        extra.update(lj_well_depth=lj["well_depth"], yukawa_strength=lj["yukawa_strength"],
                     debye_length=lj["debye_length"], lj_num=lj["lj_num"])
        cut = lj["cutoff"]
    if "ao" in kw:
        mask |= PO_AO
        extra.update(ao_well_depth=kw["ao"]["well_depth"])
        cut = kw["ao"]["cutoff"]
    return ForceParams.make(mask, cut, kT=1.0, **extra)


def make(case, path="auto"):
    kw, n, L, S, K, every, seed, _ = CASES[case]
    geom = ps.SystemGeometry(n=n, box=L, cells_per_dim=S, cell_size=L / S, integer=False)
    integ = ps.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=1e-3, vel_freq=4, seed=seed)
    sysg = ps.GpuSystem(geom, modules(kw), integ, kernel_path=path)
    pos = rsa_positions(n, L, seed)
    vel = np.random.default_rng(seed).standard_normal((n, 3))
    sysg.set_state(pos, vel=vel)
    return sysg, pos, vel


@pytest.mark.parametrize("path", ["sparse", "dense"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_adjoint_matches_oracle_adjoint(oracle, case, path):
    kw, n, L, S, K, every, seed, names = CASES[case]
    print(f"adjoint case={case} path={path}", flush=True)
    sysg, pos, vel = make(case, path)
    w = np.random.default_rng(99).standard_normal((n, 3))
    with sysg, ps.Adjoint(sysg, K, every) as adj:
        assert adj.forward() == 0
        xk = sysg.get_state()["pos"]
        g = adj.backward(w, with_gx0=True)
        assert np.array_equal(sysg.get_state()["pos"], xk)   # left at step K
    c = oracle.gb_coeffs(1.0, 1.0, 1.0, 1e-3)
    p0 = np.where(np.abs(pos) > L / 2, np.where(pos < 0, -L, L), 0.0)
    st, dth, gx0, oxk = oracle.adjoint(oracle_params(kw), c, dict(pos=pos, pos0=p0, vel=vel),
                                       K, L, S, L / S, w, vel_freq=4, seed=seed)
    assert st == 0
    assert np.abs(xk - oxk).max() < 1e-10
    for k, name in enumerate(ps.engine.ADJOINT_PARAMS):
        assert g[name] == pytest.approx(dth[k], rel=1e-8, abs=1e-14), name
    assert np.abs(g["gx0"] - gx0).max() <= 1e-8 * np.abs(gx0).max()


@pytest.mark.parametrize("path", ["sparse", "dense"])
def test_cutoff_window_matches_oracle(oracle, path):
    """The cutoff-crossing term (window 0.01, about 14 pairs per step at this
    size) against the oracle's: same trajectory to rounding, so the same
    window pairs and the same sum to 1e-8."""
    case, win = "ao", 0.01
    kw, n, L, S, K, every, seed, _ = CASES[case]
    print(f"adjoint window path={path}", flush=True)
    sysg, pos, vel = make(case, path)
    w = np.random.default_rng(98).standard_normal((n, 3))
    with sysg, ps.Adjoint(sysg, K, every) as adj:
        adj.set_cutoff_window(win)
        assert adj.forward() == 0
        g = adj.backward(w)
        adj.set_cutoff_window(0.0)
        g0 = adj.backward(w)
    c = oracle.gb_coeffs(1.0, 1.0, 1.0, 1e-3)
    p0 = np.where(np.abs(pos) > L / 2, np.where(pos < 0, -L, L), 0.0)
    st, dth, _, _ = oracle.adjoint(oracle_params(kw), c, dict(pos=pos, pos0=p0, vel=vel), K, L,
                                   S, L / S, w, vel_freq=4, seed=seed, cut_window=win)
    assert st == 0
    for k, name in enumerate(ps.engine.ADJOINT_PARAMS):
        assert g[name] == pytest.approx(dth[k], rel=1e-8, abs=1e-14), name
    # only the cutoff derivative moves, and by much more than the tolerance
    assert abs(g["cutoff"] - g0["cutoff"]) > 1e-3 * abs(g0["cutoff"])
    assert all(g[k] == g0[k] for k in ps.engine.ADJOINT_PARAMS if k != "cutoff")


def test_cutoff_window_beyond_the_cell_edge_is_refused():
    sysg, _, _ = make("ao")      # cutoff 1.12, cell edge 16/13 = 1.2308
    with sysg, ps.Adjoint(sysg, 3, 2) as adj:
        assert adj.forward() == 0
        adj.set_cutoff_window(0.2)
        with pytest.raises(ps.PsimError):
            adj.backward(np.zeros((sysg.n, 3)))
        adj.set_cutoff_window(0.1)
        adj.backward(np.zeros((sysg.n, 3)))
        with pytest.raises(ps.PsimError):
            adj.set_cutoff_window(-1.0)


def test_record_all_matches_checkpointing():
    """Every state recorded by the forward pass (no recomputation) gives the
    gradients of the checkpointed segments bit for bit, and switching back
    works."""
    case = "ao"
    kw, n, L, S, K, every, seed, names = CASES[case]
    w = np.random.default_rng(8).standard_normal((n, 3))
    sysg, _, _ = make(case)
    with sysg, ps.Adjoint(sysg, K, every) as adj:
        res = []
        for on in (False, True, False):
            adj.set_record_all(on)
            assert adj.forward() == 0
            xk = sysg.get_state()["pos"]
            res.append((adj.backward(w, with_gx0=True), xk))
    (g0, x0), (g1, x1), (g2, x2) = res
    assert np.array_equal(x0, x1) and np.array_equal(x0, x2)
    for name in ps.engine.ADJOINT_PARAMS:
        assert g1[name] == g0[name] and g2[name] == g0[name], name
    assert np.array_equal(g1["gx0"], g0["gx0"]) and np.array_equal(g2["gx0"], g0["gx0"])


def test_checkpoint_interval_does_not_change_the_gradient():
    case = "ao"
    kw, n, L, S, K, every, seed, names = CASES[case]
    w = np.random.default_rng(7).standard_normal((n, 3))
    res = []
    for ev in (1, 7, K):
        sysg, _, _ = make(case)
        with sysg, ps.Adjoint(sysg, K, ev) as adj:
            assert adj.forward() == 0
            res.append(adj.backward(w))
    assert res[0] == res[1] == res[2]


@pytest.mark.parametrize("case", sorted(CASES))
def test_adjoint_matches_gpu_central_differences(case):
    kw, n, L, S, K, every, seed, names = CASES[case]
    print(f"adjoint FD case={case}", flush=True)
    w = np.random.default_rng(11).standard_normal((n, 3))
    sysg, _, _ = make(case)
    with sysg, ps.Adjoint(sysg, K, every) as adj:
        assert adj.forward() == 0
        g = adj.backward(w)
        for name in names:
            key = {"ao_well_depth": ("ao", "well_depth"), "cutoff": ("ao", "cutoff"),
                   "lj_well_depth": ("lj", "well_depth"),
                   "yukawa_strength": ("lj", "yukawa_strength"),
                   "debye_length": ("lj", "debye_length")}[name]
            base = kw[key[0]][key[1]]
            # the cutoff step moves pairs across the cutoff unless it is tiny
            # (about 700 pairs per unit distance sit at r = cutoff each step)
            h = (1e-6 if key[1] == "cutoff" else 1e-4 if key[0] == "ao" else 1e-3) * base
            xs = []
            for sgn in (1, -1):
                k2 = {m: dict(v) for m, v in kw.items()}
                k2[key[0]][key[1]] = base + sgn * h
                if key[1] == "cutoff":
                    for m in k2:
                        k2[m]["cutoff"] = base + sgn * h
                sysg.set_forces(modules(k2))
                assert adj.forward() == 0
                xs.append(sysg.get_state()["pos"])
            d = xs[0] - xs[1]
            d -= L * np.round(d / L)
            fd = float(np.sum(w * d)) / (2 * h)
            assert g[name] == pytest.approx(fd, rel=1e-4), (name, g[name], fd)
        sysg.set_forces(modules(kw))


def test_backward_requires_the_forward_parameters():
    sysg, _, _ = make("ao")
    with sysg, ps.Adjoint(sysg, 5, 2) as adj:
        assert adj.forward() == 0
        sysg.set_forces(ps.AOPotential(kT=1.0, well_depth=0.5, cutoff=1.12))
        with pytest.raises(ps.PsimError):
            adj.backward(np.zeros((sysg.n, 3)))
