"""Neighbour sets of every pair kernel against the oracle, exactly (needs a GPU).

The reference's pair test is the strict r^2 < cutOff^2 of iterCells with
util::pbcDist over the 27-cell stencil (LJPotential.cpp:64-95, 139-177;
utilities.cpp:79-104).  The kernels cull with a guard-banded fast r^2 and
re-test survivors near the cutoff with pbcDist's arithmetic; each kernel's
statistics instantiation records every pair its acceptance admits.  Per
particle the partner count and the partner-hash sum must equal the oracle's
exactly, for the step kernels on every path (thread per particle, cell per
workgroup, brick), the trajectory adjoint's pair visitor, and at the full
sizes of the large workloads (N = 2,097,152 Yukawa + LJ, N = 262,144 AO),
where every particle's force is also held to 1e-6 relative.
"""
import os

import numpy as np
import pytest

import proteinsim2015_amd as ps
from cases import FORCE_CASES, PHI_P_C2, force_params
from helpers import force_rel_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FORCE_TOL = 1e-6
PATHS = ["sparse", "dense", "brick"]


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def system_for(plugin, cfg, n, box, cell, scale, path="auto"):
    geom = ps.SystemGeometry(n=n, box=float(box), cells_per_dim=int(scale),
                             cell_size=float(cell))
    integ = ps.BrownianIntegrator(kT=cfg.get("kT", 1.0), gamma=1.0, dt=1e-3, seed=1)
    mods = ps.PLUGINS[plugin].from_config(ps.Config(cfg))
    return ps.GpuSystem(geom, mods, integ, kernel_path=path)


def assert_same_sets(cnt, hs, ocnt, ohs, what):
    bad = np.nonzero((cnt != ocnt) | (hs != ohs))[0]
    assert len(bad) == 0, (f"{what}: {len(bad)} particles differ, first {bad[:5]}, counts "
                           f"{cnt[bad[:5]]} vs {ocnt[bad[:5]]}")


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("case", sorted(FORCE_CASES))
def test_neighbour_sets_golden_cases(oracle, case, path):
    plugin, mask, cfg, n, conc, scale = FORCE_CASES[case]
    g = load(f"forces_{case}.npz")
    pos, box, cell = g["pos"], int(g["box"]), int(g["cell"])
    print(f"case={case} n={n} path={path}", flush=True)
    fp = force_params(mask, cfg)
    ocnt, ohs = oracle.pair_stats(fp, pos, scale, cell, box)
    with system_for(plugin, cfg, n, box, cell, scale, path=path) as sysg:
        sysg.set_state(pos)
        cnt, hs = sysg.pair_stats()
        acnt, ahs = sysg.pair_stats(adjoint=True)
    assert ocnt.sum() > 0
    assert_same_sets(cnt, hs, ocnt, ohs, f"{case}/{path} step kernel")
    assert_same_sets(acnt, ahs, ocnt, ohs, f"{case} adjoint visitor")


@pytest.mark.parametrize("path", PATHS)
def test_neighbour_sets_at_the_cutoff(oracle, path):
    """Pairs placed at r = cutOff exactly and one rounding step either side,
    inside cells and across the periodic faces: the strict test decides."""
    box, S, rc = 24.0, 8, 2.5
    rng = np.random.default_rng(5)
    base = rng.random((400, 3)) * box
    pts = [base]
    for k in range(200):
        a = rng.random(3) * box
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        r = [rc, np.nextafter(rc, 0), np.nextafter(rc, 4)][k % 3]
        pts.append(np.stack([a, np.mod(a + r * u, box)]))
    # exact axis-aligned pairs across each face (r = rc exactly in FP64)
    for d in range(3):
        a = np.full(3, 7.0)
        a[d] = 1.0
        b = a.copy()
        b[d] = box - 1.5
        pts.append(np.stack([a, b]))
    pos = np.minimum(np.concatenate(pts), np.nextafter(box, 0))
    n = len(pos)
    cfg = {"cutOff": rc}
    fp = force_params(1, cfg)
    ocnt, ohs = oracle.pair_stats(fp, pos, S, box / S, box)
    with system_for("Cal", cfg, n, box, box / S, S, path=path) as sysg:
        sysg.set_state(pos)
        cnt, hs = sysg.pair_stats()
        acnt, ahs = sysg.pair_stats(adjoint=True)
    assert_same_sets(cnt, hs, ocnt, ohs, f"cutoff/{path}")
    assert_same_sets(acnt, ahs, ocnt, ohs, "cutoff adjoint")


def full_force_check(oracle, fp, f, pos, geom):
    """Every particle's force within FORCE_TOL of the oracle's, relative to
    max(|F_i|, sum_j |f_ij|) with the oracle's own pair sums."""
    sto, fo, _ = oracle.forces(fp, pos, geom.cells_per_dim, geom.cell_size, geom.box,
                               half_uv=geom.box / 2)
    assert sto == 0
    _, _, fs = oracle.pair_stats(fp, pos, geom.cells_per_dim, geom.cell_size, geom.box,
                                 abs_sum=True)
    err = force_rel_err(f, fo, fs)
    assert err.max() <= FORCE_TOL, f"max rel err {err.max():.3g} at {np.argmax(err)}"


@pytest.mark.parametrize("path", ["auto", "sparse"])
def test_full_size_yukawa_all_particles(oracle, path):
    """N = 2,097,152 Yukawa + LJ core at phi 0.30 (cutoff 3, 51^3 cells): after
    three steps, every particle's neighbour set equals the oracle's and every
    force is within 1e-6 relative."""
    n = 2097152
    geom = ps.SystemGeometry.exact(n, 0.30, 3.0)
    pos = ps.rsa_positions(n, geom.box, 1234)
    cfg = FORCE_CASES["ljyuk"][2]
    lj = ps.LennardJones.from_config(ps.Config(cfg))
    integ = ps.BrownianIntegrator(kT=1.0, gamma=1.0, dt=1e-3, seed=90210)
    print(f"yukawa 2M: path={path}", flush=True)
    with ps.GpuSystem(geom, lj, integ, kernel_path=path) as sysg:
        sysg.set_state(pos, vel=np.random.default_rng(3).standard_normal((n, 3)))
        assert sysg.step(3) == 0
        out = sysg.get_state()
        cnt, hs = sysg.pair_stats()
        st, f = sysg.forces_now()
    assert st == 0
    fp = force_params(2, cfg)
    ocnt, ohs = oracle.pair_stats(fp, out["pos"], geom.cells_per_dim, geom.cell_size, geom.box)
    assert_same_sets(cnt, hs, ocnt, ohs, f"yukawa 2M/{path}")
    print(f"yukawa 2M: in-range pairs per particle {ocnt.mean():.2f}", flush=True)
    full_force_check(oracle, fp, f, out["pos"], geom)


@pytest.mark.parametrize("path", ["auto", "brick"])
def test_full_size_ao_all_particles(oracle, path):
    """N = 262,144 AO (q = 0.1, 4 kT at contact, phi 0.15, cutoff 1.1): after
    five steps, every neighbour set equals the oracle's and every force is
    within 1e-6 relative."""
    n = 262144
    geom = ps.SystemGeometry.exact(n, 0.15, 1.1)
    pos = ps.rsa_positions(n, geom.box, 1234)
    ao = ps.AOPotential(kT=1.0, well_depth=PHI_P_C2, cutoff=1.1)
    integ = ps.BrownianIntegrator(kT=1.0, gamma=1.0, dt=5e-4, seed=90210)
    print(f"ao 262144: path={path}", flush=True)
    with ps.GpuSystem(geom, ao, integ, kernel_path=path) as sysg:
        sysg.set_state(pos, vel=np.random.default_rng(4).standard_normal((n, 3)))
        assert sysg.step(5) == 0
        out = sysg.get_state()
        cnt, hs = sysg.pair_stats()
        acnt, ahs = sysg.pair_stats(adjoint=True)
        st, f = sysg.forces_now()
    assert st == 0
    fp = force_params(4, {"cutOff": 1.1, "kT": 1.0, "wellDepth": PHI_P_C2})
    ocnt, ohs = oracle.pair_stats(fp, out["pos"], geom.cells_per_dim, geom.cell_size, geom.box)
    assert ocnt.sum() > 0
    assert_same_sets(cnt, hs, ocnt, ohs, f"ao 262144/{path}")
    assert_same_sets(acnt, ahs, ocnt, ohs, "ao 262144 adjoint")
    full_force_check(oracle, fp, f, out["pos"], geom)
