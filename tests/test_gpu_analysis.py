"""On-device analysis vs the oracle (needs a GPU).

Contacts, coordination and cluster labels are integer results: bit-exact
against the oracle's restatement of the reference's contact scan on the
same positions.  Temperature, mean displacement and the tracked MSD are
reductions: within 1e-12 relative (summation order).  The pair histogram
counts are exact; the smoothed histogram and its gradient are within 1e-10
of the brute-force oracle."""
import numpy as np
import pytest

import proteinsim2015_amd as ps
from cases import PHI_P_C2
from helpers import rsa_positions
from oracle.oracle import PO_AO, ForceParams

pytestmark = pytest.mark.gpu


def ao_system(n, box, S, seed=5, path="auto", vel_scale=1.0):
    geom = ps.SystemGeometry(n=n, box=box, cells_per_dim=S, cell_size=box / S, integer=False)
    integ = ps.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=1e-3, vel_freq=3, seed=seed)
    ao = ps.AOPotential(kT=1.0, well_depth=PHI_P_C2, cutoff=1.1)
    sysg = ps.GpuSystem(geom, ao, integ, kernel_path=path)
    pos = rsa_positions(n, box, seed)
    vel = np.random.default_rng(seed).standard_normal((n, 3)) * vel_scale
    sysg.set_state(pos, vel=vel)
    return sysg, pos, vel


@pytest.mark.parametrize("path", ["sparse", "dense"])
def test_observables_and_tracker_match_oracle(oracle, path):
    n, L, S = 2000, 20.0, 16
    print(f"observables path={path} n={n}", flush=True)
    sysg, pos, vel = ao_system(n, L, S, path=path, vel_scale=10.0)
    with sysg:
        sysg.track_start()
        nsteps = 60
        assert sysg.step(nsteps) == 0
        ob = sysg.observables(1.2, coordination=True)
        out = sysg.get_state()
    # the same trajectory on the oracle, with the crossing counters
    c = oracle.gb_coeffs(1.0, 1.0, 1.0, 1e-3)
    fp = ForceParams.make(PO_AO, 1.1, kT=1.0, ao_well_depth=PHI_P_C2)
    st = dict(pos=pos.copy(), pos0=np.zeros((n, 3)), vel=vel.copy(), frc=np.zeros((n, 3)),
              frc0=np.zeros((n, 3)), memx=np.zeros((n, 3)), time=0.0)
    pbc = np.zeros((n, 3), np.int32)
    for _ in range(nsteps):
        s, _ = oracle.run(fp, c, st, 1, L, S, L / S, vel_freq=3, seed=5)
        assert s == 0
        oracle.track_update(L, st["pos0"], pbc)
    assert np.abs(out["pos"] - st["pos"]).max() < 1e-10
    assert np.abs(pbc).sum() > 0
    msd = np.mean(np.sum((st["pos"] + L * pbc - pos) ** 2, axis=1))
    assert ob["msd"] == pytest.approx(msd, rel=1e-9)
    assert ob["temperature"] == pytest.approx(oracle.temperature(out["vel"]), rel=1e-12)
    md = np.mean(np.sum((out["pos"] - out["pos0"]) ** 2, axis=1))
    assert ob["mean_disp2"] == pytest.approx(md, rel=1e-12)
    coord = oracle.contacts(out["pos"], S, L / S, L, rc=1.2)
    assert np.array_equal(ob["coordination"], coord)
    assert ob["contacts"] == coord.sum()
    assert ob["step"] == nsteps


def test_clusters_match_oracle(oracle):
    n, L, S = 3000, 17.0, 14
    print(f"clusters n={n}", flush=True)
    sysg, pos, vel = ao_system(n, L, S)
    with sysg:
        assert sysg.step(20) == 0
        lab, ncl, mean = sysg.clusters(1.2, 5)
        out = sysg.get_state()
    _, olab = oracle.contacts(out["pos"], S, L / S, L, rc=1.2, labels=True)
    assert np.array_equal(lab, olab)
    sizes = np.bincount(olab, minlength=n)
    big = sizes[sizes >= 5]
    assert ncl == len(big)
    assert mean == (int(big.sum()) // len(big) if len(big) else 0)


@pytest.mark.parametrize("L,S", [(16.0, 12), (6.5, 5)])   # histogram grid / all pairs
def test_pair_histogram_matches_oracle(oracle, L, S):
    n = 600 if L > 10 else 80
    print(f"histogram L={L} n={n}", flush=True)
    sysg, pos, vel = ao_system(n, L, S)
    r_lo, w, nb = 0.8, 0.02, 110 if L > 10 else 65   # 65 bins: reach > L/3, all pairs
    with sysg:
        assert sysg.step(5) == 0
        x = sysg.get_state()["pos"]
        counts = sysg.pair_histogram(r_lo, w, nb, 0.0)
        smooth = sysg.pair_histogram(r_lo, w, nb, 0.02)
        rng = np.random.default_rng(1)
        wts = rng.standard_normal(nb)
        grad = sysg.pair_histogram_grad(r_lo, w, nb, 0.02, wts)
    assert np.array_equal(counts, oracle.pair_histogram_smooth(x, L, r_lo, w, nb, 0.0))
    ref = oracle.pair_histogram_smooth(x, L, r_lo, w, nb, 0.02)
    assert np.abs(smooth - ref).max() <= 1e-10 * max(1.0, ref.max())
    gref = oracle.pair_histogram_grad(x, L, r_lo, w, nb, 0.02, wts)
    assert np.abs(grad - gref).max() <= 1e-10 * max(1.0, np.abs(gref).max())


@pytest.mark.parametrize("w,sigma,nb", [(0.1, 0.02, 30), (0.02, 0.05, 100), (0.05, 0.01, 40)])
def test_pair_histogram_bin_to_sigma_ratios(oracle, w, sigma, nb):
    """Bin widths of 5, 0.4 and 5 sigma: the bin masses then take several
    Gauss-Legendre sub-intervals per bin, or one shorter than sigma."""
    print(f"histogram w={w} sigma={sigma}", flush=True)
    L = 16.0
    sysg, pos, vel = ao_system(600, L, 12)
    with sysg:
        assert sysg.step(5) == 0
        x = sysg.get_state()["pos"]
        smooth = sysg.pair_histogram(0.8, w, nb, sigma)
    ref = oracle.pair_histogram_smooth(x, L, 0.8, w, nb, sigma)
    assert np.abs(smooth - ref).max() <= 1e-10 * max(1.0, ref.max())
    assert abs(smooth.sum() - ref.sum()) <= 1e-10 * ref.sum()


def test_histogram_rejects_reach_beyond_half_box():
    sysg, _, _ = ao_system(80, 6.5, 5)
    with sysg:
        with pytest.raises(ps.PsimError):
            sysg.pair_histogram(0.8, 0.1, 30, 0.0)   # reaches 3.8 > 3.25
