"""Pins the oracle's analysis restatements (CPU only).

* crossing counters: the case table of analysisTracker.cpp:38-56 and the
  unwrapped trajectory rebuilt from per-step minimum-image displacements;
* contacts / clusters (systemHandling.cpp:94-161, analysisCluster.cpp:26-83):
  brute-force periodic neighbour counts and scipy's connected components;
* smoothed pair histogram: sigma -> 0 is the plain histogram of every
  ordered pair, the total mass is the number of pairs in range, and its
  gradient equals central differences of the histogram.
"""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components
from scipy.spatial import cKDTree

from cases import PHI_P_C2
from helpers import rsa_positions
from oracle.oracle import PO_AO, ForceParams


def test_track_update_case_table(oracle):
    L = 10.0
    # pos0 re-imaged below 0 (crossed the upper face going up), above L
    # (crossed the lower face going down), inside (no crossing), exactly L
    pos0 = np.array([[-0.1, 10.2, 5.0], [10.0, 0.0, -1e-9]])
    pbc = np.zeros((2, 3), np.int32)
    oracle.track_update(L, pos0, pbc)
    assert pbc.tolist() == [[1, -1, 0], [0, 0, 1]]


def test_tracked_displacement_unwraps_trajectory(oracle):
    """x + L * crossings equals the trajectory unwrapped from its per-step
    minimum-image increments."""
    n, L, S = 100, 12.0, 10
    pos = rsa_positions(n, L, 21)
    rng = np.random.default_rng(2)
    c = oracle.gb_coeffs(1.0, 1.0, 1.0, 1e-3)
    fp = ForceParams.make(PO_AO, 1.1, kT=1.0, ao_well_depth=PHI_P_C2)
    st = dict(pos=pos.copy(), pos0=np.zeros((n, 3)), vel=rng.standard_normal((n, 3)) * 10,
              frc=np.zeros((n, 3)), frc0=np.zeros((n, 3)), memx=np.zeros((n, 3)), time=0.0)
    start = pos.copy()
    unwrapped = pos.copy()
    pbc = np.zeros((n, 3), np.int32)
    prev = pos.copy()
    for _ in range(100):
        s, _ = oracle.run(fp, c, st, 1, L, S, L / S, seed=5)
        assert s == 0
        oracle.track_update(L, st["pos0"], pbc)
        d = st["pos"] - prev
        d -= L * np.round(d / L)
        unwrapped += d
        prev = st["pos"].copy()
    assert np.abs(pbc).sum() > 0, "no particle crossed a face; the test would pin nothing"
    np.testing.assert_allclose(st["pos"] + L * pbc, unwrapped, atol=1e-9)
    msd = np.mean(np.sum((st["pos"] + L * pbc - start) ** 2, axis=1))
    assert msd > 0


def _brute_contacts(pos, L, rc):
    t = cKDTree(np.mod(pos, L), boxsize=L)
    pairs = t.query_pairs(rc, output_type="ndarray")
    coord = np.zeros(len(pos), np.int64)
    np.add.at(coord, pairs[:, 0], 1)
    np.add.at(coord, pairs[:, 1], 1)
    return coord, pairs


@pytest.mark.parametrize("S", [3, 8])
def test_contacts_match_brute_force(oracle, S):
    n, L = 700, 14.0
    pos = rsa_positions(n, L, 4, dmin=0.95)
    coord, lab = oracle.contacts(pos, S, L / S, L, rc=1.2, labels=True)
    ref, pairs = _brute_contacts(pos, L, 1.2)
    assert np.array_equal(coord, ref)
    g = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n))
    ncomp, comp = connected_components(g, directed=False)
    # canonical labels: smallest member of each component
    smallest = np.full(ncomp, n)
    np.minimum.at(smallest, comp, np.arange(n))
    assert np.array_equal(lab, smallest[comp])


def test_smoothed_histogram_limits(oracle):
    n, L = 250, 11.0
    pos = rsa_positions(n, L, 8)
    r_lo, w, nb = 0.8, 0.05, 60
    plain = oracle.pair_histogram(pos, L, r_lo, w, nb)
    counts = oracle.pair_histogram_smooth(pos, L, r_lo, w, nb, 0.0)
    assert np.array_equal(counts, 2 * plain)     # ordered pairs
    tiny = oracle.pair_histogram_smooth(pos, L, r_lo, w, nb, 1e-9)
    assert np.abs(tiny - counts).max() < 1e-6
    # total Gaussian mass = pairs whose mass lies inside the range
    sig = 0.02
    sm = oracle.pair_histogram_smooth(pos, L, r_lo, w, nb, sig)
    t = cKDTree(pos, boxsize=L)
    d = t.sparse_distance_matrix(t, r_lo + w * nb + 1.0).toarray()
    r = d[(d > 0)]
    from scipy.special import ndtr
    expect = np.sum(ndtr((r_lo + w * nb - r) / sig) - ndtr((r_lo - r) / sig))
    assert abs(sm.sum() - expect) < 1e-9 * expect


def test_histogram_gradient_central_differences(oracle):
    n, L = 60, 7.0
    pos = rsa_positions(n, L, 12)
    r_lo, w, nb, sig = 0.8, 0.1, 20, 0.05
    rng = np.random.default_rng(0)
    wts = rng.standard_normal(nb)
    g = oracle.pair_histogram_grad(pos, L, r_lo, w, nb, sig, wts)
    h = 1e-6
    for i, k in [(0, 0), (7, 1), (33, 2), (59, 0)]:
        p1, p2 = pos.copy(), pos.copy()
        p1[i, k] += h
        p2[i, k] -= h
        fd = (wts @ oracle.pair_histogram_smooth(p1, L, r_lo, w, nb, sig)
              - wts @ oracle.pair_histogram_smooth(p2, L, r_lo, w, nb, sig)) / (2 * h)
        assert abs(fd - g[i, k]) <= 1e-6 * max(1.0, abs(g[i, k])), (i, k, fd, g[i, k])
