"""The oracle's neighbour sets (po_pair_stats) against brute force.

The reference decides a pair inside iterCells (LJPotential.cpp:64-95) with
util::pbcDist (utilities.cpp:79-104) and a strict r^2 < cutOff^2 over the
27 cells of getAcceleration's stencil (LJPotential.cpp:139-177).  With
S >= 3 the stencil covers every cell within one cell edge, so for a cutoff
no larger than the cell edge the accepted set is exactly {j != i :
pbcDist(i, j) < cutOff^2} over all particles: an O(N^2) scan with the same
arithmetic pins the count and the partner-hash sum of every particle.
"""
import numpy as np
import pytest

from oracle.oracle import PO_AO, PO_CAL, PO_LJ, ForceParams

M64 = (1 << 64) - 1


def nbr_hash(o):
    """SplitMix64 finaliser of (o + 0x9E3779B97F4A7C15), Python integers."""
    z = (int(o) + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_nbr_hash_known_values():
    # SplitMix64 seeded with 0 returns these as its first outputs (Vigna's
    # splitmix64.c: the state advances by the golden gamma before mixing)
    assert nbr_hash(0) == 0xE220A8397B1DCDAF
    assert nbr_hash(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def brute_force_stats(pos, box, rc):
    """Per particle: partner count and hash sum with pbcDist's arithmetic."""
    n = len(pos)
    half = box / 2.0
    cnt = np.zeros(n, np.int64)
    hs = [0] * n
    h = [nbr_hash(j) for j in range(n)]
    for i in range(n):
        d = np.abs(pos[i] - pos)
        d = np.where(d > half, d - box, d)
        r2 = (d[:, 0] * d[:, 0]) + (d[:, 1] * d[:, 1]) + (d[:, 2] * d[:, 2])
        js = np.nonzero(r2 < rc * rc)[0]
        js = js[js != i]
        cnt[i] = len(js)
        hs[i] = sum(h[j] for j in js) & M64
    return cnt, np.array(hs, dtype=np.uint64)


@pytest.mark.parametrize("S,n,rc_frac,seed", [(3, 300, 1.0, 1), (5, 500, 0.9, 2),
                                              (7, 700, 1.0, 3), (4, 400, 0.6, 4)])
def test_pair_stats_match_brute_force(oracle, S, n, rc_frac, seed):
    rng = np.random.default_rng(seed)
    box = 12.0 if S != 3 else 7.5
    cell = box / S
    rc = rc_frac * cell
    pos = rng.random((n, 3)) * box
    fp = ForceParams.make(PO_CAL, rc)
    cnt, hs = oracle.pair_stats(fp, pos, S, cell, box)
    bc, bh = brute_force_stats(pos, box, rc)
    assert np.array_equal(cnt, bc)
    assert np.array_equal(hs, bh)
    # symmetric relation: every pair is counted from both sides
    assert cnt.sum() % 2 == 0


def test_pair_stats_strict_cutoff_and_wrap(oracle):
    """Pairs at exactly r = cutOff are excluded (strict <); pairs across the
    periodic boundary are included."""
    box, S = 9.0, 3
    rc = 1.5
    pos = np.array([[1.0, 1.0, 1.0],
                    [2.5, 1.0, 1.0],        # r = 1.5 exactly: excluded
                    [1.0, 2.25, 1.0],       # r = 1.25: included
                    [8.5, 1.0, 1.0],        # across x = 0: r = 1.5 exactly, excluded
                    [1.0, 1.0, 8.75]])      # across z = 0: r = 1.25, included
    fp = ForceParams.make(PO_CAL, rc)
    cnt, hs = oracle.pair_stats(fp, pos, S, box / S, box)
    bc, bh = brute_force_stats(pos, box, rc)
    assert np.array_equal(cnt, bc)
    assert np.array_equal(hs, bh)
    assert cnt[0] == 2 and cnt[1] == 0
    assert hs[0] == (nbr_hash(2) + nbr_hash(4)) & M64


@pytest.mark.parametrize("mask,kw", [(PO_LJ, dict(lj_well_depth=1.0, yukawa_strength=8.0,
                                                  debye_length=0.5, lj_num=18)),
                                     (PO_AO, dict(ao_well_depth=0.25)),
                                     (PO_CAL, {}),
                                     (PO_CAL | PO_LJ, dict(lj_well_depth=1.0,
                                                           yukawa_strength=0.0, lj_num=6))])
def test_pair_abs_sum_is_pair_force_magnitude(oracle, mask, kw):
    """Two particles: sum_j |fNet| equals |F| on either particle, the pair
    force itself (the oracle's forces are pinned against the reference)."""
    box, S = 10.0, 4
    rc = 1.1 if mask == PO_AO else 2.4
    fp = ForceParams.make(mask, rc, **kw)
    for r in (1.0, 1.05, 1.09):
        pos = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
        cnt, _, fs = oracle.pair_stats(fp, pos, S, box / S, box, abs_sum=True)
        st, f, _ = oracle.forces(fp, pos, S, box / S, box, half_uv=box / 2)
        assert st == 0
        assert list(cnt) == [1, 1]
        np.testing.assert_allclose(fs, np.linalg.norm(f, axis=1), rtol=1e-15)
