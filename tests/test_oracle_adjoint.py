"""Pins the oracle's trajectory adjoint (CPU only).

* pair-force derivatives: dfNet/dr and dfNet/dparameter against central
  differences of the oracle's own pair forces (which are bit-exact with the
  reference plugins), evaluated on two-particle systems;
* the adjoint: dL/dparameter and dL/dx_0 against central differences of
  the loss of forward po_run trajectories.  The loss is linear in x_K with
  random weights, and the finite difference is taken on the position
  difference x_K(+h) - x_K(-h) before weighting, so rounding of the large
  absolute positions does not swamp it.  The pair set inside the cutoff is
  the same in the perturbed runs at these sizes (the adjoint is pathwise).
"""
import numpy as np
import pytest

from cases import PHI_P_C2
from helpers import rsa_positions
from oracle.oracle import PO_AO, PO_CAL, PO_LJ, ForceParams


def _two_body_force(oracle, fp, r, L=20.0, S=6):
    """x-force on particle 0 from particle 1 at distance r along +x."""
    pos = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
    st, f, _ = oracle.forces(fp, pos, S, L / S, L, half_uv=L / 2)
    assert st == 0
    return f[0, 0]


PAIR_CASES = {
    "ao": (dict(mask=PO_AO, cutoff=1.12, kT=1.0, ao_well_depth=0.37),
           ("ao_well_depth", "cutoff"), (0, 1), [0.97, 1.03, 1.1]),
    "lj": (dict(mask=PO_LJ, cutoff=2.5, kT=1.3, lj_well_depth=2.0, yukawa_strength=30.0,
                debye_length=0.3, lj_num=6),
           ("lj_well_depth", "yukawa_strength", "debye_length"), (2, 3, 4), [0.9, 1.2, 2.2]),
    "cal": (dict(mask=PO_CAL, cutoff=2.0), (), (), [0.9, 1.3]),
}


def _fp(kw):
    kw = dict(kw)
    mask, cutoff = kw.pop("mask"), kw.pop("cutoff")
    return ForceParams.make(mask, cutoff, **kw)


@pytest.mark.parametrize("case", sorted(PAIR_CASES))
def test_pair_derivatives_central_differences(oracle, case):
    kw, names, idx, radii = PAIR_CASES[case]
    fp = _fp(kw)
    for r in radii:
        f, dfdr, dth = oracle.pair_fnet_d(fp, r)
        assert f == pytest.approx(_two_body_force(oracle, fp, r), rel=1e-12)
        h = 1e-6
        fd = (_two_body_force(oracle, fp, r + h) - _two_body_force(oracle, fp, r - h)) / (2 * h)
        assert dfdr == pytest.approx(fd, rel=1e-7)
        for name, k in zip(names, idx):
            hp = 1e-6 * kw[name]
            kp, km = dict(kw), dict(kw)
            kp[name] += hp
            km[name] -= hp
            fdp = (_two_body_force(oracle, _fp(kp), r)
                   - _two_body_force(oracle, _fp(km), r)) / (2 * hp)
            assert dth[k] == pytest.approx(fdp, rel=1e-7, abs=1e-12), (name, r)


def _fresh_pos0(pos, L):
    return np.where(np.abs(pos) > L / 2, np.where(pos < 0, -L, L), 0.0)


ADJ_CASES = {
    # mask, force kwargs, parameters checked (name, index), n, L, S, K, relative
    # FD step: small enough that no pair crosses the cutoff between the two
    # runs (a cutoff step of 1e-3 moves pairs across it here), large enough
    # that rounding differences of the stiff cores stay below the tolerance
    "ao": (dict(mask=PO_AO, cutoff=1.12, kT=1.0, ao_well_depth=0.4),
           [("ao_well_depth", 0), ("cutoff", 1)], 200, 12.0, 10, 15, 1e-4),
    "cal_lj_yukawa_s3": (dict(mask=PO_CAL | PO_LJ, cutoff=2.5, kT=1.0, lj_well_depth=1.0,
                              yukawa_strength=30.0, debye_length=0.3, lj_num=6),
                         [("lj_well_depth", 2), ("yukawa_strength", 3), ("debye_length", 4)],
                         150, 11.0, 3, 12, 1e-3),
}


@pytest.mark.parametrize("case", sorted(ADJ_CASES))
def test_adjoint_matches_central_differences(oracle, case):
    kw, params, n, L, S, K, hrel = ADJ_CASES[case]
    pos = rsa_positions(n, L, 7)
    rng = np.random.default_rng(3)
    vel = rng.standard_normal((n, 3))
    w = rng.standard_normal((n, 3))
    c = oracle.gb_coeffs(1.0, 1.0, 1.0, 1e-3)
    state = dict(pos=pos, pos0=_fresh_pos0(pos, L), vel=vel)

    def final(kwx, dpos=None):
        p = pos if dpos is None else pos + dpos
        st = dict(pos=p.copy(), pos0=_fresh_pos0(p, L), vel=vel.copy(), frc=np.zeros((n, 3)),
                  frc0=np.zeros((n, 3)), memx=np.zeros((n, 3)), time=0.0)
        s, _ = oracle.run(_fp(kwx), c, st, K, L, S, L / S, seed=11)
        assert s == 0
        return st["pos"]

    s, dth, gx0, xk = oracle.adjoint(_fp(kw), c, state, K, L, S, L / S, w, seed=11)
    assert s == 0
    assert np.array_equal(xk, final(kw))
    for name, k in params:
        h = hrel * kw[name]
        kp, km = dict(kw), dict(kw)
        kp[name] += h
        km[name] -= h
        d = final(kp) - final(km)
        d -= L * np.round(d / L)          # a wrap in one run only
        fd = float(np.sum(w * d)) / (2 * h)
        assert dth[k] == pytest.approx(fd, rel=1e-5), (name, dth[k], fd)
    # dL/dx_0 for a few coordinates
    for i, a in [(3, 0), (50, 2)]:
        e = np.zeros((n, 3))
        e[i, a] = 1e-6
        d = final(kw, e) - final(kw, -e)
        d -= L * np.round(d / L)
        fd = float(np.sum(w * d)) / 2e-6
        assert gx0[i, a] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def _pair_distances(p, L):
    d = p[:, None, :] - p[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(axis=2))
    iu = np.triu_indices(len(p), 1)
    return r[iu]


def test_cutoff_window_matches_central_differences(oracle):
    """The cutoff window's boundary term, against the secant over
    [c - w, c + w] when exactly one pair of the trajectory lies in the window
    (at the first step, at r = c - w/3) and every other pair distance of
    every step stays 1e-4 from c, so the kick of that one pair's force (a
    position change of order dt^2 F = 1e-6 per step) toggles nothing else.
    The secant holds that pair's force jump propagated through the later
    steps, with the pair's force taken at the coefficients of c + w:
    f(r; c + w) = f(r; c) + w df/dc, so
        secant = window estimate - P / 2   (to O(w^2)),
    P being the pair's own pathwise term (in the window estimate with weight
    1), measured as the pathwise derivative at c minus that at r - w/3,
    where the pair is outside.  (Over pairs spread evenly across the window
    this term averages out.)"""
    n, L, S, K, w = 100, 12.0, 10, 3, 1e-5
    kw = dict(mask=PO_AO, cutoff=1.1, kT=1.0, ao_well_depth=0.4)
    pos = rsa_positions(n, L, 21)
    rng = np.random.default_rng(5)
    vel = rng.standard_normal((n, 3))
    gw = rng.standard_normal((n, 3))
    c = oracle.gb_coeffs(1.0, 1.0, 1.0, 1e-3)

    def traj(kwx):
        st = dict(pos=pos.copy(), pos0=_fresh_pos0(pos, L), vel=vel.copy(),
                  frc=np.zeros((n, 3)), frc0=np.zeros((n, 3)), memx=np.zeros((n, 3)),
                  time=0.0)
        xs = [st["pos"].copy()]
        for _ in range(K):
            s, _ = oracle.run(_fp(kwx), c, st, 1, L, S, L / S, seed=17)
            assert s == 0
            xs.append(st["pos"].copy())
        return xs

    # c = r + w/3 for a first-step pair at r in [1.03, 1.17] (x_0 does not
    # depend on c) whose trajectory at that c keeps the margin everywhere else
    d0 = _pair_distances(pos, L)
    cut = None
    for k in np.argsort(np.abs(d0 - 1.1)):
        cand = float(d0[k] + w / 3)
        if abs(cand - 1.1) > 0.07:
            break
        if np.abs(np.delete(d0, k) - cand).min() <= 1e-4:
            continue
        kc = dict(kw, cutoff=cand)
        if all(np.abs(_pair_distances(x, L) - cand).min() > 1e-4 for x in traj(kc)[1:K]):
            cut = cand
            break
    assert cut is not None, "no isolated pair near the cutoff"
    kw["cutoff"] = cut
    state = dict(pos=pos, pos0=_fresh_pos0(pos, L), vel=vel)
    s, dth0, _, _ = oracle.adjoint(_fp(kw), c, state, K, L, S, L / S, gw, seed=17)
    s1, dthw, _, _ = oracle.adjoint(_fp(kw), c, state, K, L, S, L / S, gw, seed=17,
                                    cut_window=w)
    assert s == 0 and s1 == 0
    kp, km = dict(kw), dict(kw)
    kp["cutoff"] += w
    km["cutoff"] -= w
    d = traj(kp)[K] - traj(km)[K]
    d -= L * np.round(d / L)
    fd = float(np.sum(gw * d)) / (2 * w)
    # the jump dominates: the pathwise derivative alone misses it
    assert abs(fd - dth0[1]) > abs(dth0[1])
    sx, dthx, _, _ = oracle.adjoint(_fp(dict(kw, cutoff=cut - 2 * w / 3)), c, state, K, L, S,
                                    L / S, gw, seed=17)
    assert sx == 0
    est = dthw[1] - 0.5 * (dth0[1] - dthx[1])
    assert est == pytest.approx(fd, rel=1e-4), (dthw[1], est, fd, dth0[1])
    assert np.array_equal(dthw[[0, 2, 3, 4]], dth0[[0, 2, 3, 4]])
