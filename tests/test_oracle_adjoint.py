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
