"""Checks of the oracle that do not depend on the reference's code: the
mathematics fixes them (energy derivatives, invariants, brute force, series,
known-answer vectors, Langevin closed forms)."""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from cases import FORCE_CASES, PHI_P_C2, YUK_A_C3, force_params
from helpers import rsa_positions
from oracle.oracle import PO_AO, PO_CAL, PO_LJ, ForceParams


# ---- pair potentials written from the physics (energies, not forces) -------
def u_lj_yukawa(r, kT, eps, n, A, lam, s=1.0):
    # U = kT eps [4((s/r)^2n - (s/r)^n) + A lam exp(-r/lam)/r]
    return kT * eps * (4.0 * ((s / r) ** (2 * n) - (s / r) ** n) + A * lam * np.exp(-r / lam) / r)


def u_cal(r):
    return r ** -36.0


def u_ao(r, kT, phi_p, q, sigma=1.0):
    # Asakura-Oosawa depletion potential for r < sigma (1 + q)
    c = sigma * (1.0 + q)
    return -phi_p * kT * ((1.0 + q) / q) ** 3 * (1.0 - 1.5 * r / c + 0.5 * (r / c) ** 3)


def fd(fun, r, h=1e-6):
    return (fun(r + h) - fun(r - h)) / (2 * h)


def two_particle_fnet(oracle, fp, r, box=20.0, S=6, across=False, seed=0):
    """Place two particles at distance r along a random direction; return
    (fNet along the pair, force vectors).  fNet > 0 means attraction."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    a = np.array([0.3, box / 2, box / 2]) if across else np.array([box / 2] * 3)
    b = np.mod(a + r * u, box)
    if across:
        b = np.mod(a - r * np.abs(u) * np.array([1, 0, 0]) - r * u * np.array([0, 1, 1]), box)
        d = b - a
        d -= box * np.round(d / box)
        u = d / np.linalg.norm(d)
    pos = np.vstack([a, b])
    st, f, _ = oracle.forces(fp, pos, S, box / S, box, half_uv=box / 2)
    assert st == 0
    return float(f[0] @ u), f, u


@pytest.mark.parametrize("n,A,lam", [(6, 0.0, 0.5), (18, 8.0, 0.5), (18, YUK_A_C3, 0.2)])
@pytest.mark.parametrize("r", [0.95, 1.02, 1.3, 2.2])
def test_lj_yukawa_force_is_energy_derivative(oracle, n, A, lam, r):
    kT, eps = 1.3, 0.7
    fp = ForceParams.make(PO_LJ, 3.0, kT=kT, lj_well_depth=eps, yukawa_strength=A,
                          debye_length=lam, lj_num=n)
    fnet, f, u = two_particle_fnet(oracle, fp, r)
    expect = fd(lambda x: u_lj_yukawa(x, kT, eps, n, A, lam), r)
    assert fnet == pytest.approx(expect, rel=1e-6, abs=1e-9)
    # Newton's third law and central force
    assert np.allclose(f[0], -f[1], rtol=0, atol=1e-12 * max(1, abs(fnet)))
    assert np.linalg.norm(np.cross(f[0], u)) < 1e-9 * max(1.0, abs(fnet))


@pytest.mark.parametrize("r", [0.85, 0.95, 1.05, 1.6])
def test_calibration_force_is_energy_derivative(oracle, r):
    fp = ForceParams.make(PO_CAL, 2.0)
    fnet, f, u = two_particle_fnet(oracle, fp, r)
    assert fnet == pytest.approx(fd(u_cal, r, h=1e-7 * r), rel=1e-6)
    assert fnet < 0   # repulsive: the force on particle 0 points away from 1


@pytest.mark.parametrize("q,u_contact", [(0.1, 4.0), (0.12, 5.0), (0.05, 2.0)])
@pytest.mark.parametrize("x", [0.85, 0.97, 1.01, 1.04])
def test_ao_force_is_energy_derivative_over_r_plus_core(oracle, q, u_contact, x):
    """fNet_AO = (dU_AO/dr)/r - 36 r^-38 (AOPotential.cpp:95-102 as written)."""
    phi_p = 2 * q * u_contact / (3 + 2 * q)
    c = 1 + q
    r = min(x, c - 1e-3)
    fp = ForceParams.make(PO_AO, c, kT=1.0, ao_well_depth=phi_p)
    fnet, f, u = two_particle_fnet(oracle, fp, r)
    expect = fd(lambda y: u_ao(y, 1.0, phi_p, q), r, h=1e-7) / r - 36.0 * r ** -38
    assert fnet == pytest.approx(expect, rel=1e-6, abs=1e-9)
    # the contact value that fixes the mapping phi_p <-> well depth
    assert u_ao(1.0, 1.0, phi_p, q) == pytest.approx(-u_contact, rel=1e-12)


def test_pair_across_periodic_boundary(oracle):
    fp = ForceParams.make(PO_LJ, 2.5, kT=1.0, lj_well_depth=1.0, yukawa_strength=0.0, lj_num=6)
    a, _, _ = two_particle_fnet(oracle, fp, 1.1, across=False, seed=3)
    b, _, _ = two_particle_fnet(oracle, fp, 1.1, across=True, seed=3)
    assert b == pytest.approx(a, rel=1e-9)


def brute_force(pos, box, rc, fnet):
    """O(N^2) minimum-image forces with fNet from a numerical energy derivative."""
    d = pos[None, :, :] - pos[:, None, :]          # j - i
    d -= box * np.round(d / box)
    r = np.linalg.norm(d, axis=2)
    np.fill_diagonal(r, np.inf)
    m = r < rc
    f = np.zeros_like(r)
    f[m] = fnet(r[m]) / r[m]
    return np.einsum("ij,ijk->ik", f, d), m.sum(axis=1)


@pytest.mark.parametrize("case", ["lj6", "ljyuk", "ao", "cal"])
def test_cell_list_forces_equal_brute_force(oracle, case):
    plugin, mask, cfg, n, conc, scale = FORCE_CASES[case]
    box, cell, _ = oracle.geometry(n, conc, scale)
    pos = rsa_positions(min(n, 1500), box, seed=77)
    fp = force_params(mask, cfg)
    st, f, _ = oracle.forces(fp, pos, scale, cell, box)
    assert st == 0
    rc = cfg["cutOff"]
    if mask == PO_LJ:
        fn = lambda r: fd(lambda x: u_lj_yukawa(x, 1.0, cfg["wellDepth"], cfg["ljNum"],  # noqa: E731
                                                cfg["yukawaStrength"], cfg["debyeLength"]), r,
                          h=1e-6)
    elif mask == PO_AO:
        q = rc - 1
        fn = lambda r: fd(lambda x: u_ao(x, 1.0, cfg["wellDepth"], q), r, h=1e-7) / r \
            - 36.0 * r ** -38  # noqa: E731
    else:
        fn = lambda r: fd(u_cal, r, h=1e-7)  # noqa: E731
    fb, _ = brute_force(pos, box, rc, fn)
    scale_f = np.abs(fb).max()
    assert np.abs(f - fb).max() <= 2e-6 * scale_f


def test_total_force_zero_and_translation_invariance(oracle):
    plugin, mask, cfg, n, conc, scale = FORCE_CASES["ljyuk"]
    box, cell, _ = oracle.geometry(n, conc, scale)
    pos = rsa_positions(n, box, seed=5)
    fp = force_params(mask, cfg)
    st, f, _ = oracle.forces(fp, pos, scale, cell, box)
    assert st == 0
    fs = np.abs(f).sum()
    assert np.abs(f.sum(axis=0)).max() <= 1e-12 * fs
    shifted = np.mod(pos + np.array([0.37, 5.23, -2.1]), box)
    st, f2, _ = oracle.forces(fp, shifted, scale, cell, box)
    assert np.abs(f2 - f).max() <= 1e-11 * np.abs(f).max()
    perm = np.random.default_rng(1).permutation(n)
    st, f3, _ = oracle.forces(fp, pos[perm], scale, cell, box)
    assert np.abs(f3 - f[perm]).max() <= 1e-12 * np.abs(f).max()


def test_overlap_and_nan_errors(oracle):
    fp = ForceParams.make(PO_CAL, 2.0)
    pos = np.array([[5.0, 5.0, 5.0], [5.79, 5.0, 5.0]])   # r = 0.79 < 0.8 (r_i + r_j)
    st, f, err = oracle.forces(fp, pos, 5, 2.0, 10.0)
    assert st == 7704
    pos[1, 0] = 5.81
    st, f, err = oracle.forces(fp, pos, 5, 2.0, 10.0)
    assert st == 0
    fp = ForceParams.make(PO_AO, 1.1, ao_well_depth=float("nan"))
    pos[1, 0] = 6.0
    st, f, err = oracle.forces(fp, pos, 5, 2.0, 10.0)
    assert st == 7705


# ---- integrator ----------------------------------------------------------------
def decimal_coeffs(gamma, dt, kT=1.0):
    getcontext().prec = 60
    y = Decimal(gamma) * Decimal(dt)
    e = (-y).exp()
    aa1 = 1 - e
    aa2 = Decimal("0.5") * y * (1 + e) - aa1
    aa3 = y - aa1
    s1 = 2 * y - 3 + 4 * e - (-2 * y).exp()
    s2 = -2 * y - 3 + 4 * y.exp() - (2 * y).exp()
    gn = y.exp() - 2 * y - (-y).exp()
    g2 = Decimal(gamma) ** 2
    sig1 = (Decimal(kT) * s1 / g2).sqrt()
    sig2 = (-Decimal(kT) * s2 / g2).sqrt()
    return dict(coeff0=e, coeff1=aa1 / y, coeff2=aa2 / (y * y), coeff3=aa3 / (y * y),
                sig1=sig1, sig2=sig2, corr=(Decimal(kT) / g2) * (gn / (sig1 * sig2)),
                goy2=gn / (y * y), goy3=gn / (y * y * y), hn=y / (y.exp() - (-y).exp()))


@pytest.mark.parametrize("gamma,dt", [(1.0, 1e-3), (0.04, 1e-3), (0.5, 5e-4), (10.0, 1e-2)])
def test_gb_coefficients_against_high_precision(oracle, gamma, dt):
    c = oracle.gb_coeffs(1.0, 1.0, gamma, dt)
    ref = decimal_coeffs(gamma, dt)
    y = gamma * dt
    eps = np.finfo(float).eps
    assert c.coeff1 == pytest.approx(float(ref["coeff1"]), rel=1e-13)
    if gamma < 0.05:
        # series branch (brownianIntegrator.cpp:135-166): accurate where the
        # closed forms cancel
        for k in ("coeff2", "coeff3", "sig1", "sig2", "goy2", "goy3", "hn"):
            assert getattr(c, k) == pytest.approx(float(ref[k]), rel=1e-9), k
    else:
        # the closed forms (:111-133) cancel at small y; the reference's own
        # rounding is reproduced, so agreement is bounded by the conditioning
        assert c.coeff0 == pytest.approx(math.exp(-y), rel=1e-15)
        tol = max(1e-12, 50 * eps / y ** 3)
        for k in ("coeff2", "coeff3", "sig1", "sig2", "goy2", "goy3", "hn", "corr"):
            assert getattr(c, k) == pytest.approx(float(ref[k]), rel=tol), k


def test_free_particle_zero_noise_velocity_decay(oracle):
    """No force, no noise: x_{n+1} - x_n = exp(-gamma dt) (x_n - x_{n-1})."""
    gamma, dt, box = 2.0, 1e-3, 100.0
    c = oracle.gb_coeffs(1.0, 1.0, gamma, dt)
    pos = np.array([[50.0, 50.0, 50.0]])
    vel = np.array([[1.0, -2.0, 0.5]])
    z = np.zeros((1, 3))
    st, p1, p01, mx = oracle.gb_first_step(c, box, pos, np.zeros((1, 3)), vel, z, z)
    disp = [p1 - pos]
    prev, cur, cur0 = pos, p1, p01
    for _ in range(50):
        st, p, p0, v, mx, _ = oracle.gb_normal_step(c, box, cur, cur0, vel, z, z, mx, z, z, 0)
        disp.append(p - cur)
        prev, cur, cur0 = cur, p, p0
    ratios = np.array([d2 / d1 for d1, d2 in zip(disp[:-1], disp[1:])])
    assert np.allclose(ratios, math.exp(-gamma * dt), rtol=1e-9, atol=0)


def test_constant_force_terminal_drift(oracle):
    """Constant F, no noise: the drift velocity tends to F / (m gamma)."""
    gamma, dt, box, m = 1.5, 1e-2, 1000.0, 2.0
    c = oracle.gb_coeffs(1.0, m, gamma, dt)
    F = np.array([[0.3, -0.2, 0.1]])
    z = np.zeros((1, 3))
    cur = np.array([[500.0, 500.0, 500.0]])
    st, p, p0, mx = oracle.gb_first_step(c, box, cur, z, z, F, z)
    for _ in range(3000):
        st, pn, p0, v, mx, _ = oracle.gb_normal_step(c, box, p, p0, z, F, F, mx, z, z, 0)
        last = pn - p
        p = pn
    assert np.allclose(last / dt, F / (m * gamma), rtol=1e-6)


@pytest.mark.slow
def test_ideal_gas_msd_and_temperature(oracle):
    """Free Langevin particles: MSD(t) = 6 kT/(m gamma) [t - (1 - e^{-gamma t})/gamma]
    from equilibrium velocities, and <v^2>/3 = kT from the velocity steps."""
    n, box, S, gamma, dt, kT = 4096, 40.0, 10, 1.0, 1e-3, 1.0
    c = oracle.gb_coeffs(kT, 1.0, gamma, dt)
    fp = ForceParams.make(PO_CAL, 1e-6)
    rng = np.random.default_rng(11)
    pos = rng.random((n, 3)) * box
    vel = rng.standard_normal((n, 3)) * math.sqrt(kT)
    st0 = dict(pos=pos.copy(), pos0=np.zeros((n, 3)), vel=vel, frc=np.zeros((n, 3)),
               frc0=np.zeros((n, 3)), memx=np.zeros((n, 3)))
    unwrapped = np.zeros((n, 3))
    prev = pos.copy()
    chunk, nchunks = 100, 15
    temps = []
    for _ in range(nchunks):
        st, err = oracle.run(fp, c, st0, chunk, box, S, box / S, vel_freq=0, seed=4242)
        assert st == 0
        d = st0["pos"] - prev
        d -= box * np.round(d / box)
        unwrapped += d
        prev = st0["pos"].copy()
        temps.append(oracle.temperature(st0["vel"]))
    t = chunk * nchunks * dt
    msd = (unwrapped ** 2).sum(axis=1)
    expect = 6 * kT / gamma * (t - (1 - math.exp(-gamma * t)) / gamma)
    se = msd.std() / math.sqrt(n)
    assert abs(msd.mean() - expect) < 3 * se
    se_t = math.sqrt(2.0 / (3 * n)) * kT
    assert abs(np.mean(temps[-5:]) - kT) < 3 * se_t * 2


def test_philox_known_answers(oracle):
    # Random123 kat_vectors for philox4x32_10
    assert oracle.philox_block([0, 0, 0, 0], [0, 0]) == \
        [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle.philox_block([0xffffffff] * 4, [0xffffffff] * 2) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle.philox_block([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                               [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_philox_normals_moments(oracle):
    z = oracle.philox_normals(90210, 7, 200000)
    assert abs(z.mean()) < 5 / math.sqrt(z.size)
    assert abs(z.var() - 1) < 5 * math.sqrt(2 / z.size)
    # distinct steps and ids give independent draws
    z2 = oracle.philox_normals(90210, 8, 200000)
    assert abs(np.corrcoef(z[0], z2[0])[0, 1]) < 0.02


def test_geometry_reference_settings(oracle):
    # ClusteredUI/settings.cfg: nParticles 2500, conc 0.05, scale 5
    box, cell, phi = oracle.geometry(2500, 0.05, 5)
    vP = 2500 * 4.0 / 3.0 * math.pi * 0.125
    assert (box, cell) == (25, 5) and int(np.cbrt(vP / 0.05)) == 29
    assert phi == pytest.approx(vP / 25 ** 3, rel=1e-15)


def test_wrap_and_image_of_previous_position(oracle):
    """particle::setX: single wrap; pos0 re-imaged (particle.cpp:61-70)."""
    c = oracle.gb_coeffs(1.0, 1.0, 1.0, 1e-3)
    box = 10.0
    pos = np.array([[9.9995, 0.0002, 5.0]])
    vel = np.array([[1.0, -1.0, 0.0]])
    z = np.zeros((1, 3))
    st, p, p0, mx = oracle.gb_first_step(c, box, pos, z, vel, z, z)
    assert st == 0
    assert 0 <= p[0, 0] < 0.001 and p0[0, 0] == pos[0, 0] - box
    assert box - 0.001 < p[0, 1] < box and p0[0, 1] == pos[0, 1] + box
    assert p0[0, 2] == pos[0, 2]
    vel = np.array([[3e4, 0.0, 0.0]])   # a jump beyond one box: still out after one wrap
    st, p, p0, mx = oracle.gb_first_step(c, box, pos, z, vel, z, z)
    assert st == 7703
