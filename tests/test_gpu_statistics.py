"""10^4-step statistics: the GPU engine against the reference's own compiled
CPU path (needs a GPU and oracle/_ref).

Both runs start from the same seeded RSA configuration and Maxwell
velocities; the reference draws its noise from fastRandom, the GPU from
Philox, so the trajectories differ and only their statistics can agree.
Workload: AO depletion (q = 0.1, 4 kT contact, phi_p = 0.25, built-in core),
N = 2292 in the reference's integer geometry (box 20, cell 2, scale 10,
phi = 0.150), dt = 5e-4, gamma = 1, kT = 1, velocities every step,
10^4 steps, sampled every 100 steps.

Checks (two standard errors of the difference, SE combined over both runs):
* kinetic temperature averaged over steps 2000..10^4 -- SE from the spread
  of per-particle v^2/3 in one frame over sqrt(N), an upper bound on the SE
  of a time average of positively correlated frames;
* mean-square displacement at 5000 and 10^4 steps -- SE from the spread of
  the per-particle squared displacement over sqrt(N);
* g(r), bins of 0.05 on [0.8, 3.0), averaged over steps 2000..10^4 -- SE per
  bin from 8 blocks of 1000 steps (several structural relaxation times
  each); at least 85 % of the bins within 2 SE and chi^2 per bin <= 2.
The GPU's own displacement tracker is also checked against the unwrapped
sampled trajectory.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import proteinsim2015_amd as ps
from helpers import maxwell_velocities, rsa_positions

pytestmark = pytest.mark.gpu

N, BOX, CELL, SCALE = 2292, 20, 2, 10
DT, NSTEPS, EVERY = 5e-4, 10000, 100
PHI_P, CUT = 2 * 0.1 * 4.0 / (3 + 2 * 0.1), 1.1
SEED = 90210
EDGES = np.arange(0.8, 3.0 + 1e-9, 0.05)


def unwrap(frames):
    d = np.diff(frames, axis=0)
    d -= BOX * np.round(d / BOX)
    return np.concatenate([frames[:1], frames[0] + np.cumsum(d, axis=0)])


def gofr(x):
    t = cKDTree(x, boxsize=BOX)
    pairs = t.query_pairs(EDGES[-1], output_type="ndarray")
    d = x[pairs[:, 0]] - x[pairs[:, 1]]
    d -= BOX * np.round(d / BOX)
    r = np.linalg.norm(d, axis=1)
    h, _ = np.histogram(r, EDGES)
    shell = 4 / 3 * np.pi * (EDGES[1:] ** 3 - EDGES[:-1] ** 3)
    return 2 * h / (N * (N / BOX ** 3) * shell)


def stats(frames, vel):
    first = 2000 // EVERY
    g = np.array([gofr(np.mod(x, BOX)) for x in frames[first:]])
    nb = 8
    blocks = g[: (len(g) // nb) * nb].reshape(nb, -1, g.shape[1]).mean(axis=1)
    u = unwrap(frames)
    sq = np.sum((u - u[0]) ** 2, axis=2)          # [frame, particle]
    v2 = np.sum(vel ** 2, axis=2) / 3.0            # [frame, particle]
    return dict(g=blocks.mean(axis=0), g_se=blocks.std(axis=0, ddof=1) / np.sqrt(nb),
                T=v2[first:].mean(), T_se=v2[first:].std(axis=1).mean() / np.sqrt(N),
                msd=sq.mean(axis=1), msd_se=sq.std(axis=1) / np.sqrt(N), unwrapped=u)


def test_gofr_msd_temperature_match_reference_statistically(reference):
    pos = rsa_positions(N, BOX, 1234)
    vel = maxwell_velocities(N, 1.0, 4321)
    cfg = {"cutOff": CUT, "kT": 1.0, "wellDepth": PHI_P, "gamma": 1.0, "mass": 1.0,
           "timeStep": DT, "velFreq": 0, "seed": SEED, "nParticles": N}
    print("statistics: reference run", flush=True)
    ref = reference.run_sampled("AO", cfg, pos, vel, 0.5, 1.0, BOX, CELL, SCALE, DT, NSTEPS,
                                EVERY, nthreads=16)
    print("statistics: GPU run", flush=True)
    geom = ps.SystemGeometry(n=N, box=float(BOX), cells_per_dim=SCALE, cell_size=float(CELL))
    integ = ps.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=DT, vel_freq=0, seed=SEED)
    frames, vels = [], []
    with ps.GpuSystem(geom, ps.AOPotential(kT=1.0, well_depth=PHI_P, cutoff=CUT),
                      integ) as sysg:
        sysg.set_state(pos, vel=vel)
        sysg.track_start()
        for k in range(NSTEPS // EVERY + 1):
            if k:
                assert sysg.step(EVERY) == 0
            st = sysg.get_state()
            frames.append(st["pos"])
            vels.append(st["vel"])
        ob = sysg.observables(0.0)
    a = stats(ref["pos"], ref["vel"])
    b = stats(np.array(frames), np.array(vels))
    # the tracker's crossing counters reproduce the unwrapped trajectory
    assert ob["msd"] == pytest.approx(b["msd"][-1], rel=1e-9)
    # temperature
    se = np.hypot(a["T_se"], b["T_se"])
    assert abs(a["T"] - b["T"]) <= 2 * se, (a["T"], b["T"], se)
    # mean-square displacement at 5000 and 10000 steps
    for k in (5000 // EVERY, NSTEPS // EVERY):
        se = np.hypot(a["msd_se"][k], b["msd_se"][k])
        assert abs(a["msd"][k] - b["msd"][k]) <= 2 * se, (k, a["msd"][k], b["msd"][k], se)
    # g(r)
    se = np.hypot(a["g_se"], b["g_se"])
    ok = se > 0
    z = np.abs(a["g"] - b["g"])[ok] / se[ok]
    print(f"g(r): {np.mean(z <= 2):.2f} of bins within 2 SE, chi2/bin {np.mean(z ** 2):.2f}, "
          f"peak {a['g'].max():.3f} vs {b['g'].max():.3f}", flush=True)
    assert np.mean(z <= 2) >= 0.85
    assert np.mean(z ** 2) <= 2.0
