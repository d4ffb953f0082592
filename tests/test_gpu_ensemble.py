"""Replica ensemble (needs a GPU): every replica follows the oracle, and the
graph replay and the autotuner's choice do not change a single bit."""
import numpy as np
import pytest

import proteinsim2015_amd as ps
from oracle.oracle import PO_AO, ForceParams

pytestmark = pytest.mark.gpu


def ao_replicas(seed, R, nmin, nmax):
    """Seeded replicas in the batched-ensemble parameter ranges: volume
    fraction 0.05-0.35, AO contact depth 2-8 kT, q 0.05-0.2."""
    rng = np.random.default_rng(seed)
    reps = []
    for _ in range(R):
        n = int(rng.integers(nmin, nmax + 1))
        phi = rng.uniform(0.05, 0.35)
        u = rng.uniform(2.0, 8.0)
        q = rng.uniform(0.05, 0.2)
        g = ps.SystemGeometry.exact(n, phi, 1.0 + q)
        ao = ps.AOPotential(kT=1.0, well_depth=ps.AOPotential.well_depth_for_contact(u, q),
                            cutoff=1.0 + q)
        reps.append(ps.Replica(g, ao))
    return reps


def initial_state(reps, seed):
    pos = np.concatenate([ps.rsa_positions(r.geometry.n, r.geometry.box, seed + k)
                          for k, r in enumerate(reps)])
    vel = np.random.default_rng(seed).standard_normal((len(pos), 3))
    return pos, vel


INTEG = ps.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=5e-4, vel_freq=3, seed=4242)


def test_ensemble_replicas_follow_oracle(oracle):
    reps = ao_replicas(11, 6, 150, 1500)
    pos, vel = initial_state(reps, 100)
    nsteps = 10
    with ps.Ensemble(reps, INTEG) as ens:
        ens.set_state(pos, vel)
        ens.set_graph(4)
        ens.step(nsteps)
        assert ens.last_error()[0] == 0
        out = ens.get_state()
        offsets = ens.offsets
    c = oracle.gb_coeffs(1.0, 1.0, 1.0, 5e-4)
    for r, rep in enumerate(reps):
        g = rep.geometry
        sl = slice(offsets[r], offsets[r + 1])
        p0 = np.where(np.abs(pos[sl]) > g.box / 2, np.where(pos[sl] < 0, -g.box, g.box), 0.0)
        st = dict(pos=pos[sl].copy(), pos0=p0, vel=vel[sl].copy(), frc=np.zeros((g.n, 3)),
                  frc0=np.zeros((g.n, 3)), memx=np.zeros((g.n, 3)))
        fp = ForceParams.make(PO_AO, rep.forces.cutoff, kT=1.0,
                              ao_well_depth=rep.forces.well_depth)
        code, _ = oracle.run(fp, c, st, nsteps, g.box, g.cells_per_dim, g.cell_size,
                             vel_freq=3, seed=4242, id0=int(offsets[r]))
        assert code == 0
        assert np.abs(out["pos"][sl] - st["pos"]).max() <= 1e-10, f"replica {r}"
        assert np.abs(out["vel"][sl] - st["vel"]).max() <= 1e-7, f"replica {r}"


def test_graph_replay_and_autotune_are_bit_identical():
    reps = ao_replicas(5, 12, 100, 800)
    pos, vel = initial_state(reps, 7)
    results = []
    for mode in ("eager", "graph", "tuned"):
        with ps.Ensemble(reps, INTEG) as ens:
            ens.set_state(pos, vel)
            if mode == "graph":
                ens.set_graph(5)
                ens.step(12)
            elif mode == "tuned":
                # the tuning's whole steps: one warm-up and one timed, for
                # one launch per class and for one launch in all (4)
                kernel, tile, ms = ens.autotune(1)
                allowed = [("sparse", 64), ("sparse", 128), ("sparse", 256),
                           ("brick", 96), ("brick", 192), ("brick", 384)]
                assert (kernel, tile) in allowed
                assert ms > 0
                choices = ens.class_choices()
                assert sum(c[0] for c in choices) == len(reps)
                assert all((k, t) in allowed for n, k, t in choices if n)
                assert ens.merged in (True, False)
                ens.step(8)
            else:
                ens.step(12)
            ens.sync()
            assert ens.last_error()[0] == 0
            results.append(ens.get_state())
    for k in ("pos", "pos0", "vel", "frc"):
        assert np.array_equal(results[0][k], results[1][k]), k
        assert np.array_equal(results[0][k], results[2][k]), k


def test_ensemble_reports_overlap():
    reps = ao_replicas(3, 3, 100, 200)
    pos, vel = initial_state(reps, 9)
    s1 = slice(reps[0].geometry.n, reps[0].geometry.n + reps[1].geometry.n)
    p1 = pos[s1].copy()
    p1[1] = p1[0] + np.array([0.5, 0.0, 0.0])     # r = 0.5 < 0.8 (r_i + r_j)
    p1[1] = np.mod(p1[1], reps[1].geometry.box)
    pos[s1] = p1
    with ps.Ensemble(reps, INTEG) as ens:
        ens.set_state(pos, vel)
        ens.step(1)
        code, i, j = ens.last_error()
    assert code == 7704
    assert {i, j} == {s1.start, s1.start + 1}


def test_full_512_replica_ensemble(oracle):
    """The batched workload at its own size: 512 seeded AO replicas (N 500-20000,
    phi 0.05-0.35, 2-8 kT, q 0.05-0.2; 5.25 M particles), autotuned and
    graph-replayed.  Every replica: positions inside its box, its net force
    ~ 0 and its neighbour sets equal to the oracle's after the run; eight
    seeded replicas (the smallest, the largest, the densest and five more)
    follow the oracle's trajectory."""
    from proteinsim2015_amd import bench_ensemble as be
    reps = be.make_replicas(512)
    pos_l, vel_l = be.initial_states(reps)
    pos, vel = np.concatenate(pos_l), np.concatenate(vel_l)
    integ = ps.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=5e-4, vel_freq=1000,
                                  seed=90210)
    print(f"ensemble 512: {len(pos)} particles", flush=True)
    with ps.Ensemble(reps, integ) as ens:
        ens.set_state(pos, vel)
        kernel, tile, _ = ens.autotune(1)      # whole steps: 2 x (one warm-up, one timed)
        ens.set_graph(4)
        ens.step(8)
        ens.sync()
        assert ens.last_error()[0] == 0
        out = ens.get_state()
        cnt, hs = ens.pair_stats()
        offsets = ens.offsets
    nsteps = 4 + 8
    print(f"ensemble 512: kernel {kernel} tile {tile}", flush=True)
    phis = np.array([r.geometry.volume_fraction for r in reps])
    sizes = np.array([r.geometry.n for r in reps])
    sample = {int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argmax(phis))}
    rng = np.random.default_rng(77)
    while len(sample) < 8:
        sample.add(int(rng.integers(len(reps))))
    c = oracle.gb_coeffs(1.0, 1.0, 1.0, 5e-4)
    for r, rep in enumerate(reps):
        g = rep.geometry
        sl = slice(offsets[r], offsets[r + 1])
        p = out["pos"][sl]
        assert p.min() >= 0.0 and p.max() < g.box, f"replica {r}"
        f = out["frc"][sl]
        assert np.abs(f.sum(axis=0)).max() <= 1e-9 * max(1.0, np.abs(f).sum()), f"replica {r}"
        fp = ForceParams.make(PO_AO, rep.forces.cutoff, kT=1.0,
                              ao_well_depth=rep.forces.well_depth)
        oc, oh = oracle.pair_stats(fp, p, g.cells_per_dim, g.cell_size, g.box,
                                   id0=int(offsets[r]))
        assert np.array_equal(cnt[sl], oc) and np.array_equal(hs[sl], oh), f"replica {r}"
        if r in sample:
            p0 = np.where(np.abs(pos[sl]) > g.box / 2,
                          np.where(pos[sl] < 0, -g.box, g.box), 0.0)
            st = dict(pos=pos[sl].copy(), pos0=p0, vel=vel[sl].copy(),
                      frc=np.zeros((g.n, 3)), frc0=np.zeros((g.n, 3)), memx=np.zeros((g.n, 3)))
            code, _ = oracle.run(fp, c, st, nsteps, g.box, g.cells_per_dim, g.cell_size,
                                 vel_freq=1000, seed=90210, id0=int(offsets[r]))
            assert code == 0
            assert np.abs(p - st["pos"]).max() <= 1e-10, f"replica {r}"
