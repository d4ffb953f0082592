"""Brick-per-workgroup pair kernel (needs a GPU).

The brick kernel runs the thread-per-particle gather's arithmetic against
LDS-staged operands, in the same order, so with the same lanes per particle
it must give the same bits as the thread-per-particle kernel: positions,
forces, memory terms and velocities are compared with np.array_equal.  A
halo too populous for one LDS stage is processed in chunks, which reorders
the partial sums; that case is held to 1e-12."""
import numpy as np
import pytest

import proteinsim2015_amd as ps

pytestmark = pytest.mark.gpu

INTEG = ps.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=5e-4, vel_freq=3, seed=777)


def ao(q=0.1, u=4.0):
    return ps.AOPotential(kT=1.0, well_depth=ps.AOPotential.well_depth_for_contact(u, q),
                          cutoff=1.0 + q)


def lj_yukawa(cutoff=3.0):
    return ps.LennardJones(kT=1.0, well_depth=1.0, yukawa_strength=14841.3, lj_num=18,
                           cutoff=cutoff, debye_length=0.2)


def run(geom, forces, path, pos, vel, steps=6):
    print(f"path={path} n={geom.n} S={geom.cells_per_dim}", flush=True)
    with ps.GpuSystem(geom, forces, INTEG, kernel_path=path) as s:
        info = s.kernel_info()
        print(f"  kernel {info}", flush=True)
        assert info["path"] == path
        s.set_state(pos, vel=vel)
        st0, f0 = s.forces_now()
        s.step_device_noise(steps, 0)
        s.sync()
        code = s.last_error()[0]
        out = s.get_state()
        st1, f1 = s.forces_now()
    assert code == 0 and st0 == 0 and st1 == 0
    out["f0"], out["f1"] = f0, f1
    return out


def same_bits(a, b):
    for k in ("pos", "pos0", "vel", "frc", "f0", "f1"):
        assert np.array_equal(a[k], b[k]), k


def state(n, box, seed):
    pos = ps.rsa_positions(n, box, seed)
    vel = np.random.default_rng(seed).standard_normal((n, 3))
    return pos, vel


CASES = {
    # name: (n, phi, cutoff, forces)
    "ao_small": (3000, 0.15, 1.1, lambda: ao()),
    "ao_dense_small": (2500, 0.33, 1.05, lambda: ao(0.05, 6.0)),
    "ljyuk_small": (4000, 0.2, 3.0, lambda: lj_yukawa()),
    "cal_lj": (3000, 0.2, 2.5, lambda: ps.ForceSet([ps.Calibration(cutoff=2.5), ps.LennardJones(
        kT=1.0, well_depth=1.0, yukawa_strength=0.0, lj_num=12, cutoff=2.5,
        debye_length=0.5)])),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_brick_equals_thread_per_particle(case):
    n, phi, rc, make = CASES[case]
    geom = ps.SystemGeometry.exact(n, phi, rc)
    pos, vel = state(n, geom.box, 31)
    same_bits(run(geom, make(), "brick", pos, vel), run(geom, make(), "sparse", pos, vel))


def test_brick_equals_thread_per_particle_s3():
    """S = 3: every neighbour cell is a periodic image (exact minimum image)."""
    geom = ps.SystemGeometry(n=60, box=7.5, cells_per_dim=3, cell_size=2.5)
    pos, vel = state(60, 7.5, 5)
    f = lambda: ps.LennardJones(kT=1.0, well_depth=1.0, yukawa_strength=0.0, lj_num=12,  # noqa
                                cutoff=2.5, debye_length=0.5)
    same_bits(run(geom, f(), "brick", pos, vel), run(geom, f(), "sparse", pos, vel))


def test_brick_equals_thread_per_particle_one_lane():
    """262,144 particles: one lane per particle in both kernels."""
    n = 262144
    geom = ps.SystemGeometry.exact(n, 0.15, 1.1)
    pos, vel = state(n, geom.box, 8)
    same_bits(run(geom, ao(), "brick", pos, vel, 3), run(geom, ao(), "sparse", pos, vel, 3))


def test_brick_chunked_halo():
    """A 14^3 lattice cluster (spacing 1.0) in a box of 30: bricks inside
    the cluster stage their halo in several chunks."""
    g = np.arange(14, dtype=float) * 1.0 + 8.25
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = len(pos)
    geom = ps.SystemGeometry(n=n, box=30.0, cells_per_dim=27, cell_size=30.0 / 27,
                             integer=False)
    vel = np.random.default_rng(3).standard_normal((n, 3))
    a = run(geom, ao(), "brick", pos, vel, 3)
    b = run(geom, ao(), "sparse", pos, vel, 3)
    for k in ("pos", "pos0", "vel"):
        assert np.allclose(a[k], b[k], rtol=0, atol=1e-12), k
    for k in ("f0", "f1"):
        scale = np.abs(b[k]).max()
        assert np.abs(a[k] - b[k]).max() <= 1e-12 * scale, k


def test_ensemble_brick_equals_thread_per_particle():
    rng = np.random.default_rng(21)
    reps = []
    for _ in range(10):
        n = int(rng.integers(150, 1500))
        phi = rng.uniform(0.05, 0.35)
        q = rng.uniform(0.05, 0.2)
        g = ps.SystemGeometry.exact(n, phi, 1.0 + q)
        reps.append(ps.Replica(g, ao(q, rng.uniform(2.0, 8.0))))
    pos = np.concatenate([ps.rsa_positions(r.geometry.n, r.geometry.box, 50 + k)
                          for k, r in enumerate(reps)])
    vel = np.random.default_rng(2).standard_normal((len(pos), 3))
    out = []
    for path in ("brick", "sparse"):
        print(f"ensemble path={path}", flush=True)
        with ps.Ensemble(reps, INTEG, kernel_path=path) as ens:
            ens.set_state(pos, vel)
            ens.step(8)
            ens.sync()
            assert ens.last_error()[0] == 0
            out.append(ens.get_state())
    for k in ("pos", "pos0", "vel", "frc"):
        assert np.array_equal(out[0][k], out[1][k]), k
