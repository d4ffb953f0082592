"""Graph replay of the single-system step (needs a GPU): the same bits as eager
launches, across first steps, velocity steps, the lagged state buffers being
read back in between, parameter changes and a remainder of eager steps."""
import numpy as np
import pytest

import proteinsim2015_amd as ps

pytestmark = pytest.mark.gpu


def make(path, n=3000, phi=0.25, vel_freq=2, seed=5):
    geom = ps.SystemGeometry.exact(n, phi, 2.5)
    mods = [ps.Calibration(cutoff=2.5),
            ps.LennardJones(kT=1.0, well_depth=1.0, yukawa_strength=0.0, lj_num=6,
                            cutoff=2.5, debye_length=0.5)]
    integ = ps.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=1e-3, vel_freq=vel_freq,
                                  seed=seed)
    pos = ps.rsa_positions(n, geom.box, 321)
    vel = np.random.default_rng(7).standard_normal((n, 3))
    sysg = ps.GpuSystem(geom, mods, integ, kernel_path=path)
    sysg.set_state(pos, vel=vel)
    return sysg


def run(sysg, plan):
    """plan: list of ("step", k) / ("state",) / ("obs",) / ("lj", eps)."""
    out = []
    for item in plan:
        if item[0] == "step":
            assert sysg.step(item[1]) == 0
        elif item[0] == "state":
            out.append(sysg.get_state())
        elif item[0] == "obs":
            out.append(sysg.observables(1.2))
        elif item[0] == "lj":
            sysg.set_forces([ps.Calibration(cutoff=2.5),
                             ps.LennardJones(kT=1.0, well_depth=item[1], yukawa_strength=0.0,
                                             lj_num=6, cutoff=2.5, debye_length=0.5)])
    out.append(sysg.get_state())
    return out


PLAN = [("step", 13), ("state",), ("step", 6), ("obs",), ("lj", 1.5), ("step", 9)]


@pytest.mark.parametrize("path", ["sparse", "dense", "brick"])
def test_graph_replay_is_bit_identical(path):
    print(f"graph: path={path}", flush=True)
    with make(path) as a:
        eager = run(a, PLAN)
    with make(path) as b:
        b.set_graph(4)
        graph = run(b, PLAN)
    for e, g in zip(eager, graph):
        if isinstance(e, dict):
            for k in e:
                if isinstance(e[k], np.ndarray):
                    if k == "frc0":
                        continue
                    assert np.array_equal(e[k], g[k]), k
                else:
                    same_nan = e[k] != e[k] and g[k] != g[k]
                    assert e[k] == g[k] or same_nan, k
        else:
            assert e == g


def test_graph_replay_full_step_counts():
    """Replays only (no eager remainder), an odd number of replays, and a
    velocity step on every step."""
    with make("auto", vel_freq=0) as a:
        assert a.step(30) == 0
        ea = a.get_state()
    with make("auto", vel_freq=0) as b:
        b.set_graph(10)
        assert b.step(30) == 0
        gb = b.get_state()
    for k in ("pos", "pos0", "vel", "frc", "memx"):
        assert np.array_equal(ea[k], gb[k]), k
    assert ea["step_index"] == gb["step_index"] == 30
    assert ea["time"] == gb["time"]
