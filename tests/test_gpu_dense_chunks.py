"""Cell-per-workgroup pair kernel past one LDS stage (needs a GPU).

The dense kernel stages at most 512 neighbour candidates at a time and takes
at most 256 home particles per pass.  These systems have cells coarser than
the cutoff, so the 27-cell neighbourhood holds about 890 candidates (two
staging chunks) or about 8,400 candidates with 312 home particles per cell
(seventeen chunks, two home-particle passes).  Forces are compared with the
oracle at the parity tolerance (1e-6 relative, metric
|dF_i| / max(|F_i|, sum_j |f_ij|)) on both the dense and the
thread-per-particle kernel."""
import math

import pytest

import proteinsim2015_amd as ps
from cases import force_params
from helpers import force_rel_err, pair_abs_sum
from oracle.oracle import PO_LJ
from test_gpu_parity import FORCE_TOL, pair_fmag

pytestmark = pytest.mark.gpu

CFG = {"cutOff": 2.5, "kT": 1.0, "wellDepth": 1.0, "yukawaStrength": 0.0, "ljNum": 6,
       "debyeLength": 0.5}


@pytest.mark.parametrize("n, cells", [(4096, 5), (20000, 4)])
def test_dense_forces_past_one_stage(oracle, n, cells):
    box = (n * math.pi / 6.0 / 0.30) ** (1.0 / 3.0)
    cell = box / cells
    assert cell >= CFG["cutOff"]
    geom = ps.SystemGeometry(n=n, box=box, cells_per_dim=cells, cell_size=cell, integer=False)
    pos = ps.rsa_positions(n, box, 4321)
    lj = ps.LennardJones.from_config(ps.Config(CFG))
    integ = ps.BrownianIntegrator(kT=1.0, gamma=1.0, dt=1e-3, seed=7)
    print(f"dense chunks: n={n} S={cells} occupancy={n / cells ** 3:.1f}", flush=True)
    forces = {}
    for path in ("dense", "sparse"):
        with ps.GpuSystem(geom, lj, integ, kernel_path=path) as sysg:
            sysg.set_state(pos)
            assert sysg.kernel_info()["path"] == path
            st, f = sysg.forces_now()
        assert st == 0
        forces[path] = f
    fp = force_params(PO_LJ, CFG)
    sto, fo, _ = oracle.forces(fp, pos, cells, cell, box, half_uv=box / 2)
    assert sto == 0
    denom = pair_abs_sum(pos, box, fp.cutoff, pair_fmag(fp))
    for path, f in forces.items():
        err = force_rel_err(f, fo, denom)
        assert err.max() <= FORCE_TOL, f"{path}: max rel err {err.max():.3g}"
