"""psim_gpu_forces_noalloc (forces at device-resident positions, no device
allocation) against psim_gpu_forces on a context, through the C ABI (needs a
GPU).  Two calls in a row, the second with a smaller input on another kernel
path: the forces must be bit-identical to the context's, and within 1e-6
relative of the oracle."""
import ctypes as C

import numpy as np
import pytest

import proteinsim2015_amd as ps
from proteinsim2015_amd import _lib
from cases import PHI_P_C2, force_params
from helpers import force_rel_err

pytestmark = pytest.mark.gpu


def case_systems():
    """(modules, geometry, positions, oracle force params): a dense-path
    Calibration + LJ 12-6 system of 4096 particles, then a smaller sparse AO
    system of 1500 particles."""
    n1, box1, S1 = 4096, 21.0, 7
    g1 = ps.SystemGeometry(n=n1, box=box1, cells_per_dim=S1, cell_size=box1 / S1)
    m1 = [ps.Calibration(cutoff=2.5),
          ps.LennardJones(kT=1.0, well_depth=1.0, yukawa_strength=0.0, lj_num=6, cutoff=2.5,
                          debye_length=0.5)]
    p1 = ps.rsa_positions(n1, box1, 1234)
    f1 = force_params(1 | 2, {"cutOff": 2.5, "kT": 1.0, "wellDepth": 1.0,
                              "yukawaStrength": 0.0, "ljNum": 6, "debyeLength": 0.5})
    n2 = 1500
    g2 = ps.SystemGeometry.exact(n2, 0.15, 1.1)
    m2 = [ps.AOPotential(kT=1.0, well_depth=PHI_P_C2, cutoff=1.1)]
    p2 = ps.rsa_positions(n2, g2.box, 99)
    f2 = force_params(4, {"cutOff": 1.1, "kT": 1.0, "wellDepth": PHI_P_C2})
    return [(m1, g1, p1, f1, "dense"), (m2, g2, p2, f2, "sparse")]


def test_noalloc_forces_equal_context_forces(oracle):
    import torch
    lib = _lib.load()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    integ = ps.BrownianIntegrator(kT=1.0, gamma=1.0, dt=1e-3, seed=1)
    for mods, geom, pos, fp, path in case_systems():
        print(f"noalloc: n={geom.n} path={path}", flush=True)
        with ps.GpuSystem(geom, mods, integ) as sysg:
            assert sysg.kernel_info()["path"] == path
            sysg.set_state(pos)
            st_ref, f_ref = sysg.forces_now()
            params = sysg.params
        assert st_ref == 0
        d_pos = torch.tensor(pos, dtype=torch.float64, device=dev).contiguous()
        d_f = torch.full_like(d_pos, float("nan"))
        torch.cuda.synchronize()
        st = lib.psim_gpu_forces_noalloc(C.byref(params), C.c_void_p(d_pos.data_ptr()),
                                         C.c_void_p(d_f.data_ptr()),
                                         C.c_void_p(stream.cuda_stream))
        assert st == 0
        f = d_f.cpu().numpy()
        assert np.array_equal(f, f_ref), f"max |dF| {np.abs(f - f_ref).max():.3g}"
        _, _, fs = oracle.pair_stats(fp, pos, geom.cells_per_dim, geom.cell_size, geom.box,
                                     abs_sum=True)
        sto, fo, _ = oracle.forces(fp, pos, geom.cells_per_dim, geom.cell_size, geom.box,
                                   half_uv=geom.box / 2)
        assert sto == 0
        assert force_rel_err(f, fo, fs).max() <= 1e-6


def test_noalloc_reports_overlap_and_rejects_oversize():
    import torch
    lib = _lib.load()
    dev = torch.device("cuda:0")
    integ = ps.BrownianIntegrator(kT=1.0, gamma=1.0, dt=1e-3, seed=1)
    geom = ps.SystemGeometry(n=3, box=10.0, cells_per_dim=5, cell_size=2.0)
    pos = np.array([[5.0, 5.0, 5.0], [5.79, 5.0, 5.0], [1.0, 1.0, 1.0]])
    with ps.GpuSystem(geom, ps.Calibration(cutoff=2.0), integ) as sysg:
        params = sysg.params
    d_pos = torch.tensor(pos, dtype=torch.float64, device=dev)
    d_f = torch.empty_like(d_pos)
    assert lib.psim_gpu_forces_noalloc(C.byref(params), C.c_void_p(d_pos.data_ptr()),
                                       C.c_void_p(d_f.data_ptr()), None) == 7704
    # a system whose cell build does not fit the scratch array: refused
    # before any device work
    big = _lib.PsimParams.from_buffer_copy(params)
    big.n_particles = 4_000_000
    big.cells_per_dim = 100
    big.box = 200.0
    big.cell_size = 2.0
    assert lib.psim_gpu_forces_noalloc(C.byref(big), C.c_void_p(d_pos.data_ptr()),
                                       C.c_void_p(d_f.data_ptr()), None) == _lib.E_NOMEM
