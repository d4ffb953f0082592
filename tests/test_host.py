"""Host-side logic and the C ABI surface (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy.spatial import cKDTree

import proteinsim2015_amd as ps
from proteinsim2015_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(psim_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = declared_functions(os.path.join(ROOT, "include", "psim_gpu.h"))
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTS)
    assert lib.psim_gpu_abi_version() == _lib.ABI_VERSION


def test_status_strings():
    lib = _lib.load()
    for code in (0, 7701, 7703, 7704, 7705, -1, -2, -3, -4):
        assert lib.psim_gpu_status_string(code)


def test_create_rejects_bad_parameters():
    lib = _lib.load()
    p = _lib.PsimParams()
    ctx = ctypes.c_void_p()
    assert lib.psim_gpu_create(ctypes.byref(p), ctypes.byref(ctx)) == -1
    p.n_particles, p.cells_per_dim, p.box, p.cell_size = 10, 2, 4.0, 2.0
    p.cutoff, p.dt, p.mass, p.force_mask = 1.0, 1e-3, 1.0, 1
    assert lib.psim_gpu_create(ctypes.byref(p), ctypes.byref(ctx)) == -4   # S < 3
    assert not ctx.value


def test_config_parsing_matches_reference_semantics(tmp_path):
    # ClusteredUI/settings.cfg style: trailing ';', spaces, '#' comments
    f = tmp_path / "settings.cfg"
    f.write_text("# comment\nnParticles = 2500\nconc = 0.05\nscale = 5\nwellDepth = 10;\n"
                 "debyeLength = 0.5;\nljNum = 18;\nforce = LJ\n  kT=1\n")
    cfg = ps.Config(str(f))
    assert cfg.get_param("nParticles", 1000) == 2500
    assert cfg.get_param("wellDepth", 1.0) == 10.0
    assert cfg.get_param("ljNum", 18) == 18
    assert cfg.get_param("force", "") == "LJ"
    assert cfg.get_param("kT", 0.5) == 1.0
    assert cfg.get_param("missing", 7) == 7
    g = ps.SystemGeometry.from_config(cfg)
    assert (g.box, g.cell_size, g.cells_per_dim) == (25.0, 5.0, 5)
    lj = ps.LennardJones.from_config(cfg)
    assert (lj.well_depth, lj.lj_num, lj.cutoff, lj.yukawa_strength) == (10.0, 18, 2.5, 8.0)
    integ = ps.BrownianIntegrator.from_config(cfg)
    assert (integ.gamma, integ.vel_freq, integ.seed) == (0.5, 1000, 90210)


def test_geometry_matches_oracle(oracle):
    for n, conc, scale in [(4096, 0.2, 7), (262144, 0.15, 48), (2097152, 0.3, 51), (2500, 0.05, 5)]:
        g = ps.SystemGeometry.from_config(ps.Config({"nParticles": n, "conc": conc,
                                                     "scale": scale}))
        box, cell, phi = oracle.geometry(n, conc, scale)
        assert (g.box, g.cell_size) == (box, cell)
        assert g.volume_fraction == pytest.approx(phi, rel=1e-12)


def test_exact_geometry():
    g = ps.SystemGeometry.exact(2097152, 0.30, 3.0)
    assert g.cells_per_dim == 51 and g.cell_size >= 3.0
    assert g.volume_fraction == pytest.approx(0.30, rel=1e-12)


def test_ao_well_depth_mapping():
    assert ps.AOPotential.well_depth_for_contact(4.0, 0.1) == pytest.approx(0.25)


@pytest.mark.parametrize("n,phi", [(3000, 0.2), (20000, 0.3), (50, 0.01)])
def test_rsa_no_overlap_and_deterministic(n, phi):
    g = ps.SystemGeometry.exact(n, phi, 1.0)
    a = ps.rsa_positions(n, g.box, 17)
    b = ps.rsa_positions(n, g.box, 17)
    assert np.array_equal(a, b)
    assert a.min() >= 0 and a.max() < g.box
    d, _ = cKDTree(a, boxsize=g.box).query(a, k=2)
    assert d[:, 1].min() >= 1.0
    c = ps.rsa_positions(n, g.box, 18)
    assert not np.array_equal(a, c)


def test_rsa_impossible_density_reports_7701():
    with pytest.raises(ps.PsimError) as e:
        ps.rsa_positions(600, 6.0, 1.0, 1)   # phi ~ 1.45
    assert e.value.status == 7701
