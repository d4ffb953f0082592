"""The PSim::IForce plugin adaptors: exported symbols (no GPU needed)."""
import ctypes
import os

import pytest

PLUGIN_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "proteinsim2015_amd", "plugins")


@pytest.mark.parametrize("name", ["LJ", "AO", "Cal"])
def test_plugin_exports_factory_and_batch_entry(name):
    path = os.path.join(PLUGIN_DIR, f"{name}.so")
    assert os.path.exists(path), "build the plugins with python -m proteinsim2015_amd.build"
    lib = ctypes.CDLL(path)          # resolves libpsim_gpu.so through its rpath
    assert hasattr(lib, "getForce")  # extern "C" getForce(config*), LJPotential.h:74-77
    assert hasattr(lib, "psim_force_batch")


def test_demo_host_reports_missing_plugin(tmp_path):
    import subprocess
    demo = os.path.join(PLUGIN_DIR, "psim_host_demo")
    cfg = tmp_path / "settings.cfg"
    cfg.write_text("nParticles = 2\nconc = 0.01\nscale = 3\n")
    pos = tmp_path / "pos.f64"
    pos.write_bytes(b"\0" * 48)
    r = subprocess.run([demo, str(tmp_path / "missing.so"), str(cfg), str(pos),
                        str(tmp_path / "f.f64")], capture_output=True, text=True)
    assert r.returncode == 100   # runSim.cpp:41-45
