"""Snapshot files (CPU): the reference's recovery.txt, movie frame and
sysConfig.cfg, written and read by libpsim_gpu.so's host code.

Pinned against the compiled reference: its RecoverySystem
(RecoverySystem.cpp:27-43) loads a recovery file written here and writes the
loaded particles back out through its own writeSystem; that text must equal
what read_recovery + write_recovery give for the same file (so both the
reader's re-imaging of the previous position and the writer's number format
are the reference's).  The movie frame and sysConfig.cfg are compared with
the files the reference's system constructor writes.
"""
import os

import numpy as np
import pytest

import proteinsim2015_amd as ps
from proteinsim2015_amd import _lib
from proteinsim2015_amd.engine import reference_initial_state
from proteinsim2015_amd.snapshot import read_recovery, write_recovery, write_xyz

CFG = {"nParticles": 300, "conc": 0.2, "scale": 3, "seed": 90210, "kT": 1.0, "radius": 0.5,
       "mass": 1.0, "timeStep": 0.001}
BOX = 9


def moved_state(seed=3):
    """Reference initial positions one 'step' later: the previous position
    is the old one in the new image, so near the faces it lies outside the
    box; forces span many magnitudes and signs."""
    st = reference_initial_state(300, BOX, 0.5, 90210, 0, 1.0)
    rng = np.random.default_rng(seed)
    old = st["pos"]
    new = old + rng.uniform(-0.4, 0.4, old.shape)
    pos = np.where(new < 0, new + BOX, np.where(new >= BOX, new - BOX, new))
    pos0 = old + np.where(new < 0, BOX, np.where(new >= BOX, -BOX, 0))
    frc = rng.standard_normal(old.shape) * 10.0 ** rng.integers(-8, 6, old.shape)
    frc0 = rng.standard_normal(old.shape) * 10.0 ** rng.integers(-8, 6, old.shape)
    return pos, pos0, frc, frc0


def test_recovery_round_trip_matches_the_reference_loader(reference, tmp_path):
    pos, pos0, frc, frc0 = moved_state()
    assert (pos0 < 0).any() and (pos0 >= BOX).any()
    snap = tmp_path / "trial" / "snapshots" / "time-7"
    snap.mkdir(parents=True)
    write_recovery(str(snap / "recovery.txt"), pos, pos0, frc, frc0, 1.0, 0.5)
    rewind = reference.recover(CFG, str(tmp_path / "trial"), "7")
    ref_text = open(os.path.join(rewind, "initialState.txt")).read()

    st = read_recovery(str(snap / "recovery.txt"), 300, BOX)
    write_recovery(str(tmp_path / "mine.txt"), st["pos"], st["pos0"], st["frc"], st["frc0"],
                   st["mass"][0], st["radius"][0])
    assert open(tmp_path / "mine.txt").read() == ref_text
    # the previous positions come back in the current image
    d = st["pos"] - st["pos0"]
    assert np.abs(d).max() < BOX / 2
    assert np.all((st["pos"] >= 0) & (st["pos"] < BOX))
    assert np.all(st["mass"] == 1.0) and np.all(st["radius"] == 0.5)


def test_recovery_text_layout(tmp_path):
    pos = np.array([[1.0, 2.5, 8.999999], [0.0, 1e-7, 123456.789]])
    pos0 = pos - 0.25
    p = tmp_path / "r.txt"
    write_recovery(str(p), pos, pos0, None, None, 1.0, 0.5)
    lines = open(p).read().split("\n")
    assert len(lines) == 2 and not open(p).read().endswith("\n")
    assert lines[0] == "1 2.5 9 0.75 2.25 8.75 0 0 0 0 0 0 1 0.5"
    assert lines[1].split()[2] == "123457" and lines[1].split()[1] == "1e-07"


def test_movie_frame_matches_the_reference(reference, tmp_path):
    trial = str(tmp_path / "t")
    reference.init(CFG, trial)
    st = reference_initial_state(300, BOX, 0.5, 90210, 0, 1.0)
    write_xyz(str(tmp_path / "mine.xyz"), st["pos"], 0.0)
    assert open(tmp_path / "mine.xyz").read() == \
        open(os.path.join(trial, "movie", "system-initial.xyz")).read()
    write_xyz(str(tmp_path / "t2.xyz"), st["pos"][:2], 2.5)
    assert open(tmp_path / "t2.xyz").read().split("\n")[:2] == ["2", "Current Time: 3"]


def test_sys_config_matches_the_reference(reference, tmp_path):
    trial = str(tmp_path / "t")
    reference.init(CFG, trial)
    ref = open(os.path.join(trial, "sysConfig.cfg")).read().split("\n")
    conc = 300 * (4.0 / 3.0) * np.arctan(1.0) * 4.0 * 0.125 / BOX ** 3
    p = str(tmp_path / "mine.cfg")
    _lib.check(_lib.load().psim_write_sys_config(p.encode(), trial.encode(), 300, conc, BOX, 3,
                                                 3, 1.0, 0.001, 1000, 0.25, 90210))
    mine = open(p).read().split("\n")
    # the reference writes cycleHour before the run sets it (an uninitialised
    # member); every other line must agree
    keep = [k for k, line in enumerate(ref) if not line.startswith("cycleHour")]
    assert [mine[k] for k in keep] == [ref[k] for k in keep]
    assert len(mine) == len(ref)


def test_recovery_errors(tmp_path):
    pos, pos0, frc, frc0 = moved_state()
    p = str(tmp_path / "r.txt")
    write_recovery(p, pos, pos0, frc, frc0)
    with pytest.raises(ps.PsimError) as e:
        read_recovery(p, 299, BOX)          # line count differs from nParticles
    assert e.value.status == _lib.E_IO
    with pytest.raises(ps.PsimError) as e:
        read_recovery(str(tmp_path / "missing.txt"), 300, BOX)
    assert e.value.status == _lib.E_IO
    bad = pos.copy()
    bad[5, 1] = 2.5 * BOX                   # still outside the box after one wrap
    write_recovery(p, bad, pos0, frc, frc0)
    with pytest.raises(ps.PsimError) as e:
        read_recovery(p, 300, BOX)
    assert e.value.status == ps.ERR_BOUNDS
    with open(p, "w") as f:
        f.write("1 2 3 4 5\n")
    with pytest.raises(ps.PsimError) as e:
        read_recovery(p, 1, BOX)
    assert e.value.status == _lib.E_IO


def test_cli_rejects_an_unknown_force(tmp_path):
    """runSim.cpp:41-45: a force that cannot be loaded ends the program with
    status 100 (before any device work)."""
    import subprocess
    from proteinsim2015_amd.snapshot import cli_path
    (tmp_path / "settings.cfg").write_text("trialName = x\nforce = Spring\nnParticles = 10\n")
    r = subprocess.run([cli_path()], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 100
    assert "force" in r.stderr
