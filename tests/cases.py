"""Parity cases shared by the fixture generator and the tests.

Each force case names the reference plugin it pins, the reference config
keys handed to that plugin (the same keys the reference reads, see
LJPotential.cpp:34-54, AOPotential.cpp:40-47, Calibration.cpp:36), and a
reference-integer geometry (system.cpp:50-64).
"""
from __future__ import annotations

import math

from oracle.oracle import PO_AO, PO_CAL, PO_LJ, ForceParams

# AO contact depth 4 kT at q = 0.1: phi_p = 2 q U / (3 + 2 q)
PHI_P_C2 = 2 * 0.1 * 4.0 / (3 + 2 * 0.1)
PHI_P_C5 = 2 * 0.12 * 5.0 / (3 + 2 * 0.12)
# Yukawa contact strength 20 kT with lambda = 0.2: A = 20 / (lambda e^{-1/lambda})
YUK_A_C3 = 20.0 / (0.2 * math.exp(-1.0 / 0.2))

FORCE_CASES = {
    # name: (plugin, mask, reference config, n, conc, scale)
    "cal": ("Cal", PO_CAL, {"cutOff": 2.5}, 4096, 0.2, 7),
    "lj6": ("LJ", PO_LJ, {"cutOff": 2.5, "kT": 1.0, "wellDepth": 1.0,
                          "yukawaStrength": 0.0, "ljNum": 6, "debyeLength": 0.5},
            4096, 0.2, 7),
    "ljyuk": ("LJ", PO_LJ, {"cutOff": 3.0, "kT": 1.0, "wellDepth": 1.0,
                            "yukawaStrength": YUK_A_C3, "ljNum": 18, "debyeLength": 0.2},
              5306, 0.2999, 7),
    "ljdefault": ("LJ", PO_LJ, {"cutOff": 2.5, "kT": 1.0, "wellDepth": 10.0,
                                "yukawaStrength": 8.0, "ljNum": 18, "debyeLength": 0.5},
                  2500, 0.05, 5),
    "ao": ("AO", PO_AO, {"cutOff": 1.1, "kT": 1.0, "wellDepth": PHI_P_C2}, 4096, 0.15, 12),
    "ao12": ("AO", PO_AO, {"cutOff": 1.12, "kT": 1.0, "wellDepth": PHI_P_C5}, 4096, 0.15, 12),
    # S = 3: the per-pair reference minimum image path
    "lj_s3": ("LJ", PO_LJ, {"cutOff": 2.5, "kT": 1.0, "wellDepth": 1.0,
                            "yukawaStrength": 0.0, "ljNum": 6, "debyeLength": 0.5},
              100, 0.1, 3),
}

# trajectory cases: (plugin, mask, config incl. integrator keys, n, conc, scale, nsteps)
TRAJ_CASES = {
    "lj6_vel0": ("LJ", PO_LJ, {"cutOff": 2.5, "kT": 1.0, "wellDepth": 1.0,
                               "yukawaStrength": 0.0, "ljNum": 6, "debyeLength": 0.5,
                               "gamma": 1.0, "mass": 1.0, "timeStep": 0.001, "velFreq": 0},
                 512, 0.1, 4, 5),
    "ljyuk_vel1": ("LJ", PO_LJ, {"cutOff": 3.0, "kT": 1.0, "wellDepth": 1.0,
                                 "yukawaStrength": YUK_A_C3, "ljNum": 18, "debyeLength": 0.2,
                                 "gamma": 1.0, "mass": 1.0, "timeStep": 0.001, "velFreq": 1},
                   512, 0.1, 4, 5),
    "ao_lowgamma": ("AO", PO_AO, {"cutOff": 1.1, "kT": 1.0, "wellDepth": PHI_P_C2,
                                  "gamma": 0.01, "mass": 2.0, "timeStep": 0.0005,
                                  "velFreq": 2},
                    512, 0.1, 6, 5),
    "cal_default": ("Cal", PO_CAL, {"cutOff": 1.1, "kT": 1.0, "gamma": 0.5, "mass": 1.0,
                                    "timeStep": 0.001, "velFreq": 1000},
                    512, 0.1, 4, 5),
}

NOISE_SEED = 90210


def force_params(mask, cfg) -> ForceParams:
    wd_default = 0.261 if mask == PO_AO else 10.0
    return ForceParams.make(mask, cfg["cutOff"], kT=cfg.get("kT", 1.0),
                            lj_well_depth=cfg.get("wellDepth", wd_default),
                            yukawa_strength=cfg.get("yukawaStrength", 8.0),
                            debye_length=cfg.get("debyeLength", 0.5),
                            lj_num=cfg.get("ljNum", 18),
                            ao_well_depth=cfg.get("wellDepth", 0.261))


def noise_from_streams(streams, nsteps):
    """Arrange per-particle g250 streams into [nsteps][6][n]: the first step
    draws 3 (memX, brownianIntegrator.cpp:206-208), later steps draw 6 (memCorr
    then memX, :238-249)."""
    import numpy as np
    n = streams.shape[0]
    noise = np.zeros((nsteps, 6, n))
    noise[0, 3:6] = streams[:, 0:3].T
    for s in range(1, nsteps):
        o = 3 + 6 * (s - 1)
        noise[s, 0:3] = streams[:, o:o + 3].T
        noise[s, 3:6] = streams[:, o + 3:o + 6].T
    return noise
