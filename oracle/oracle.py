"""ctypes wrapper around oracle/libpsim_oracle.so (the C restatement).

TEST INFRASTRUCTURE ONLY -- see oracle/psim_oracle.h for what each function
restates and how it is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsim_oracle.so")
SRC_PATH = os.path.join(HERE, "psim_oracle.c")

PO_CAL, PO_LJ, PO_AO = 1, 2, 4


def build_oracle(force: bool = False) -> str:
    """Compile the C restatement (gcc, no FMA contraction, OpenMP)."""
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < os.path.getmtime(SRC_PATH):
        cmd = ["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fopenmp",
               "-fno-fast-math", "-o", LIB_PATH, SRC_PATH, "-lm"]
        subprocess.check_call(cmd, cwd=HERE)
    return LIB_PATH


class ForceParams(C.Structure):
    _fields_ = [("mask", C.c_int), ("lj_num", C.c_int), ("cutoff", C.c_double),
                ("kT", C.c_double), ("lj_well_depth", C.c_double),
                ("yukawa_strength", C.c_double), ("debye_length", C.c_double),
                ("ao_well_depth", C.c_double)]

    @classmethod
    def make(cls, mask, cutoff, kT=1.0, lj_well_depth=10.0, yukawa_strength=8.0,
             debye_length=0.5, lj_num=18, ao_well_depth=0.261):
        # defaults are the reference's getParam defaults (LJPotential.cpp:34-54,
        # AOPotential.cpp:40-47)
        return cls(mask, lj_num, cutoff, kT, lj_well_depth, yukawa_strength,
                   debye_length, ao_well_depth)


class GBCoeffs(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "kT", "mass", "gamma", "dt", "dt_inv", "y", "coeff0", "coeff1", "coeff2",
        "coeff3", "sig1", "sig2", "corr", "dev", "goy2", "goy3", "hn", "gn")]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Oracle:
    def __init__(self, path: str | None = None):
        self.lib = C.CDLL(path or build_oracle())
        L = self.lib
        L.po_geometry.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double,
                                  C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]
        L.po_hash.argtypes = [C.c_int, _dp, C.c_double, C.c_int, _ip]
        L.po_build_cells.argtypes = [C.c_int, _ip, C.c_int, _ip, _ip, _ip]
        L.po_reorder.argtypes = [C.c_int, _ip, _dp, C.c_double, _dp]
        L.po_forces.argtypes = [C.POINTER(ForceParams), C.c_int, _dp, _ip, _ip, _ip,
                                C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                                _ip, C.c_int]
        L.po_forces.restype = C.c_int
        L.po_pair_stats.argtypes = [C.POINTER(ForceParams), C.c_int, _dp, _ip, _ip, _ip,
                                    C.c_int, C.c_double, C.c_double, _ip, _up, C.c_void_p,
                                    C.c_int, C.c_int]
        L.po_gb_coeffs_init.argtypes = [C.c_double] * 4 + [C.POINTER(GBCoeffs)]
        L.po_gb_first_step.argtypes = [C.POINTER(GBCoeffs), C.c_int, C.c_double,
                                       _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                       C.POINTER(C.c_int)]
        L.po_gb_first_step.restype = C.c_int
        L.po_gb_normal_step.argtypes = [C.POINTER(GBCoeffs), C.c_int, C.c_double,
                                        _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                        C.c_int, C.POINTER(C.c_int)]
        L.po_gb_normal_step.restype = C.c_int
        L.po_fastrandom_streams.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
        L.po_fastrandom_streams.restype = C.c_int
        L.po_philox_normals.argtypes = [C.c_ulonglong, C.c_ulonglong, C.c_int,
                                        C.c_int, _dp]
        L.po_philox4x32_10.argtypes = [C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                       C.POINTER(C.c_uint)]
        L.po_run.argtypes = [C.POINTER(ForceParams), C.POINTER(GBCoeffs), C.c_int,
                             C.c_double, C.c_double, C.c_int, C.c_double,
                             C.c_double, C.c_int, C.POINTER(C.c_int),
                             C.POINTER(C.c_double), _dp, _dp, _dp, _dp, _dp, _dp,
                             C.c_int, C.c_void_p, C.c_ulonglong,
                             C.POINTER(C.c_ulonglong), _ip, C.c_int]
        L.po_run.restype = C.c_int
        L.po_run_ids.argtypes = L.po_run.argtypes + [C.c_int]
        L.po_run_ids.restype = C.c_int
        L.po_temperature.argtypes = [C.c_int, _dp]
        L.po_temperature.restype = C.c_double
        L.po_pair_histogram.argtypes = [C.c_int, _dp, C.c_double, C.c_double,
                                        C.c_double, C.c_int, _dp]
        L.po_track_update.argtypes = [C.c_int, C.c_double, _dp, _ip]
        L.po_contacts.argtypes = [C.c_int, _dp, _ip, _ip, _ip, C.c_int, C.c_double,
                                  C.c_double, C.c_double, _ip, C.c_void_p]
        L.po_pair_histogram_smooth.argtypes = [C.c_int, _dp, C.c_double, C.c_double,
                                               C.c_double, C.c_int, C.c_double, _dp]
        L.po_pair_histogram_grad.argtypes = [C.c_int, _dp, C.c_double, C.c_double,
                                             C.c_double, C.c_int, C.c_double, _dp, _dp]
        L.po_pair_fnet_d.argtypes = [C.POINTER(ForceParams), C.c_double, C.c_double,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), _dp]
        L.po_adjoint.argtypes = [C.POINTER(ForceParams), C.POINTER(GBCoeffs), C.c_int,
                                 C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                 C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp,
                                 C.c_int, C.c_ulonglong, C.c_ulonglong, _dp, _dp,
                                 C.c_void_p, C.c_void_p, C.c_int, C.c_double]
        L.po_adjoint.restype = C.c_int

    # -- geometry / cells -------------------------------------------------
    def geometry(self, n, conc, scale, radius=0.5):
        b, c, phi = C.c_int(), C.c_int(), C.c_double()
        self.lib.po_geometry(n, conc, scale, radius, C.byref(b), C.byref(c), C.byref(phi))
        return b.value, c.value, phi.value

    def hash(self, pos, cell_size, S):
        pos = _f64(pos).reshape(-1, 3)
        key = np.empty(len(pos), np.int32)
        self.lib.po_hash(len(pos), pos.ravel(), float(cell_size), int(S), key)
        return key

    def build_cells(self, key, ncell):
        key = _i32(key)
        n = len(key)
        perm = np.empty(n, np.int32)
        cs = np.empty(ncell, np.int32)
        ce = np.empty(ncell, np.int32)
        self.lib.po_build_cells(n, key, ncell, perm, cs, ce)
        return perm, cs, ce

    def reorder(self, perm, pos, radius):
        pos = _f64(pos).reshape(-1, 3)
        out = np.empty(4 * len(pos))
        self.lib.po_reorder(len(pos), _i32(perm), pos.ravel(), radius, out)
        return out.reshape(-1, 4)

    # -- forces -------------------------------------------------------------
    def forces_table(self, fp, sorted4, perm, cs, ce, S, cell_size, box,
                     half_uv=None, nthreads=0):
        """Forces from a given cell table (reference sortedParticles layout)."""
        sorted4 = _f64(sorted4).reshape(-1, 4)
        n = len(sorted4)
        if half_uv is None:
            half_uv = float(int(box) // 2) if float(box).is_integer() else box / 2.0
        out = np.zeros(3 * n)
        err = np.full(2, -1, np.int32)
        st = self.lib.po_forces(C.byref(fp), n, sorted4.ravel(), _i32(perm), _i32(cs),
                                _i32(ce), int(S), float(cell_size), float(box),
                                float(half_uv), out, err, int(nthreads))
        return st, out.reshape(-1, 3), err

    def forces(self, fp, pos, S, cell_size, box, radius=0.5, half_uv=None, nthreads=0):
        """Cell build + forces (orig order)."""
        key = self.hash(pos, cell_size, S)
        perm, cs, ce = self.build_cells(key, S ** 3)
        sorted4 = self.reorder(perm, pos, radius)
        return self.forces_table(fp, sorted4, perm, cs, ce, S, cell_size, box,
                                 half_uv, nthreads)

    def pair_stats(self, fp, pos, S, cell_size, box, radius=0.5, abs_sum=False, id0=0,
                   nthreads=0):
        """Per-particle partner count and partner-hash sum (uint64) of the
        force loop's neighbour sets on a fresh cell table, by original index;
        with abs_sum, also sum_j |fNet(r_ij)|.  Partners are named by original
        index + id0."""
        pos = _f64(pos).reshape(-1, 3)
        n = len(pos)
        key = self.hash(pos, cell_size, S)
        perm, cs, ce = self.build_cells(key, S ** 3)
        sorted4 = self.reorder(perm, pos, radius)
        cnt = np.empty(n, np.int32)
        hs = np.empty(n, np.uint64)
        fs = np.empty(n) if abs_sum else None
        self.lib.po_pair_stats(C.byref(fp), n, sorted4.ravel(), perm, cs, ce, int(S),
                               float(cell_size), float(box), cnt, hs,
                               None if fs is None else fs.ctypes.data_as(C.c_void_p),
                               int(id0), int(nthreads))
        return (cnt, hs, fs) if abs_sum else (cnt, hs)

    # -- integrator ---------------------------------------------------------
    def gb_coeffs(self, kT, mass, gamma, dt):
        c = GBCoeffs()
        self.lib.po_gb_coeffs_init(kT, mass, gamma, dt, C.byref(c))
        return c

    def gb_first_step(self, c, box, pos, pos0, vel, frc, noise_b):
        pos, pos0 = _f64(pos).copy(), _f64(pos0).copy()
        n = pos.size // 3
        memx = np.zeros(3 * n)
        memcorr = np.zeros(3 * n)
        ei = C.c_int(-1)
        st = self.lib.po_gb_first_step(C.byref(c), n, box, pos.ravel(), pos0.ravel(),
                                       _f64(vel).ravel(), _f64(frc).ravel(), memx,
                                       memcorr, _f64(noise_b).ravel(), C.byref(ei))
        return st, pos.reshape(-1, 3), pos0.reshape(-1, 3), memx.reshape(-1, 3)

    def gb_normal_step(self, c, box, pos, pos0, vel, frc, frc0, memx, noise_a,
                       noise_b, do_vel):
        pos, pos0, vel, memx = (_f64(a).copy() for a in (pos, pos0, vel, memx))
        n = pos.size // 3
        memcorr = np.zeros(3 * n)
        ei = C.c_int(-1)
        st = self.lib.po_gb_normal_step(C.byref(c), n, box, pos.ravel(), pos0.ravel(),
                                        vel.ravel(), _f64(frc).ravel(),
                                        _f64(frc0).ravel(), memx.ravel(), memcorr,
                                        _f64(noise_a).ravel(), _f64(noise_b).ravel(),
                                        int(do_vel), C.byref(ei))
        return (st, pos.reshape(-1, 3), pos0.reshape(-1, 3), vel.reshape(-1, 3),
                memx.reshape(-1, 3), memcorr.reshape(-1, 3))

    # -- noise --------------------------------------------------------------
    def fastrandom_streams(self, seed, n, ndraws):
        out = np.empty(n * ndraws)
        st = self.lib.po_fastrandom_streams(int(seed), int(n), int(ndraws), out)
        return st, out.reshape(n, ndraws)

    def philox_normals(self, seed, step, n, id0=0):
        out = np.empty(6 * n)
        self.lib.po_philox_normals(int(seed), int(step), int(id0), int(n), out)
        return out.reshape(6, n)

    def philox_block(self, ctr, key):
        c = (C.c_uint * 4)(*ctr)
        k = (C.c_uint * 2)(*key)
        o = (C.c_uint * 4)()
        self.lib.po_philox4x32_10(c, k, o)
        return list(o)

    # -- full run -------------------------------------------------------------
    def run(self, fp, c, state, nsteps, box, S, cell_size, radius=0.5,
            vel_freq=1000, noise=None, seed=0, half_uv=None, nthreads=0, id0=0):
        """Advance ``state`` (dict of AoS3 arrays + counters) in place."""
        n = state["pos"].shape[0]
        if half_uv is None:
            half_uv = float(int(box) // 2) if float(box).is_integer() else box / 2.0
        arrs = {k: np.ascontiguousarray(state[k], dtype=np.float64).reshape(-1)
                for k in ("pos", "pos0", "vel", "frc", "frc0", "memx")}
        vc = C.c_int(state.get("vel_counter", 0))
        t = C.c_double(state.get("time", 0.0))
        si = C.c_ulonglong(state.get("step_index", 0))
        err = np.full(2, -1, np.int32)
        nz = None
        if noise is not None:
            nz = _f64(noise).reshape(-1)
            assert nz.size == nsteps * 6 * n
        st = self.lib.po_run_ids(C.byref(fp), C.byref(c), n, float(box), float(half_uv),
                                 int(S), float(cell_size), float(radius), int(vel_freq),
                                 C.byref(vc), C.byref(t), arrs["pos"], arrs["pos0"],
                                 arrs["vel"], arrs["frc"], arrs["frc0"], arrs["memx"],
                                 int(nsteps),
                                 None if nz is None else nz.ctypes.data_as(C.c_void_p),
                                 int(seed), C.byref(si), err, int(nthreads), int(id0))
        for k, v in arrs.items():
            state[k] = v.reshape(-1, 3)
        state["vel_counter"], state["time"], state["step_index"] = vc.value, t.value, si.value
        return st, err

    # -- analysis -------------------------------------------------------------
    def temperature(self, vel):
        vel = _f64(vel).reshape(-1, 3)
        return self.lib.po_temperature(len(vel), vel.ravel())

    def pair_histogram(self, pos, box, r_lo, width, nbins):
        pos = _f64(pos).reshape(-1, 3)
        out = np.empty(nbins)
        self.lib.po_pair_histogram(len(pos), pos.ravel(), box, r_lo, width, nbins, out)
        return out

    def track_update(self, box, pos0, pbc):
        """analysisTracker.cpp:38-64 on AoS3 arrays; updates pbc in place."""
        pos0 = _f64(pos0).reshape(-1)
        assert pbc.dtype == np.int32 and pbc.flags.c_contiguous
        self.lib.po_track_update(pos0.size // 3, float(box), pos0, pbc.reshape(-1))

    def contacts(self, pos, S, cell_size, box, rc=1.2, radius=0.5, labels=False):
        """Coordination by original index (and component labels) from the
        reference's 27-cell contact scan on a fresh cell table."""
        pos = _f64(pos).reshape(-1, 3)
        n = len(pos)
        key = self.hash(pos, cell_size, S)
        perm, cs, ce = self.build_cells(key, S ** 3)
        sorted4 = self.reorder(perm, pos, radius)
        coord = np.empty(n, np.int32)
        lab = np.empty(n, np.int32) if labels else None
        self.lib.po_contacts(n, sorted4.ravel(), perm, cs, ce, int(S), float(cell_size),
                             float(box), float(rc), coord,
                             None if lab is None else lab.ctypes.data_as(C.c_void_p))
        return (coord, lab) if labels else coord

    def pair_histogram_smooth(self, pos, box, r_lo, width, nbins, sigma):
        pos = _f64(pos).reshape(-1, 3)
        out = np.empty(nbins)
        self.lib.po_pair_histogram_smooth(len(pos), pos.ravel(), float(box), float(r_lo),
                                          float(width), int(nbins), float(sigma), out)
        return out

    def pair_histogram_grad(self, pos, box, r_lo, width, nbins, sigma, w):
        pos = _f64(pos).reshape(-1, 3)
        out = np.empty(pos.size)
        self.lib.po_pair_histogram_grad(len(pos), pos.ravel(), float(box), float(r_lo),
                                        float(width), int(nbins), float(sigma), _f64(w),
                                        out)
        return out.reshape(-1, 3)

    def pair_fnet_d(self, fp, r, size=1.0):
        f, d = C.c_double(), C.c_double()
        th = np.empty(5)
        self.lib.po_pair_fnet_d(C.byref(fp), float(r), float(size), C.byref(f), C.byref(d),
                                th)
        return f.value, d.value, th

    def adjoint(self, fp, c, state, nsteps, box, S, cell_size, gx, radius=0.5,
                vel_freq=1000, seed=0, half_uv=None, nthreads=0, cut_window=0.0):
        """dL/dtheta (5), dL/dx_0 and x_K for nsteps from ``state`` (not
        modified); cut_window > 0 adds the cutoff-crossing term to dL/dcutOff."""
        n = state["pos"].shape[0]
        if half_uv is None:
            half_uv = float(int(box) // 2) if float(box).is_integer() else box / 2.0
        arr = {k: _f64(state.get(k, np.zeros((n, 3)))).reshape(-1)
               for k in ("pos", "pos0", "vel", "frc", "memx")}
        dth = np.empty(5)
        gx0 = np.empty(3 * n)
        xk = np.empty(3 * n)
        st = self.lib.po_adjoint(C.byref(fp), C.byref(c), n, float(box), float(half_uv),
                                 int(S), float(cell_size), float(radius), int(vel_freq),
                                 int(state.get("vel_counter", 0)),
                                 float(state.get("time", 0.0)), arr["pos"], arr["pos0"],
                                 arr["vel"], arr["frc"], arr["memx"], int(nsteps), int(seed),
                                 int(state.get("step_index", 0)), _f64(gx).reshape(-1), dth,
                                 gx0.ctypes.data_as(C.c_void_p),
                                 xk.ctypes.data_as(C.c_void_p), int(nthreads),
                                 float(cut_window))
        return st, dth, gx0.reshape(-1, 3), xk.reshape(-1, 3)


_ORACLE = None


def load_oracle() -> Oracle:
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = Oracle()
    return _ORACLE
