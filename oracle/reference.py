"""ctypes wrapper around oracle/_ref/libpsimref.so, the harness that runs the
reference's own compiled plugins, integrator and RNG.

TEST INFRASTRUCTURE ONLY.  oracle/_ref/ is built from the reference sources by
oracle/ref/Makefile (``make -C oracle/ref REF=<reference checkout>``); it is
absent wherever the reference checkout is absent, and the tests that need it
skip there.
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
PLUGINS = {"LJ": "LJ.so", "AO": "AO.so", "Cal": "Cal.so"}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def reference_available() -> bool:
    return all(os.path.exists(os.path.join(REF_DIR, f))
               for f in ["libpsimref.so", *PLUGINS.values()])


def write_cfg(params: dict) -> str:
    """Write a ``key = value`` settings file (config.cpp:27-69 format)."""
    fd, path = tempfile.mkstemp(prefix="psimref_", suffix=".cfg")
    with os.fdopen(fd, "w") as f:
        for k, v in params.items():
            f.write(f"{k} = {v!r}\n" if isinstance(v, float) else f"{k} = {v}\n")
    return path


class Reference:
    def __init__(self):
        if not reference_available():
            raise RuntimeError("oracle/_ref is not built")
        self.lib = C.CDLL(os.path.join(REF_DIR, "libpsimref.so"))
        L = self.lib
        L.ref_fastrandom.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
        L.ref_forces.argtypes = [C.c_char_p, C.c_char_p, C.c_int, _dp, C.c_double,
                                 C.c_int, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip,
                                 _ip, _dp]
        L.ref_forces.restype = C.c_int
        L.ref_run.argtypes = [C.c_char_p, C.c_char_p, C.c_int, _dp, _dp, C.c_double,
                              C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                              C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_char_p,
                              C.c_int]
        L.ref_run.restype = C.c_int
        L.ref_run_sampled.argtypes = [C.c_char_p, C.c_char_p, C.c_int, _dp, _dp, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
        L.ref_run_sampled.restype = C.c_int
        L.ref_init.argtypes = [C.c_char_p, C.c_char_p]
        L.ref_init.restype = C.c_int
        L.ref_recover.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
        L.ref_recover.restype = C.c_int

    @staticmethod
    def plugin(name):
        """A reference plugin by name, or any plugin .so by path."""
        if name in PLUGINS:
            return os.path.join(REF_DIR, PLUGINS[name]).encode()
        return os.path.abspath(name).encode()

    def fastrandom(self, seed, n, ndraws):
        out = np.empty(n * ndraws)
        self.lib.ref_fastrandom(int(seed), int(n), int(ndraws), out)
        return out.reshape(n, ndraws)

    def forces(self, plugin, cfg: dict, pos, radius, box, cell, scale, nthreads=8):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = len(pos)
        ncell = scale ** 3
        f = np.empty(3 * n)
        perm = np.empty(n, np.int32)
        cs = np.empty(ncell, np.int32)
        ce = np.empty(ncell, np.int32)
        sorted4 = np.empty(4 * n)
        path = write_cfg(cfg)
        try:
            st = self.lib.ref_forces(self.plugin(plugin), path.encode(), n, pos.ravel(),
                                     float(radius), int(box), int(cell), int(scale),
                                     int(nthreads), f, perm, cs, ce, sorted4)
        finally:
            os.unlink(path)
        if st:
            raise RuntimeError(f"ref_forces failed ({st})")
        return f.reshape(-1, 3), perm, cs, ce, sorted4.reshape(-1, 4)

    def run(self, plugin, cfg: dict, pos, vel, radius, mass, box, cell, scale, dt,
            nsteps, nthreads=8, record=True, integrator=None, skip=0):
        """Trajectory dict (record=True) or the step-loop seconds (record=False)
        of the steps after the first ``skip``.  integrator: path of a library
        exporting getIntegrator(config*) to use instead of the reference's
        brownianIntegrator."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        n = len(pos)
        keys = ("pos", "pos0", "vel", "frc", "frc0")
        tr = {k: np.empty((nsteps + 1) * 3 * n) for k in keys} if record else None
        ptr = (lambda k: tr[k].ctypes.data_as(C.c_void_p)) if record else (lambda k: None)
        secs = C.c_double(0.0)
        path = write_cfg(cfg)
        try:
            st = self.lib.ref_run(self.plugin(plugin), path.encode(), n, pos.ravel(),
                                  vel.ravel(), float(radius), float(mass), int(box),
                                  int(cell), int(scale), float(dt), int(nthreads),
                                  int(nsteps), *(ptr(k) for k in keys), C.byref(secs),
                                  None if integrator is None
                                  else os.path.abspath(integrator).encode(), int(skip))
        finally:
            os.unlink(path)
        if st:
            raise RuntimeError(f"ref_run failed ({st})")
        if not record:
            return secs.value
        return {k: v.reshape(nsteps + 1, n, 3) for k, v in tr.items()}

    def run_sampled(self, plugin, cfg: dict, pos, vel, radius, mass, box, cell, scale, dt,
                    nsteps, every, nthreads=8):
        """Positions, re-imaged previous positions and velocities every
        ``every`` steps: arrays [nsteps // every + 1, n, 3]."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        n = len(pos)
        nf = nsteps // every + 1
        tr = {k: np.empty(nf * 3 * n) for k in ("pos", "pos0", "vel")}
        path = write_cfg(cfg)
        try:
            st = self.lib.ref_run_sampled(self.plugin(plugin), path.encode(), n, pos.ravel(),
                                          vel.ravel(), float(radius), float(mass), int(box),
                                          int(cell), int(scale), float(dt), int(nthreads),
                                          int(nsteps), int(every),
                                          *(tr[k].ctypes.data_as(C.c_void_p)
                                            for k in ("pos", "pos0", "vel")))
        finally:
            os.unlink(path)
        if st:
            raise RuntimeError(f"ref_run_sampled failed ({st})")
        return {k: v.reshape(nf, n, 3) for k, v in tr.items()}

    def init(self, cfg: dict, trial_dir: str) -> None:
        """Construct the reference's system in trial_dir (which receives
        sysConfig.cfg, initialState.txt and movie/system-initial.xyz)."""
        os.makedirs(os.path.join(trial_dir, "movie"), exist_ok=True)
        path = write_cfg(cfg)
        try:
            st = self.lib.ref_init(path.encode(), trial_dir.encode())
        finally:
            os.unlink(path)
        if st:
            raise RuntimeError(f"ref_init failed ({st})")


    def recover(self, cfg: dict, sys_state: str, stamp: str) -> str:
        """Load sys_state/snapshots/time-<stamp>/recovery.txt with the
        reference's RecoverySystem; returns the directory holding the
        initialState.txt it writes for the loaded particles."""
        rewind = os.path.join(sys_state, f"-rewind-{stamp}")
        # RecoverySystem::createRewindDir asks on stdin before writing into an
        # existing directory: the directory must not exist yet
        if os.path.exists(rewind):
            raise FileExistsError(rewind)
        path = write_cfg(cfg)
        try:
            st = self.lib.ref_recover(path.encode(), sys_state.encode(), stamp.encode())
        finally:
            os.unlink(path)
        if st:
            raise RuntimeError(f"ref_recover failed ({st})")
        return rewind
