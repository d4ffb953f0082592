"""Host-side mirror of the reference's interfaces over the C ABI.

Vocabulary follows the reference:
  * ``Config``       -- PSim::config, ``key = value`` files with typed
                        ``get_param(key, default)`` (config.cpp:27-175);
  * ``LennardJones``, ``AOPotential``, ``Calibration`` -- the three IForce
                        plugins' parameters with the reference's keys and
                        defaults (LJPotential.cpp:29-62, AOPotential.cpp:33-59,
                        Calibration.cpp:30-39);
  * ``BrownianIntegrator`` -- brownianIntegrator's parameters
                        (brownianIntegrator.cpp:27-99);
  * ``SystemGeometry`` -- systemState's box/cell geometry
                        (system.cpp:27-67);
  * ``GpuSystem``      -- one simulated system on the GPU: the state, the
                        force evaluation and the step loop of system::run.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import FORCE_AO, FORCE_CAL, FORCE_LJ, PAIRS_ADJOINT, PAIRS_STEP, PsimParams, check

_NUM = re.compile(r"^\s*[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?)")
_INT = re.compile(r"^\s*[+-]?\d+")


class Config:
    """PSim::config: ``key = value`` lines, ``#`` comment lines, whitespace
    trimmed (config.cpp:27-56); values converted like std::stoi/std::stod,
    which read the leading number (so ``wellDepth = 10;`` is 10)."""

    def __init__(self, source=None, **overrides):
        self.options: dict[str, str] = {}
        if isinstance(source, dict):
            self.options.update({k: str(v) for k, v in source.items()})
        elif source is not None:
            with open(source) as f:
                for line in f:
                    self._parse_line(line.rstrip("\n"))
        self.options.update({k: str(v) for k, v in overrides.items()})

    def _parse_line(self, line: str) -> None:
        if not line or line[0] == "#":
            return
        key, _, val = line.partition("=")
        self.options[key.strip()] = val.strip()

    def contains_key(self, key: str) -> bool:
        return key in self.options

    def get_param(self, key: str, default):
        if key not in self.options:
            return default
        raw = self.options[key]
        if isinstance(default, bool):
            return bool(int(_INT.match(raw).group(0)))
        if isinstance(default, int):
            m = _INT.match(raw)
            if not m:
                raise ValueError(f"config: {key} = {raw!r} is not an int")
            return int(m.group(0))
        if isinstance(default, float):
            m = _NUM.match(raw)
            if not m:
                raise ValueError(f"config: {key} = {raw!r} is not a number")
            return float(m.group(0))
        return raw


@dataclass
class Calibration:
    """Calibration.cpp:30-39: soft shell F = -36 r^-37 inside cutOff."""
    cutoff: float = 1.1
    mask: int = FORCE_CAL

    @classmethod
    def from_config(cls, cfg: Config):
        return cls(cutoff=cfg.get_param("cutOff", 1.1))


@dataclass
class LennardJones:
    """LJPotential.cpp:29-62: LJ (2n-n) plus Yukawa, both scaled by kT*wellDepth."""
    kT: float = 1.0
    well_depth: float = 10.0
    yukawa_strength: float = 8.0
    lj_num: int = 18
    cutoff: float = 2.5
    debye_length: float = 0.5
    mask: int = FORCE_LJ

    @classmethod
    def from_config(cls, cfg: Config):
        return cls(kT=cfg.get_param("kT", 1.0), well_depth=cfg.get_param("wellDepth", 10.0),
                   yukawa_strength=cfg.get_param("yukawaStrength", 8.0),
                   lj_num=cfg.get_param("ljNum", 18), cutoff=cfg.get_param("cutOff", 2.5),
                   debye_length=cfg.get_param("debyeLength", 0.5))


@dataclass
class AOPotential:
    """AOPotential.cpp:33-59: Asakura-Oosawa attraction (polymer reservoir
    fraction wellDepth) plus the r^-38 core; cutoff = 1 + q."""
    kT: float = 1.0
    well_depth: float = 0.261
    cutoff: float = 1.1
    mask: int = FORCE_AO

    @classmethod
    def from_config(cls, cfg: Config):
        return cls(kT=cfg.get_param("kT", 1.0), well_depth=cfg.get_param("wellDepth", 0.261),
                   cutoff=cfg.get_param("cutOff", 1.1))

    @staticmethod
    def well_depth_for_contact(u_contact_kT: float, q: float) -> float:
        """phi_p giving an AO contact energy of -u kT: U_AO(1) = -phi_p (3+2q)/(2q)."""
        return 2.0 * q * u_contact_kT / (3.0 + 2.0 * q)


PLUGINS = {"LJ": LennardJones, "AO": AOPotential, "Cal": Calibration,
           "LennardJones": LennardJones, "AOPotential": AOPotential,
           "Calibration": Calibration}


@dataclass
class ForceSet:
    """A sum of force modules sharing one cutoff (the reference's config has
    a single ``cutOff`` key read by every plugin)."""
    modules: list = field(default_factory=list)

    @property
    def mask(self) -> int:
        m = 0
        for mod in self.modules:
            m |= mod.mask
        return m

    @property
    def cutoff(self) -> float:
        cs = {mod.cutoff for mod in self.modules}
        if len(cs) != 1:
            raise ValueError(f"force modules disagree on cutOff: {sorted(cs)}")
        return cs.pop()

    def fill(self, p: PsimParams) -> None:
        p.force_mask = self.mask
        p.cutoff = self.cutoff
        p.lj_num, p.lj_well_depth, p.yukawa_strength, p.debye_length = 18, 10.0, 8.0, 0.5
        p.ao_well_depth = 0.261
        for m in self.modules:
            if isinstance(m, LennardJones):
                p.lj_num, p.lj_well_depth = m.lj_num, m.well_depth
                p.yukawa_strength, p.debye_length = m.yukawa_strength, m.debye_length
            elif isinstance(m, AOPotential):
                p.ao_well_depth = m.well_depth


@dataclass
class BrownianIntegrator:
    """brownianIntegrator.cpp:27-99 parameters (coefficients are computed in
    the native library with the reference's formulas)."""
    kT: float = 1.0
    mass: float = 1.0
    gamma: float = 0.5
    dt: float = 0.001
    vel_freq: int = 1000
    seed: int = 90210

    @classmethod
    def from_config(cls, cfg: Config):
        return cls(kT=cfg.get_param("kT", 1.0), mass=cfg.get_param("mass", 1.0),
                   gamma=cfg.get_param("gamma", 0.5), dt=cfg.get_param("timeStep", 0.001),
                   vel_freq=cfg.get_param("velFreq", 1000), seed=cfg.get_param("seed", 90210))


@dataclass
class SystemGeometry:
    n: int
    box: float
    cells_per_dim: int
    cell_size: float
    radius: float = 0.5
    integer: bool = True

    @property
    def ncell(self) -> int:
        return self.cells_per_dim ** 3

    @property
    def volume_fraction(self) -> float:
        return self.n * (4.0 / 3.0) * math.pi * self.radius ** 3 / self.box ** 3

    @classmethod
    def from_config(cls, cfg: Config):
        """system::setSystemConstants (system.cpp:27-67): integer box and
        cell, box = cellSize * scale."""
        n = cfg.get_param("nParticles", 1000)
        conc = cfg.get_param("conc", 0.01)
        scale = cfg.get_param("scale", 4)
        r = cfg.get_param("radius", 0.5)
        vP = n * (4.0 / 3.0) * math.atan(1.0) * 4.0 * r * r * r
        box = int(np.cbrt(vP / conc))
        cell = box // scale
        box = cell * scale
        return cls(n=n, box=float(box), cells_per_dim=scale, cell_size=float(cell),
                   radius=r, integer=True)

    @classmethod
    def exact(cls, n: int, phi: float, min_cell: float, radius: float = 0.5):
        """Exact box for volume fraction phi; S = floor(L / min_cell) cells of
        edge L / S >= min_cell (normally the cutoff)."""
        box = (n * (4.0 / 3.0) * math.pi * radius ** 3 / phi) ** (1.0 / 3.0)
        S = int(math.floor(box / min_cell))
        return cls(n=n, box=box, cells_per_dim=S, cell_size=box / S, radius=radius,
                   integer=False)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _aos3(a, n, name):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if a.size != 3 * n:
        raise ValueError(f"{name}: expected {n}x3 values, got {a.size}")
    return a


def reference_initial_state(n: int, box: int, radius: float = 0.5, seed: int = 90210,
                            seed_size: int = 0, kT: float = 1.0) -> dict:
    """The reference's own initial configuration (system::initParticles,
    initCheck and maxwellVelocityInit, systemInit.cpp:31-267) for an integer
    box: the same std::mt19937 draws in the same order, so the positions and
    velocities equal those the reference would create with this seed."""
    lib = _lib.load()
    pos, pos0, vel = (np.empty(3 * n) for _ in range(3))
    check(lib.psim_init_reference(int(n), int(box), float(radius), int(seed) & 0xFFFFFFFF,
                                  int(seed_size), float(kT), _ptr(pos), _ptr(pos0), _ptr(vel)),
          "psim_init_reference")
    return {"pos": pos.reshape(n, 3), "pos0": pos0.reshape(n, 3), "vel": vel.reshape(n, 3)}


def rsa_positions(n: int, box: float, seed: int, dmin: float = 1.0) -> np.ndarray:
    """Seeded random sequential addition (native, O(N)); error 7701 if the
    packing cannot be completed."""
    out = np.empty(3 * n)
    check(_lib.load().psim_init_rsa(int(n), float(box), float(dmin), int(seed), _ptr(out)),
          "psim_init_rsa")
    return out.reshape(n, 3)


class GpuSystem:
    """One system resident in HBM, advanced by the HIP kernels."""

    KERNEL_PATHS = {"auto": 0, "sparse": 1, "dense": 2, "brick": 3}

    def __init__(self, geometry: SystemGeometry, forces, integrator: BrownianIntegrator,
                 noise_seed: int | None = None, kernel_path: str = "auto"):
        self.lib = _lib.load()
        if not isinstance(forces, ForceSet):
            forces = ForceSet(list(forces) if isinstance(forces, (list, tuple)) else [forces])
        self.geometry, self.forces, self.integrator = geometry, forces, integrator
        p = PsimParams()
        p.n_particles = geometry.n
        p.cells_per_dim = geometry.cells_per_dim
        p.box = geometry.box
        p.cell_size = geometry.cell_size
        p.half_uv = 0.0
        p.kernel_path = self.KERNEL_PATHS[kernel_path]
        p.radius = geometry.radius
        p.mass, p.kT, p.gamma, p.dt = (integrator.mass, integrator.kT, integrator.gamma,
                                       integrator.dt)
        p.vel_freq = integrator.vel_freq
        forces.fill(p)
        kts = {m.kT for m in forces.modules if hasattr(m, "kT")}
        if kts and kts != {integrator.kT}:
            raise ValueError("force modules and integrator must share kT (one config key)")
        p.kT = integrator.kT
        p.seed = integrator.seed if noise_seed is None else noise_seed
        self.params = p
        self._ctx = C.c_void_p()
        check(self.lib.psim_gpu_create(C.byref(p), C.byref(self._ctx)), "psim_gpu_create")
        self.n = geometry.n

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if self._ctx:
            self.lib.psim_gpu_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state --------------------------------------------------------------
    def set_state(self, pos, pos0=None, vel=None, frc=None, memx=None, time=0.0,
                  vel_counter=0, step_index=0):
        n = self.n
        arrs = [_aos3(a, n, k) for a, k in ((pos, "pos"), (pos0, "pos0"), (vel, "vel"),
                                            (frc, "frc"), (memx, "memx"))]
        check(self.lib.psim_gpu_set_state(self._ctx, _ptr(arrs[0]), _ptr(arrs[1]),
                                          _ptr(arrs[2]), _ptr(arrs[3]), None, _ptr(arrs[4]),
                                          float(time), int(vel_counter), int(step_index)),
              "psim_gpu_set_state")

    def get_state(self) -> dict:
        n = self.n
        out = {k: np.empty(3 * n) for k in ("pos", "pos0", "vel", "frc", "memx")}
        t = C.c_double()
        vc = C.c_int()
        si = C.c_ulonglong()
        check(self.lib.psim_gpu_get_state(self._ctx, _ptr(out["pos"]), _ptr(out["pos0"]),
                                          _ptr(out["vel"]), _ptr(out["frc"]), None,
                                          _ptr(out["memx"]), C.byref(t), C.byref(vc),
                                          C.byref(si)), "psim_gpu_get_state")
        st = {k: v.reshape(n, 3) for k, v in out.items()}
        st.update(time=t.value, vel_counter=vc.value, step_index=si.value)
        return st

    def save(self, path: str) -> None:
        """Full-precision state file (positions, previous positions,
        velocities, last forces, integrator memory, time, counters and the
        displacement tracker)."""
        check(self.lib.psim_gpu_save(self._ctx, os.fsencode(path)), f"save {path}")

    def load(self, path: str) -> None:
        check(self.lib.psim_gpu_load(self._ctx, os.fsencode(path)), f"load {path}")

    def get_tracker(self):
        """(start positions, crossing counters), AoS3 by original index."""
        start = np.empty(3 * self.n)
        cross = np.empty(3 * self.n, dtype=np.int32)
        check(self.lib.psim_gpu_get_tracker(self._ctx, _ptr(start),
                                            cross.ctypes.data_as(C.POINTER(C.c_int))),
              "psim_gpu_get_tracker")
        return start.reshape(self.n, 3), cross.reshape(self.n, 3)

    def set_tracker(self, start, crossings) -> None:
        start = _aos3(start, self.n, "start")
        cross = np.ascontiguousarray(crossings, dtype=np.int32).reshape(-1)
        if cross.size != 3 * self.n:
            raise ValueError("crossings: expected n x 3 ints")
        check(self.lib.psim_gpu_set_tracker(self._ctx, _ptr(start),
                                            cross.ctypes.data_as(C.POINTER(C.c_int))),
              "psim_gpu_set_tracker")

    # -- hot path -------------------------------------------------------------
    def step(self, nsteps: int = 1, noise=None) -> int:
        """Advance nsteps; noise None -> in-kernel Philox, else [nsteps][6][n].
        Returns the status (0 or 7703/7704/7705)."""
        nz = None
        if noise is not None:
            nz = np.ascontiguousarray(noise, dtype=np.float64).reshape(-1)
            if nz.size != nsteps * 6 * self.n:
                raise ValueError("noise must be [nsteps][6][n]")
        st = self.lib.psim_gpu_step(self._ctx, int(nsteps), _ptr(nz))
        if st < 0:
            check(st, "psim_gpu_step")
        return st

    def step_device_noise(self, nsteps: int, device_ptr: int) -> None:
        check(self.lib.psim_gpu_step_device_noise(self._ctx, int(nsteps),
                                                  C.c_void_p(device_ptr)),
              "psim_gpu_step_device_noise")

    def set_graph(self, steps_per_graph: int) -> None:
        """Replay blocks of ``steps_per_graph`` (even) steps as one captured
        hipGraph when stepping with in-kernel noise (0: eager launches)."""
        check(self.lib.psim_gpu_set_graph(self._ctx, int(steps_per_graph)), "psim_gpu_set_graph")

    def kernel_info(self) -> dict:
        """The pair kernel this context runs ("sparse", "dense" or "brick"),
        the brick's home cells and the lanes per particle."""
        path, b = C.c_int(), (C.c_int * 4)()
        check(self.lib.psim_gpu_kernel_info(self._ctx, C.byref(path), b), "psim_gpu_kernel_info")
        return {"path": {1: "sparse", 2: "dense", 3: "brick"}[path.value],
                "brick": tuple(b[:3]), "lanes_per_particle": b[3]}

    def forces_now(self):
        """(status, forces AoS3 by original index) at the current positions."""
        out = np.empty(3 * self.n)
        st = self.lib.psim_gpu_forces(self._ctx, _ptr(out))
        if st < 0:
            check(st, "psim_gpu_forces")
        return st, out.reshape(self.n, 3)

    def force_table(self, sorted4, hash_index, cell_start_end):
        """IForce::getAcceleration over the reference's own arrays."""
        s4 = np.ascontiguousarray(sorted4, dtype=np.float64).reshape(-1)
        hi = np.ascontiguousarray(hash_index, dtype=np.int32).reshape(-1)
        cse = np.ascontiguousarray(cell_start_end, dtype=np.int32).reshape(-1)
        out = np.empty(3 * self.n)
        st = self.lib.psim_gpu_force_table(self._ctx, _ptr(s4),
                                           hi.ctypes.data_as(C.POINTER(C.c_int)),
                                           cse.ctypes.data_as(C.POINTER(C.c_int)), _ptr(out))
        if st < 0:
            check(st, "psim_gpu_force_table")
        return st, out.reshape(self.n, 3)

    def pair_stats(self, adjoint: bool = False):
        """Neighbour sets as the pair kernel sees them: per particle (original
        index) the accepted partner count and the partner-hash sum (uint64);
        adjoint=True uses the trajectory adjoint's pair visitor instead of the
        step kernel's."""
        cnt = np.empty(self.n, np.int32)
        hs = np.empty(self.n, np.uint64)
        check(self.lib.psim_gpu_pair_stats(self._ctx, PAIRS_ADJOINT if adjoint else PAIRS_STEP,
                                           cnt.ctypes.data_as(C.POINTER(C.c_int)),
                                           hs.ctypes.data_as(C.POINTER(C.c_ulonglong))),
              "psim_gpu_pair_stats")
        return cnt, hs

    def cells(self):
        n, nc = self.n, self.geometry.ncell
        key = np.empty(n, np.int32)
        cs = np.empty(nc + 1, np.int32)
        ob = np.empty(n, np.int32)
        ip = C.POINTER(C.c_int)
        check(self.lib.psim_gpu_get_cells(self._ctx, key.ctypes.data_as(ip),
                                          cs.ctypes.data_as(ip), ob.ctypes.data_as(ip)),
              "psim_gpu_get_cells")
        return key, cs, ob

    def last_error(self):
        code, i, j = C.c_int(), C.c_int(), C.c_int()
        check(self.lib.psim_gpu_last_error(self._ctx, C.byref(code), C.byref(i), C.byref(j)),
              "psim_gpu_last_error")
        return code.value, i.value, j.value

    def clear_error(self):
        check(self.lib.psim_gpu_clear_error(self._ctx), "psim_gpu_clear_error")

    def sync(self):
        check(self.lib.psim_gpu_sync(self._ctx), "psim_gpu_sync")

    def set_profiling(self, on: bool):
        check(self.lib.psim_gpu_set_profiling(self._ctx, int(on)), "psim_gpu_set_profiling")

    def kernel_times(self, reset=False) -> dict:
        kt = _lib.KernelTimes()
        check(self.lib.psim_gpu_kernel_times(self._ctx, C.byref(kt), int(reset)),
              "psim_gpu_kernel_times")
        return {f: getattr(kt, f) for f, _ in kt._fields_}

    @property
    def stream(self) -> int:
        return self.lib.psim_gpu_stream(self._ctx) or 0

    # -- parameters -----------------------------------------------------------
    def set_forces(self, forces) -> None:
        """Replace the force modules' parameters (same modules or others);
        geometry, integrator and state are kept."""
        if not isinstance(forces, ForceSet):
            forces = ForceSet(list(forces) if isinstance(forces, (list, tuple)) else [forces])
        p = PsimParams.from_buffer_copy(self.params)
        forces.fill(p)
        check(self.lib.psim_gpu_set_force_params(self._ctx, C.byref(p)),
              "psim_gpu_set_force_params")
        self.params, self.forces = p, forces

    # -- analysis -------------------------------------------------------------
    def track_start(self) -> None:
        """Start the displacement tracker at the current positions."""
        check(self.lib.psim_gpu_track_start(self._ctx), "psim_gpu_track_start")

    def observables(self, contact_r: float = 1.2, coordination: bool = False):
        """Temperature, mean displacement, tracked MSD and contact counts of
        the current state (the reference's analysis definitions)."""
        ob = _lib.Observables()
        coord = np.empty(self.n, np.int32) if coordination else None
        check(self.lib.psim_gpu_observables(
            self._ctx, float(contact_r), C.byref(ob),
            None if coord is None else coord.ctypes.data_as(C.POINTER(C.c_int))),
            "psim_gpu_observables")
        out = {f: getattr(ob, f) for f, _ in ob._fields_}
        if coordination:
            out["coordination"] = coord
        return out

    def clusters(self, contact_r: float = 1.2, min_size: int = 5):
        """Connected components of the contact graph: (labels by original
        index, number of clusters of >= min_size members, mean size)."""
        lab = np.empty(self.n, np.int32)
        nc, ms = C.c_int(), C.c_int()
        ip = C.POINTER(C.c_int)
        check(self.lib.psim_gpu_clusters(self._ctx, float(contact_r), int(min_size),
                                         lab.ctypes.data_as(ip), C.byref(nc), C.byref(ms)),
              "psim_gpu_clusters")
        return lab, nc.value, ms.value

    def pair_histogram(self, r_lo: float, width: float, nbins: int, sigma: float = 0.0):
        """Pair-distance histogram over ordered pairs (Gaussian-smoothed when
        sigma > 0)."""
        out = np.empty(int(nbins))
        check(self.lib.psim_gpu_pair_histogram(self._ctx, float(r_lo), float(width),
                                               int(nbins), float(sigma), _ptr(out)),
              "psim_gpu_pair_histogram")
        return out

    def pair_histogram_grad(self, r_lo: float, width: float, nbins: int, sigma: float, w):
        """d(sum_b w_b hist_b)/d positions, AoS3 by original index."""
        wv = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
        if wv.size != nbins:
            raise ValueError("w must have nbins entries")
        out = np.empty(3 * self.n)
        check(self.lib.psim_gpu_pair_histogram_grad(self._ctx, float(r_lo), float(width),
                                                    int(nbins), float(sigma), _ptr(wv),
                                                    _ptr(out)),
              "psim_gpu_pair_histogram_grad")
        return out.reshape(self.n, 3)


ADJOINT_PARAMS = ("ao_well_depth", "cutoff", "lj_well_depth", "yukawa_strength",
                  "debye_length")


class Adjoint:
    """Reverse-mode derivative of ``nsteps`` steps of a GpuSystem (in-kernel
    noise, fixed by the seed and step index) with respect to the pair-force
    parameters, with a checkpoint every ``checkpoint_every`` steps.

    The trajectory starts from the system's state when the object is made
    (or at the last ``set_initial``)."""

    def __init__(self, system: GpuSystem, nsteps: int, checkpoint_every: int = 20):
        self.system, self.nsteps, self.every = system, int(nsteps), int(checkpoint_every)
        self.lib = system.lib
        self._adj = C.c_void_p()
        check(self.lib.psim_adj_create(system._ctx, self.nsteps, self.every,
                                       C.byref(self._adj)), "psim_adj_create")

    def close(self):
        if self._adj:
            self.lib.psim_adj_destroy(self._adj)
            self._adj = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_initial(self) -> None:
        check(self.lib.psim_adj_set_initial(self._adj), "psim_adj_set_initial")

    def set_record_all(self, on: bool) -> None:
        """Record every state in the forward pass (about K (32 N + 4 ncell)
        bytes of device memory) so that the backward pass recomputes nothing;
        off: checkpointed segments, recomputed backwards.  Same gradients."""
        check(self.lib.psim_adj_set_record_all(self._adj, int(bool(on))),
              "psim_adj_set_record_all")

    def set_cutoff_window(self, w: float) -> None:
        """w > 0 adds the pairs that cross the cutoff to dL/dcutoff, linearised
        over [cutoff - w, cutoff + w] (the force jumps there); 0 = pathwise."""
        check(self.lib.psim_adj_set_cutoff_window(self._adj, float(w)),
              "psim_adj_set_cutoff_window")

    def forward(self) -> int:
        """Run the trajectory with the system's current parameters; returns
        the status (0 or a physics error code)."""
        st = self.lib.psim_adj_forward(self._adj)
        if st < 0:
            check(st, "psim_adj_forward")
        return st

    def backward(self, gx, with_gx0: bool = False):
        """gx = dL/dx at the last step (n x 3, original order) -> dict of
        dL/dparameter (and dL/dx at the start)."""
        g = np.ascontiguousarray(gx, dtype=np.float64).reshape(-1)
        if g.size != 3 * self.system.n:
            raise ValueError("gx must be n x 3")
        dth = np.empty(_lib.DP_COUNT)
        gx0 = np.empty(3 * self.system.n) if with_gx0 else None
        check(self.lib.psim_adj_backward(self._adj, _ptr(g), _ptr(dth), _ptr(gx0)),
              "psim_adj_backward")
        out = dict(zip(ADJOINT_PARAMS, dth.tolist()))
        if with_gx0:
            out["gx0"] = gx0.reshape(-1, 3)
        return out


@dataclass
class Replica:
    """One member of a replica ensemble: its geometry and force strengths
    (the integrator and the force modules are shared)."""
    geometry: SystemGeometry
    forces: object


class Ensemble:
    """Independent systems of uneven N, box and force strength advanced
    together on the GPU (one launch per kernel stage for all replicas,
    optionally replayed from a captured hipGraph)."""

    def __init__(self, replicas, integrator: BrownianIntegrator, radius: float = 0.5,
                 kernel_path: str = "sparse"):
        self.lib = _lib.load()
        if not replicas:
            raise ValueError("empty ensemble")
        sets = [r.forces if isinstance(r.forces, ForceSet) else ForceSet(
            list(r.forces) if isinstance(r.forces, (list, tuple)) else [r.forces])
            for r in replicas]
        masks = {fs.mask for fs in sets}
        if len(masks) != 1:
            raise ValueError("all replicas must use the same force modules")
        common = PsimParams()
        common.force_mask = masks.pop()
        common.radius = radius
        common.mass, common.kT, common.gamma, common.dt = (integrator.mass, integrator.kT,
                                                           integrator.gamma, integrator.dt)
        common.vel_freq = integrator.vel_freq
        common.seed = integrator.seed
        if kernel_path not in ("brick", "sparse"):
            raise ValueError("ensemble kernel_path is 'sparse' or 'brick'")
        common.kernel_path = 3 if kernel_path == "brick" else 1
        arr = (_lib.PsimReplica * len(replicas))()
        for k, (rep, fs) in enumerate(zip(replicas, sets)):
            tmp = PsimParams()
            fs.fill(tmp)
            g = rep.geometry
            a = arr[k]
            a.n_particles, a.cells_per_dim, a.box, a.cell_size = g.n, g.cells_per_dim, g.box, \
                g.cell_size
            a.cutoff, a.lj_well_depth, a.yukawa_strength = tmp.cutoff, tmp.lj_well_depth, \
                tmp.yukawa_strength
            a.debye_length, a.ao_well_depth, a.lj_num = tmp.debye_length, tmp.ao_well_depth, \
                tmp.lj_num
        self.replicas = replicas
        self.offsets = np.concatenate([[0], np.cumsum([r.geometry.n for r in replicas])])
        self._ctx = C.c_void_p()
        check(self.lib.psim_ens_create(C.byref(common), arr, len(replicas), C.byref(self._ctx)),
              "psim_ens_create")
        self.n = int(self.offsets[-1])

    def close(self):
        if self._ctx:
            self.lib.psim_ens_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_state(self, pos, vel=None, time=0.0):
        pos = _aos3(pos, self.n, "pos")
        vel = _aos3(vel, self.n, "vel")
        check(self.lib.psim_ens_set_state(self._ctx, _ptr(pos), _ptr(vel), float(time)),
              "psim_ens_set_state")

    def get_state(self) -> dict:
        out = {k: np.empty(3 * self.n) for k in ("pos", "pos0", "vel", "frc")}
        check(self.lib.psim_ens_get_state(self._ctx, _ptr(out["pos"]), _ptr(out["pos0"]),
                                          _ptr(out["vel"]), _ptr(out["frc"])),
              "psim_ens_get_state")
        return {k: v.reshape(self.n, 3) for k, v in out.items()}

    def replica_slice(self, r: int) -> slice:
        return slice(int(self.offsets[r]), int(self.offsets[r + 1]))

    def set_graph(self, steps_per_graph: int):
        check(self.lib.psim_ens_set_graph(self._ctx, int(steps_per_graph)), "psim_ens_set_graph")

    def step(self, nsteps: int = 1):
        check(self.lib.psim_ens_step(self._ctx, int(nsteps)), "psim_ens_step")

    def sync(self):
        check(self.lib.psim_ens_sync(self._ctx), "psim_ens_sync")

    def last_error(self):
        code, i, j = C.c_int(), C.c_int(), C.c_int()
        check(self.lib.psim_ens_last_error(self._ctx, C.byref(code), C.byref(i), C.byref(j)),
              "psim_ens_last_error")
        return code.value, i.value, j.value

    def pair_stats(self):
        """Per-particle accepted partner count and partner-hash sum of every
        replica (concatenated original index), from the current pair kernel."""
        cnt = np.empty(self.n, np.int32)
        hs = np.empty(self.n, np.uint64)
        check(self.lib.psim_ens_pair_stats(self._ctx, cnt.ctypes.data_as(C.POINTER(C.c_int)),
                                           hs.ctypes.data_as(C.POINTER(C.c_ulonglong))),
              "psim_ens_pair_stats")
        return cnt, hs

    def autotune(self, trial_steps: int = 5):
        """For each occupancy class of replicas (particles per cell below 0.25,
        0.5, 1, and above), time the pair-kernel candidates (thread per
        particle with workgroup sizes 64/128/256, bricks of about 96/192/384
        home particles) on that class and keep the fastest; returns (kernel,
        tile, ms per step) of the class with the most particles, kernel
        "sparse" or "brick" (``class_choices`` lists every class).  The timed
        whole steps advance the ensemble; results do not depend on the
        choice."""
        k, t, ms = C.c_int(), C.c_int(), C.c_float()
        check(self.lib.psim_ens_autotune(self._ctx, int(trial_steps), C.byref(k), C.byref(t),
                                         C.byref(ms)), "psim_ens_autotune")
        return ("brick" if k.value == 3 else "sparse"), t.value, ms.value

    @property
    def merged(self) -> bool:
        """True when one pair launch for all replicas beat one per class."""
        st = self.lib.psim_ens_merged(self._ctx)
        if st < 0:
            check(st, "psim_ens_merged")
        return bool(st)

    def class_choices(self):
        """[(replicas, kernel, tile)] for each occupancy class, in order (the
        kernel that runs the class: its own, or the single launch when
        ``merged``)."""
        out = []
        for c in range(_lib.ENS_CLASSES):
            n, k, t = C.c_int(), C.c_int(), C.c_int()
            check(self.lib.psim_ens_class_info(self._ctx, c, C.byref(n), C.byref(k), C.byref(t)),
                  "psim_ens_class_info")
            out.append((n.value, "brick" if k.value == 3 else "sparse", t.value))
        return out
