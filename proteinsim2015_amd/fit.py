"""Structure fitting of AO depletion parameters with the trajectory adjoint.

The observable is the Gaussian-smoothed pair-distance histogram of the last
configuration of a K-step trajectory, normalised to g(r); the loss is the
squared distance to a target g(r).  ``TrajectoryGofR`` is a
``torch.autograd.Function`` whose forward runs the trajectory on the GPU
(through the C ABI) and whose backward runs the hand-written adjoint:

    theta = (U, q)  ->  AO wellDepth phi_p = 2 q U / (3 + 2 q), cutOff = 1 + q
                    ->  x_K (K steps, fixed in-kernel noise)
                    ->  h_b (smoothed histogram)  ->  g_b = h_b / (N rho V_b)

U is the AO contact depth in kT (U_AO(1) = -phi_p (3 + 2q) / (2q),
AOPotential.cpp:50-55 with sigma = 1) and q the polymer/colloid size ratio.

The q derivative.  The reference's AO force jumps at the cutoff c = 1 + q
(fNet(c) = -36 c^-38, -0.96 kT/sigma at c = 1.1), so every change of q moves
pairs across the cutoff, and in a system of this size the crossings
cascade: between trajectories at q +- h the largest position change is
4e-3 at h = 1e-5 and 0.12 at h = 3e-4 (N = 32768, 200 steps), and the
secant of one noise realisation is no derivative (from -4.2 to -7.3 for
h = 1e-7 ... 1e-2).  The adjoint therefore estimates the derivative of the
loss averaged over the noise: its pathwise part plus the crossing term of a
cutoff window w (Adjoint.set_cutoff_window; pairs with |r - c| < w weighted
by fNet / (2 w)).  Over six noise seeds the windowed estimate (w = 3e-3)
differs from the secant (mean of h = 1e-3, 3e-3, 1e-2) by +0.06 +- 0.57 on
average, the pathwise one by -0.80 +- 0.55 (secants about -6.5).
AOFitProblem uses w = 3e-3 unless told otherwise; cutoff_window=0 gives the
pathwise derivative.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from . import engine as E


@dataclass
class GofRSpec:
    r_lo: float = 0.8
    width: float = 0.02
    nbins: int = 110
    sigma: float = 0.02

    def shell_volumes(self) -> np.ndarray:
        e = self.r_lo + self.width * np.arange(self.nbins + 1)
        return 4.0 / 3.0 * math.pi * (e[1:] ** 3 - e[:-1] ** 3)

    @property
    def centres(self) -> np.ndarray:
        return self.r_lo + self.width * (np.arange(self.nbins) + 0.5)


def ao_params(u_contact: float, q: float):
    """(phi_p, cutoff) and d(phi_p)/d(U, q)."""
    phi = 2.0 * q * u_contact / (3.0 + 2.0 * q)
    dphi_du = 2.0 * q / (3.0 + 2.0 * q)
    dphi_dq = 6.0 * u_contact / (3.0 + 2.0 * q) ** 2
    return phi, 1.0 + q, dphi_du, dphi_dq


class AOFitProblem:
    """One GPU system, its trajectory adjoint and the g(r) observable."""

    def __init__(self, system: E.GpuSystem, nsteps: int, checkpoint_every: int = 20,
                 spec: GofRSpec | None = None, kT: float = 1.0,
                 cutoff_window: float = 3e-3, record_all: bool = True):
        """record_all: keep every state of the trajectory from the forward
        pass (about nsteps (32 N + 4 ncell) bytes of device memory, 0.26 GB at
        N = 32768 and 200 steps) so that the backward pass recomputes nothing;
        False: checkpoint_every-step segments, recomputed backwards."""
        self.system = system
        self.spec = spec or GofRSpec()
        self.kT = kT
        self.cutoff_window = float(cutoff_window)
        self.adjoint = E.Adjoint(system, nsteps, checkpoint_every)
        self.adjoint.set_cutoff_window(self.cutoff_window)
        self.adjoint.set_record_all(record_all)
        n, L = system.n, system.geometry.box
        self.norm = n * (n / L ** 3) * self.spec.shell_volumes()
        self._theta = None

    def close(self):
        self.adjoint.close()

    def set_theta(self, u_contact: float, q: float) -> None:
        phi, cut, _, _ = ao_params(u_contact, q)
        if cut + self.cutoff_window > self.system.geometry.cell_size:
            raise ValueError(f"cutoff {cut} + window {self.cutoff_window} exceeds the cell edge "
                             f"{self.system.geometry.cell_size}")
        self.system.set_forces(E.AOPotential(kT=self.kT, well_depth=phi, cutoff=cut))
        self._theta = (float(u_contact), float(q))

    def forward(self, u_contact: float, q: float) -> np.ndarray:
        """g(r) of the last configuration of the trajectory at (U, q)."""
        self.set_theta(u_contact, q)
        st = self.adjoint.forward()
        if st:
            raise _lib.PsimError(st, "trajectory")
        s = self.spec
        h = self.system.pair_histogram(s.r_lo, s.width, s.nbins, s.sigma)
        return h / self.norm

    def backward(self, grad_g: np.ndarray) -> np.ndarray:
        """dL/d(U, q) given dL/dg of the last forward."""
        s = self.spec
        w = np.asarray(grad_g, dtype=np.float64) / self.norm
        gx = self.system.pair_histogram_grad(s.r_lo, s.width, s.nbins, s.sigma, w)
        d = self.adjoint.backward(gx)
        u, q = self._theta
        _, _, dphi_du, dphi_dq = ao_params(u, q)
        return np.array([d["ao_well_depth"] * dphi_du,
                         d["ao_well_depth"] * dphi_dq + d["cutoff"]])


class TrajectoryGofR(torch.autograd.Function):
    """theta = (U, q) [float64 tensor] -> g(r) [nbins], differentiable."""

    @staticmethod
    def forward(ctx, theta, problem: AOFitProblem):
        u, q = (float(v) for v in theta.detach().cpu().tolist())
        g = problem.forward(u, q)
        ctx.problem = problem
        return torch.from_numpy(g).to(theta.device)

    @staticmethod
    def backward(ctx, grad_g):
        gt = ctx.problem.backward(grad_g.detach().cpu().numpy())
        return torch.from_numpy(gt).to(grad_g.device), None


def gofr_loss(problem: AOFitProblem, theta: torch.Tensor, target: torch.Tensor):
    g = TrajectoryGofR.apply(theta, problem)
    return ((g - target) ** 2).sum()


def c5_problem(n: int = 32768, phi: float = 0.15, nsteps: int = 200, every: int = 20,
               seed: int = 1234, noise_seed: int = 90210, dt: float = 1e-3,
               spec: GofRSpec | None = None, max_q: float = 0.2,
               cutoff_window: float = 3e-3, record_all: bool = True):
    """The differentiable AO fit workload: N colloids at volume fraction phi,
    seeded RSA start, K Brownian steps with fixed in-kernel noise (the cell
    edge holds the cutoff 1 + max_q plus the cutoff window)."""
    geom = E.SystemGeometry.exact(n, phi, 1.0 + max_q + cutoff_window)
    integ = E.BrownianIntegrator(kT=1.0, mass=1.0, gamma=1.0, dt=dt, vel_freq=1000,
                                 seed=noise_seed)
    phi_p, cut, _, _ = ao_params(5.0, 0.12)
    sysg = E.GpuSystem(geom, E.AOPotential(kT=1.0, well_depth=phi_p, cutoff=cut), integ)
    pos = E.rsa_positions(n, geom.box, seed)
    vel = np.random.default_rng(seed + 1).standard_normal((n, 3))
    sysg.set_state(pos, vel=vel)
    return AOFitProblem(sysg, nsteps, every, spec, cutoff_window=cutoff_window,
                        record_all=record_all)


def fit(problem: AOFitProblem, theta0, target_g: np.ndarray, iters: int = 20,
        lr: float = 0.05, scale=(1.0, 0.1)):
    """Adam on (U, q) against a target g(r), in coordinates (U, q) / scale so
    that one learning rate suits both; returns [(theta, loss, dL/dtheta)].
    The q component is the noise-averaged estimate described at the top of
    this module (one realisation's loss is not differentiable in q)."""
    sc = torch.tensor(scale, dtype=torch.float64)
    z = (torch.tensor(theta0, dtype=torch.float64) / sc).requires_grad_(True)
    target = torch.from_numpy(np.asarray(target_g, dtype=np.float64))
    opt = torch.optim.Adam([z], lr=lr)
    hist = []
    for _ in range(iters):
        opt.zero_grad()
        theta = z * sc
        loss = gofr_loss(problem, theta, target)
        loss.backward()
        hist.append((theta.detach().tolist(), loss.item(), (z.grad / sc).tolist()))
        opt.step()
        with torch.no_grad():
            z[0].clamp_(0.5 / scale[0], 12.0 / scale[0])
            z[1].clamp_(0.02 / scale[1], 0.2 / scale[1])
    return hist
