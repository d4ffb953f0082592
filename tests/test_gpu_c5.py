"""The differentiable AO fit at full size (needs a GPU): N = 32768, phi = 0.15,
K = 200 steps, checkpoints every 20, smoothed g(r) loss against a target
made at U = 5 kT, q = 0.12, gradient at U = 4 kT, q = 0.10.

* the GPU adjoint's dL/d(phi_p, cutOff), with the fit's cutoff window, against
  the oracle's FP64 CPU adjoint with the same window, fed the same dL/dx_K
  (1e-7 relative; the two trajectories differ by rounding only);
* dL/dU through the torch.autograd.Function against central differences
  of the loss, relative step 1e-6 (1e-4 relative bar);
* dL/dq against the secant averaged over noise seeds.  The AO force of the
  reference jumps by 36 c^-38 (0.96 kT/sigma at c = 1.1) at the cutoff; with
  about 7e4 pairs per unit distance at r = c in each of the 200 steps, every
  step in q moves pairs across it, and the crossings cascade (the ± runs
  differ by up to 0.1 in position at h = 3e-4), so one realisation's secant
  is no derivative.  Over six noise seeds, the windowed adjoint minus the
  secant (mean over h = 1e-3, 3e-3, 1e-2) must average zero within two
  standard errors.
"""
import numpy as np
import pytest
import torch

import proteinsim2015_amd as ps
from proteinsim2015_amd import fit
from oracle.oracle import PO_AO, ForceParams

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    print("c5: setup", flush=True)
    prob = fit.c5_problem()
    target = prob.forward(5.0, 0.12)
    yield prob, target
    prob.close()
    prob.system.close()


def test_c5_gpu_adjoint_matches_cpu_fp64_adjoint(problem, oracle):
    prob, target = problem
    u0, q0 = 4.0, 0.10
    g = prob.forward(u0, q0)
    s = prob.spec
    gx = prob.system.pair_histogram_grad(s.r_lo, s.width, s.nbins, s.sigma,
                                         2.0 * (g - target) / prob.norm)
    d = prob.adjoint.backward(gx)
    xk = prob.system.get_state()["pos"]
    # the same start as fit.c5_problem
    geom, integ = prob.system.geometry, prob.system.integrator
    pos = ps.rsa_positions(geom.n, geom.box, 1234)
    vel = np.random.default_rng(1235).standard_normal((geom.n, 3))
    L = geom.box
    p0 = np.where(np.abs(pos) > L / 2, np.where(pos < 0, -L, L), 0.0)
    phi_p, cut, _, _ = fit.ao_params(u0, q0)
    fp = ForceParams.make(PO_AO, cut, kT=1.0, ao_well_depth=phi_p)
    c = oracle.gb_coeffs(integ.kT, integ.mass, integ.gamma, integ.dt)
    print("c5: CPU adjoint", flush=True)
    st, dth, _, oxk = oracle.adjoint(fp, c, dict(pos=pos, pos0=p0, vel=vel), 200, L,
                                     geom.cells_per_dim, geom.cell_size, gx,
                                     vel_freq=integ.vel_freq, seed=integ.seed,
                                     cut_window=prob.cutoff_window)
    assert st == 0
    assert np.abs(xk - oxk).max() < 1e-9
    assert d["ao_well_depth"] == pytest.approx(dth[0], rel=1e-7)
    assert d["cutoff"] == pytest.approx(dth[1], rel=1e-7)


def test_c5_gradient_matches_central_differences(problem):
    prob, target = problem
    tgt = torch.from_numpy(target)
    theta = torch.tensor([4.0, 0.10], dtype=torch.float64, requires_grad=True)
    loss = fit.gofr_loss(prob, theta, tgt)
    loss.backward()
    grad = theta.grad.numpy()
    assert np.all(np.isfinite(grad)) and float(loss.detach()) > 0

    def lossat(u, q):
        return float(np.sum((prob.forward(u, q) - target) ** 2))

    h = 1e-6 * 4.0
    fd = (lossat(4.0 + h, 0.10) - lossat(4.0 - h, 0.10)) / (2 * h)
    print(f"c5: dL/dU adjoint {grad[0]:.9g} fd {fd:.9g}", flush=True)
    assert grad[0] == pytest.approx(fd, rel=1e-4)


def test_c5_q_gradient_matches_the_seed_averaged_secant():
    u0, q0 = 4.0, 0.10
    dw, dp = [], []
    for ns in range(6):
        print(f"c5: noise seed {1000 + ns}", flush=True)
        prob = fit.c5_problem(noise_seed=1000 + ns)
        try:
            target = prob.forward(5.0, 0.12)

            def lossat(u, q):
                return float(np.sum((prob.forward(u, q) - target) ** 2))

            grads = []
            for win in (prob.cutoff_window, 0.0):
                prob.adjoint.set_cutoff_window(win)
                g = prob.forward(u0, q0)
                grads.append(prob.backward(2.0 * (g - target))[1])
            sec = np.mean([(lossat(u0, q0 + h) - lossat(u0, q0 - h)) / (2 * h)
                           for h in (1e-3, 3e-3, 1e-2)])
        finally:
            prob.close()
            prob.system.close()
        dw.append(grads[0] - sec)
        dp.append(grads[1] - sec)
    dw, dp = np.array(dw), np.array(dp)
    sem = dw.std(ddof=1) / np.sqrt(len(dw))
    print(f"c5: dL/dq windowed - secant {dw.mean():.3g} +- {sem:.3g}; pathwise - secant "
          f"{dp.mean():.3g} +- {dp.std(ddof=1) / np.sqrt(len(dp)):.3g}", flush=True)
    assert abs(dw.mean()) <= 2 * sem


def test_fit_step_reduces_the_loss(problem):
    """One Adam step along the adjoint gradient lowers the loss."""
    prob, target = problem
    hist = fit.fit(prob, [4.0, 0.10], target, iters=3, lr=0.05)
    assert hist[-1][1] < hist[0][1]
