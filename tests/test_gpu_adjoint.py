"""Adjoint of the shallow-water step on the GPU (claw2d_adjoint_step) and the checkpointed
frozen-dt gradient (clawpack_amd.adjoint.evolve_frozen).

Checks: one reverse step against float64 torch autograd of the torch restatement of the same
step (oracle/torch_step.py) over BCs, limiters, orders and transverse modes, including fast flows
where the entropy fix is active; the checkpointed multi-step gradient against torch autograd
through the same frozen-dt steps; and the inversion workload (512^2, periodic, 400 steps, 64
gauges sampled every 10 steps) against central finite differences along 16 seeded directions.
"""
import itertools
import sys

import numpy as np
import pytest
import torch

from clawpack_amd import _lib, problems
from clawpack_amd.adjoint import adjoint_step, cfl_dt_sequence, evolve_frozen
from clawpack_amd.device import make_desc
from oracle import torch_step as ts

pytestmark = pytest.mark.gpu


def desc_for(mx, my, dx, dy, bc_lower=(1, 1), bc_upper=(1, 1), order=2, tw=2, lim=4):
    return make_desc(0, _lib.FP64, mx, my, dx, dy, 3, 3, order=order, transverse_waves=tw,
                     mthlim=(lim,) * 3, bc_lower=bc_lower, bc_upper=bc_upper, grav=1.0)


def torch_vjp(q_planes, lam_planes, dt, dx, dy, bc_lower, bc_upper, order, tw, lim):
    """(d step / d q)^T lam with the torch restatement (pyclaw layout inside)."""
    q = torch.tensor(np.ascontiguousarray(q_planes.transpose(0, 2, 1)), requires_grad=True)
    qn, _ = ts.step(q, dt, dx, dy, "swe", 1.0, mthlim=(lim,) * 3, bc_lower=bc_lower, bc_upper=bc_upper,
                    order=order, transverse_waves=tw)
    (qn * torch.tensor(np.ascontiguousarray(lam_planes.transpose(0, 2, 1)))).sum().backward()
    return q.grad.numpy().transpose(0, 2, 1)


CASES = list(itertools.product(
    [(37, 29), (130, 71)],
    [((1, 1), (1, 1)), ((2, 2), (2, 2)), ((3, 3), (3, 3)), ((1, 2), (3, 2))],
    [(2, 2, 4), (2, 1, 1), (1, 0, 0), (2, 2, 2), (2, 0, 3)],
))


@pytest.mark.parametrize("shape,bc,method", CASES)
def test_adjoint_step_matches_torch_autograd(shape, bc, method):
    mx, my = shape
    order, tw, lim = method
    print("case", shape, bc, method, flush=True)
    q = problems.random_smooth_swe(mx, my, seed=mx + 3 * my, amp=0.8)
    planes = np.ascontiguousarray(q.transpose(0, 2, 1))
    lam = np.random.default_rng(5).standard_normal(planes.shape)
    dx, dy = 1.0 / mx, 1.2 / my
    dt = 0.2 * min(dx, dy)
    d = desc_for(mx, my, dx, dy, bc[0], bc[1], order, tw, lim)
    got = adjoint_step(d, torch.tensor(planes, device="cuda"), torch.tensor(lam, device="cuda"), dt).cpu().numpy()
    ref = torch_vjp(planes, lam, dt, dx, dy, bc[0], bc[1], order, tw, lim)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err


def test_adjoint_step_in_place_and_transonic():
    mx, my = 64, 48
    q = problems.random_smooth_swe(mx, my, seed=11, amp=2.0)  # negative Roe middle depths occur
    planes = np.ascontiguousarray(q.transpose(0, 2, 1))
    lam = np.random.default_rng(6).standard_normal(planes.shape)
    dx, dy = 1.0 / mx, 1.3 / my
    dt = 0.05 * min(dx, dy)
    d = desc_for(mx, my, dx, dy, (2, 2), (2, 2))
    qd = torch.tensor(planes, device="cuda")
    out = adjoint_step(d, qd, torch.tensor(lam, device="cuda"), dt).cpu().numpy()
    lam_d = torch.tensor(lam, device="cuda")
    adjoint_step(d, qd, lam_d, dt, lam_out=lam_d)
    assert np.array_equal(out, lam_d.cpu().numpy())
    ref = torch_vjp(planes, lam, dt, dx, dy, (2, 2), (2, 2), 2, 2, 4)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()


def loss_fn(eta, desc, dts, gauges, obs, sample_every, checkpoint_every):
    zeros = torch.zeros_like(eta)
    q0 = torch.stack([1.0 + eta, zeros, zeros])
    samples, _ = evolve_frozen(q0, desc, dts, gauges, sample_every, checkpoint_every)
    return ((samples[:, 0, :] - obs) ** 2).sum()


def test_checkpointed_gradient_matches_torch_autograd():
    mx, my, nsteps = 40, 32, 30
    eta_true, gauges, dx, dy = problems.inversion_workload(mx, my, seed=3, n_gauges=12)
    d = desc_for(mx, my, dx, dy, (2, 2), (2, 2))
    q_true = torch.tensor(np.stack([1.0 + eta_true, 0 * eta_true, 0 * eta_true]), device="cuda")
    dts = cfl_dt_sequence(d, q_true, nsteps)
    g = torch.tensor(gauges)
    with torch.no_grad():
        obs = evolve_frozen(q_true, d, dts, g, 5, 7)[0][:, 0, :]
    eta = torch.tensor(0.5 * eta_true + 0.01 * np.random.default_rng(2).standard_normal(eta_true.shape),
                       device="cuda", requires_grad=True)
    loss = loss_fn(eta, d, dts, g, obs, 5, 7)
    loss.backward()
    got = eta.grad.cpu().numpy()
    # torch autograd through the torch restatement (CPU, float64), same frozen dt sequence
    e = torch.tensor(eta.detach().cpu().numpy(), requires_grad=True)
    q = torch.stack([1.0 + e, torch.zeros_like(e), torch.zeros_like(e)]).permute(0, 2, 1)
    o = obs.cpu()
    total = 0.0
    for k, dt in enumerate(dts):
        q, _ = ts.step(q, float(dt), dx, dy, "swe", 1.0, mthlim=(4, 4, 4), bc_lower=(2, 2), bc_upper=(2, 2))
        if (k + 1) % 5 == 0:
            h = q[0].permute(1, 0).reshape(-1)[g]
            total = total + ((h - o[(k + 1) // 5 - 1]) ** 2).sum()
    total.backward()
    ref = e.grad.numpy()
    assert abs(float(loss) - float(total)) <= 1e-12 * abs(float(total))
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


def inversion_setup(n, nsteps, lim):
    eta_true, gauges, dx, dy = problems.inversion_workload(n, n, seed=4)
    d = desc_for(n, n, dx, dy, (2, 2), (2, 2), lim=lim)
    q_true = torch.tensor(np.stack([1.0 + eta_true, 0 * eta_true, 0 * eta_true]), device="cuda")
    dts = cfl_dt_sequence(d, q_true, nsteps)
    g = torch.tensor(gauges)
    with torch.no_grad():
        obs = evolve_frozen(q_true, d, dts, g, 10, 20)[0][:, 0, :]
    # a generic point: away from eta0 = 0 (rest state, where the limiter is not differentiable)
    rng = np.random.default_rng(9)
    eta0 = 0.5 * eta_true + 1e-3 * rng.standard_normal(eta_true.shape)
    return eta_true, gauges, dx, dy, d, dts, g, obs, eta0, rng


def test_inversion_gradient_is_exact_vs_torch_autograd():
    """128^2, 100 frozen-dt steps, MC, gauges every 10 steps, checkpoints every 20: the adjoint
    gradient equals float64 torch autograd of the torch restatement of the same steps."""
    n, nsteps = 128, 100
    eta_true, gauges, dx, dy, d, dts, g, obs, eta0, _ = inversion_setup(n, nsteps, 4)
    eta = torch.tensor(eta0, device="cuda", requires_grad=True)
    loss = loss_fn(eta, d, dts, g, obs, 10, 20)
    loss.backward()
    got = eta.grad.cpu().numpy()
    e = torch.tensor(eta0, requires_grad=True)
    q = torch.stack([1.0 + e, torch.zeros_like(e), torch.zeros_like(e)]).permute(0, 2, 1)
    o = obs.cpu()
    total = 0.0
    for k, dt in enumerate(dts):
        q, _ = ts.step(q, float(dt), dx, dy, "swe", 1.0, mthlim=(4, 4, 4), bc_lower=(2, 2), bc_upper=(2, 2))
        if (k + 1) % 10 == 0:
            total = total + ((q[0].permute(1, 0).reshape(-1)[g] - o[(k + 1) // 10 - 1]) ** 2).sum()
    total.backward()
    ref = e.grad.numpy()
    assert abs(float(loss) - float(total)) <= 1e-12 * abs(float(total))
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def fd_check(lim, ndirs=16, eps=1e-6):
    mx = 512
    eta_true, gauges, dx, dy, d, dts, g, obs, eta0, rng = inversion_setup(mx, 400, lim)
    eta = torch.tensor(eta0, device="cuda", requires_grad=True)
    loss = loss_fn(eta, d, dts, g, obs, 10, 20)
    loss.backward()
    grad = eta.grad.cpu().numpy()
    gnorm = float(np.linalg.norm(grad))
    rel_dir, rel_norm = [], []
    for k in range(ndirs):
        v = rng.standard_normal(eta0.shape)
        v /= np.linalg.norm(v)
        with torch.no_grad():
            lp = float(loss_fn(torch.tensor(eta0 + eps * v, device="cuda"), d, dts, g, obs, 10, 20))
            lm = float(loss_fn(torch.tensor(eta0 - eps * v, device="cuda"), d, dts, g, obs, 10, 20))
        fd = (lp - lm) / (2 * eps)
        ad = float((grad * v).sum())
        rel_dir.append(abs(fd - ad) / abs(ad))
        rel_norm.append(abs(fd - ad) / gnorm)
        print("lim %d direction %2d fd %+.12e adjoint %+.12e |fd-ad|/|ad| %.2e |fd-ad|/|grad| %.2e"
              % (lim, k, fd, ad, rel_dir[-1], rel_norm[-1]), flush=True)
    return np.array(rel_dir), np.array(rel_norm)


def test_inversion_gradient_matches_central_differences():
    """The inversion workload (512^2 periodic, MC, 400 frozen-dt steps, checkpoints every 20, 64
    gauges every 10 steps): central differences, eps = 1e-6, 16 seeded unit directions.  The step
    is only piecewise smooth (limiter cases, upwind and transverse sign switches), and a random
    unit direction over 262144 cells has a directional derivative ~|grad|/512, so the error is
    measured against |grad|: |fd - adjoint| <= 1e-6 |grad| in every direction."""
    rel_dir, rel_norm = fd_check(4)
    assert rel_norm.max() <= 1e-6, rel_norm.max()
    assert np.median(rel_dir) <= 1e-3, np.median(rel_dir)


def test_inversion_gradient_central_differences_unlimited():
    """Same workload with the unlimited second-order method (no limiter kinks): the remaining
    difference is round-off in the differences, |fd - adjoint| <= 1e-8 |grad|."""
    rel_dir, rel_norm = fd_check(0, ndirs=8)
    assert rel_norm.max() <= 1e-8, rel_norm.max()


def test_gradient_independent_of_checkpoint_spacing_and_repeat_backward():
    """The backward recomputes each segment from its checkpoint on a second stream while the
    previous segment is reversed; the states it recomputes are the forward's bit for bit, so the
    gradient may not depend on the spacing (one segment per step, segments that do not divide the
    step count, one segment for the whole run), and a second backward through the same graph,
    which has no kept last segment and recomputes it too, gives the same gradient."""
    mx, my, nsteps = 96, 80, 45
    eta_true, gauges, dx, dy = problems.inversion_workload(mx, my, seed=5, n_gauges=16)
    d = desc_for(mx, my, dx, dy, (2, 2), (2, 2))
    q_true = torch.tensor(np.stack([1.0 + eta_true, 0 * eta_true, 0 * eta_true]), device="cuda")
    dts = cfl_dt_sequence(d, q_true, nsteps)
    g = torch.tensor(gauges)
    with torch.no_grad():
        obs = evolve_frozen(q_true, d, dts, g, 5, 7)[0][:, 0, :]
    eta0 = 0.5 * eta_true + 0.01 * np.random.default_rng(4).standard_normal(eta_true.shape)
    grads = {}
    for K in (1, 3, 7, 30, 45, 100):
        print("case checkpoint spacing", K, flush=True)
        eta = torch.tensor(eta0, device="cuda", requires_grad=True)
        loss = loss_fn(eta, d, dts, g, obs, 5, K)
        loss.backward(retain_graph=True)
        first = eta.grad.clone()
        eta.grad = None
        loss.backward()
        assert torch.equal(eta.grad, first), K
        grads[K] = first.cpu().numpy()
    for K, v in grads.items():
        assert np.array_equal(v, grads[7]), K
    assert np.abs(grads[7]).max() > 0
