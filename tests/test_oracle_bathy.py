"""Pins of the oracle's shallow-water-over-bathymetry f-wave solver (shallow_bathymetry_fwave_2D,
upstream riemann/src/rpn2_shallow_bathymetry_fwave.f90, absent from the reference tree) and of
the f-wave form of flux2 (solver.fwave = True).

What fixes the restatement, independently of its own formulas:
  * the f-waves of an interface sum to the flux jump plus the bathymetry source term
    f(q_r) - f(q_l) + (0, g (h_l + h_r)/2 (b_r - b_l), 0), and each lies along its Roe
    eigenvector (1, s, v) or (0, 0, 1);
  * on a flat bottom the f-waves are s_p times the Roe waves of shallow_roe_with_efix (Roe
    property), so without a limiter (and away from entropy-fix branches) a whole step equals the
    Roe step, which only holds with sign(s) (not |s|) in the f-wave second-order correction;
  * a lake at rest (u = v = 0, h + b constant) stays at rest to round-off (well balancing), for
    the unsplit and the split step;
  * total depth is conserved with periodic boundaries (momentum is not: the bathymetry source);
  * the C oracle equals the independent whole-array torch restatement over BCs (aux BCs
    different from the q BCs included), limiters, transverse modes and dry cells.
"""
import itertools

import numpy as np
import pytest
import torch

from oracle import corc
from oracle import torch_step as ts

G = 9.81
TOL = 1e-3


def bathy_desc(mx, my, dx, dy, order=2, tw=2, lim=4, grav=G, tol=TOL):
    return corc.make_desc("bathy", mx, my, dx, dy, grav=grav, order=order, transverse_waves=tw,
                          mthlim=(lim,) * 3, dry_tolerance=tol)


def swe_desc(mx, my, dx, dy, order=2, tw=2, lim=4, grav=G):
    return corc.make_desc("swe", mx, my, dx, dy, grav=grav, order=order, transverse_waves=tw,
                          mthlim=(lim,) * 3)


def random_slice(rng, L, hmin=0.5, hmax=2.0, umax=0.5):
    h = rng.uniform(hmin, hmax, L)
    return np.stack([h, h * rng.uniform(-umax, umax, L), h * rng.uniform(-umax, umax, L)])


def flux(q, mu, mv, grav):
    h, hu, hv = q[0], q[mu], q[mv]
    f = np.zeros_like(q)
    f[0] = hu
    f[mu] = hu * hu / h + 0.5 * grav * h * h
    f[mv] = hu * hv / h
    return f


@pytest.mark.parametrize("ixy", [1, 2])
def test_fwaves_sum_to_flux_jump_plus_source(ixy):
    rng = np.random.default_rng(10 + ixy)
    n, g = 40, 2
    L = n + 2 * g
    q = random_slice(rng, L)
    b = rng.uniform(-0.4, 0.3, L)[None, :]
    d = bathy_desc(n, n, 0.1, 0.1)
    wave, s, amdq, apdq = corc.rpn2(d, ixy, q, b)
    mu, mv = (1, 2) if ixy == 1 else (2, 1)
    f = flux(q, mu, mv, G)
    expect = f[:, 1:] - f[:, :-1]
    expect[mu] += G * 0.5 * (q[0, :-1] + q[0, 1:]) * (b[0, 1:] - b[0, :-1])
    scale = np.abs(expect).max()
    assert np.abs(wave.sum(axis=1)[:, 1:] - expect).max() <= 1e-13 * scale
    assert np.abs((amdq + apdq)[:, 1:] - expect).max() <= 1e-13 * scale
    # each f-wave along its eigenvector: Z1, Z3 ~ (1, s, v) with the same Roe v, Z2 ~ (0, 0, 1)
    for p in (0, 2):
        Z = wave[:, p, 1:]
        assert np.abs(Z[mu] - Z[0] * s[p, 1:]).max() <= 1e-13 * scale
    assert np.abs(wave[0, 1, 1:]).max() == 0.0 and np.abs(wave[mu, 1, 1:]).max() == 0.0
    Z1, Z3 = wave[:, 0, 1:], wave[:, 2, 1:]
    assert np.abs(Z1[mv] * Z3[0] - Z3[mv] * Z1[0]).max() <= 1e-12 * scale * scale
    # amdq takes the f-waves of negative speed, apdq those of positive speed
    neg = (s[:, 1:] < 0)[None]
    assert np.array_equal(amdq[:, 1:], np.where(neg, wave[:, :, 1:], 0.0).sum(axis=1))
    assert np.all((s[:, 1:] < 0).sum(axis=0) >= 1) and np.all((s[:, 1:] > 0).sum(axis=0) >= 1)


def test_flat_bottom_fwaves_are_speed_times_roe_waves():
    rng = np.random.default_rng(3)
    n, g = 60, 2
    L = n + 2 * g
    q = random_slice(rng, L, umax=0.3)  # subsonic: no entropy-fix branch is taken
    b = np.full((1, L), -0.7)
    for ixy in (1, 2):
        fw, fs, famdq, fapdq = corc.rpn2(bathy_desc(n, n, 0.1, 0.1), ixy, q, b)
        rw, rs, ramdq, rapdq = corc.rpn2(swe_desc(n, n, 0.1, 0.1), ixy, q)
        sl = slice(1, None)
        assert np.abs(fs[:, sl] - rs[:, sl]).max() <= 1e-14 * np.abs(rs).max()
        expect = rw[:, :, sl] * rs[None, :, sl]
        sc = np.abs(expect).max()
        assert np.abs(fw[:, :, sl] - expect).max() <= 1e-13 * sc
        assert np.abs(famdq[:, sl] - ramdq[:, sl]).max() <= 1e-13 * sc
        assert np.abs(fapdq[:, sl] - rapdq[:, sl]).max() <= 1e-13 * sc


def smooth_swe(mx, my, seed, umax=0.3):
    rng = np.random.default_rng(seed)
    x = (np.arange(mx) + 0.5) / mx
    y = (np.arange(my) + 0.5) / my
    X, Y = np.meshgrid(x, y, indexing="ij")
    h = 1.0 + 0.2 * np.sin(2 * np.pi * (X + rng.uniform())) * np.cos(2 * np.pi * (Y + rng.uniform()))
    u = umax * np.cos(2 * np.pi * (X + Y + rng.uniform()))
    v = umax * np.sin(2 * np.pi * (X - Y + rng.uniform()))
    return np.asfortranarray(np.stack([h, h * u, h * v]))


def bump(mx, my, seed, amp=0.4):
    rng = np.random.default_rng(seed)
    x = (np.arange(mx) + 0.5) / mx
    y = (np.arange(my) + 0.5) / my
    X, Y = np.meshgrid(x, y, indexing="ij")
    x0, y0 = rng.uniform(0.3, 0.7, 2)
    return np.asfortranarray((amp * np.exp(-40 * ((X - x0) ** 2 + (Y - y0) ** 2)) - 0.5
                              + 0.05 * np.sin(2 * np.pi * (X + 2 * Y)))[None])


@pytest.mark.parametrize("order,tw", [(1, 0), (1, 2), (2, 0), (2, 1), (2, 2)])
def test_flat_bottom_step_equals_roe_step_unlimited(order, tw):
    mx, my = 32, 24
    q = smooth_swe(mx, my, seed=order * 7 + tw)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.3 * min(dx, dy) / (np.sqrt(G * 1.3) + 0.5)
    c = corc.make_ctrl(bc_lower=(1, 3), bc_upper=(1, 3))
    b = np.full((1, mx, my), 0.25, order="F")
    qb, cb = corc.step_hyperbolic(bathy_desc(mx, my, dx, dy, order, tw, lim=0), c, q, dt, aux=b)
    qr, cr = corc.step_hyperbolic(swe_desc(mx, my, dx, dy, order, tw, lim=0), c, q, dt)
    assert abs(cb - cr) <= 1e-14 * cr
    assert np.abs(qb - qr).max() <= 1e-13 * np.abs(qr).max()


def test_fwave_correction_uses_sign_of_speed():
    """A planted |s| (the wave form) in the f-wave correction breaks the flat-bottom equality:
    the unlimited second-order step differs from the Roe step by O(dt/dx * |s|) instead of
    round-off, so the equality test above pins sign(s)."""
    mx, my = 32, 24
    q = smooth_swe(mx, my, seed=5)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.3 * min(dx, dy) / (np.sqrt(G * 1.3) + 0.5)
    c = corc.make_ctrl()
    b = np.zeros((1, mx, my), order="F")
    d = bathy_desc(mx, my, dx, dy, 2, 2, lim=0)
    d.fwave = 0  # f-waves treated as waves: |s| Z instead of sign(s) Z
    qb, _ = corc.step_hyperbolic(d, c, q, dt, aux=b)
    qr, _ = corc.step_hyperbolic(swe_desc(mx, my, dx, dy, 2, 2, lim=0), c, q, dt)
    assert np.abs(qb - qr).max() > 1e-6 * np.abs(qr).max()


@pytest.mark.parametrize("split,lim,bc", [(False, 4, 3), (False, 1, 1), (True, 4, 3), (True, 2, 2),
                                          (False, 3, 2)])
def test_lake_at_rest_is_preserved(split, lim, bc):
    mx, my = 40, 36
    b = bump(mx, my, seed=lim + 10 * bc)
    eta = 0.0
    h = eta - b[0]
    assert h.min() > 0.05
    q = np.asfortranarray(np.stack([h, 0 * h, 0 * h]))
    dx, dy = 1.0 / mx, 1.0 / my
    d = bathy_desc(mx, my, dx, dy, 2, 2, lim=lim)
    c = corc.make_ctrl(bc_lower=(bc, bc), bc_upper=(bc, bc), dimensional_split=split, dt_initial=0.01)
    qn, st = corc.evolve(d, c, q, nsteps=60, aux=b)
    assert st["numsteps"] == 60
    assert np.abs(qn[0] + b[0] - eta).max() <= 1e-13
    assert np.abs(qn[1:]).max() <= 1e-13
    # the plain Roe solver (no source term) does not hold the lake at rest
    qr, _ = corc.evolve(swe_desc(mx, my, dx, dy, 2, 2, lim), c, q, nsteps=5)
    assert np.abs(qr[0] + b[0] - eta).max() > 1e-4


@pytest.mark.parametrize("split", [False, True])
def test_depth_is_conserved_periodic(split):
    mx, my = 48, 40
    b = bump(mx, my, seed=4)
    q = smooth_swe(mx, my, seed=8)
    q[0] -= b[0]
    q = np.asfortranarray(q)
    d = bathy_desc(mx, my, 1.0 / mx, 1.0 / my, 2, 2, lim=4)
    c = corc.make_ctrl(bc_lower=(2, 2), bc_upper=(2, 2), dimensional_split=split, dt_initial=0.005)
    m0 = corc.component_sum(d, q, 0)
    qn, st = corc.evolve(d, c, q, nsteps=40, aux=b)
    assert st["numsteps"] == 40
    assert abs(corc.component_sum(d, qn, 0) - m0) <= 1e-13 * abs(m0)
    # momentum changes through the bathymetry source term
    assert abs(corc.component_sum(d, qn, 1) - corc.component_sum(d, q, 1)) > 1e-6


def test_both_dry_interfaces_have_no_waves_and_dry_land_stays_dry():
    n, g = 20, 2
    L = n + 2 * g
    q = np.zeros((3, L))
    q[0, :10] = 1.0  # wet on the left, dry (h = 0) on the right
    q[1, :10] = 0.2
    b = np.zeros((1, L))
    b[0, 10:] = 2.0
    wave, s, amdq, apdq = corc.rpn2(bathy_desc(n, n, 0.1, 0.1), 1, q, b)
    assert np.all(wave[:, :, 11:] == 0.0) and np.all(s[:, 11:] == 0.0)
    assert np.all(amdq[:, 11:] == 0.0) and np.all(apdq[:, 11:] == 0.0)
    # a dam break onto a dry flat bed: cells far ahead of the front stay exactly dry
    mx, my = 40, 12
    h = np.zeros((mx, my))
    h[:12] = 1.0
    qq = np.asfortranarray(np.stack([h, 0 * h, 0 * h]))
    bb = np.zeros((1, mx, my), order="F")
    d = bathy_desc(mx, my, 1.0 / mx, 1.0 / my, 2, 2, lim=1)
    c = corc.make_ctrl(bc_lower=(3, 3), bc_upper=(3, 3), dt_initial=1e-3)
    qn, st = corc.evolve(d, c, qq, nsteps=3, aux=bb)
    assert np.all(np.isfinite(qn))
    assert np.all(qn[:, 20:] == 0.0)
    assert qn[0, 12:14].min() > 0.0  # the front has moved onto the dry bed


CROSS = list(itertools.product(
    [((1, 1), (1, 1), None, None), ((2, 3), (2, 3), None, None), ((3, 1), (3, 3), (1, 1), (1, 1)),
     ((2, 2), (2, 2), None, None)],
    [(2, 2, 4), (2, 1, 1), (1, 2, 0), (2, 0, 3), (2, 2, 2), (2, 2, 5)],
))


@pytest.mark.parametrize("bcs,method", CROSS)
def test_c_oracle_matches_torch_restatement(bcs, method):
    bcl, bcu, abl, abu = bcs
    order, tw, lim = method
    mx, my = 26, 21
    q = smooth_swe(mx, my, seed=mx + lim + 3 * tw, umax=0.6)
    b = bump(mx, my, seed=lim)
    q[0] = np.maximum(q[0] - b[0] - 0.5, 0.0)
    q[0, 3:6, 4:8] = 0.0          # a dry patch, and one with h below the dry tolerance
    q[0, 12, 10] = 0.5 * TOL
    q = np.asfortranarray(q)
    dx, dy = 1.0 / mx, 1.2 / my
    dt = 0.25 * min(dx, dy) / (np.sqrt(G * 1.5) + 1.0)
    d = bathy_desc(mx, my, dx, dy, order, tw, lim)
    c = corc.make_ctrl(bc_lower=bcl, bc_upper=bcu, aux_bc_lower=abl, aux_bc_upper=abu)
    got, cfl = corc.step_hyperbolic(d, c, q, dt, aux=b)
    ref, rcfl = ts.step(torch.tensor(q), dt, dx, dy, "bathy", (G, TOL), mthlim=(lim,) * 3,
                        bc_lower=bcl, bc_upper=bcu, order=order, transverse_waves=tw,
                        b=torch.tensor(b[0]), aux_bc_lower=abl, aux_bc_upper=abu)
    ref = ref.numpy()
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    assert abs(cfl - float(rcfl)) <= 1e-14 * cfl
    # the split step too
    cs = corc.make_ctrl(bc_lower=bcl, bc_upper=bcu, aux_bc_lower=abl, aux_bc_upper=abu,
                        dimensional_split=True)
    got, cfl = corc.step_hyperbolic(d, cs, q, dt, aux=b)
    ref, rcfl = ts.step_split(torch.tensor(q), dt, dx, dy, "bathy", (G, TOL), mthlim=(lim,) * 3,
                              bc_lower=bcl, bc_upper=bcu, order=order, b=torch.tensor(b[0]),
                              aux_bc_lower=abl, aux_bc_upper=abu)
    ref = ref.numpy()
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    assert abs(cfl - float(rcfl)) <= 1e-14 * cfl


def test_aux_ghost_fill_mirrors_without_sign_change():
    mx, my = 7, 5
    d = bathy_desc(mx, my, 0.1, 0.1)
    c = corc.make_ctrl(bc_lower=(3, 1), bc_upper=(3, 1), aux_bc_lower=(3, 1), aux_bc_upper=(1, 3))
    aux = np.asfortranarray(np.arange(mx * my, dtype=float).reshape(1, mx, my) + 1.0)
    abc = corc.fill_aux_ghosts(d, c, aux)
    ref = ts.fill_aux_ghosts(torch.tensor(aux[0]), 2, (3, 1), (1, 3)).numpy()
    assert np.array_equal(abc[0], ref)
    assert np.array_equal(abc[0, 1, 2:-2], aux[0, 0])     # wall: mirror, positive
    assert np.array_equal(abc[0, 0, 2:-2], aux[0, 1])
    assert np.array_equal(abc[0, -1, 2:-2], aux[0, -1])   # extrap
