"""Seeded initial states of the workloads (pyclaw example setups restated; used by tests, smoke, bench)."""
import numpy as np


def centers(lower, upper, n):
    d = (upper - lower) / n
    return lower + (np.arange(n) + 0.5) * d, d


def radial_dam_break(mx=200, my=200, lower=-2.5, upper=2.5, r0=0.5, h_in=2.0, h_out=1.0):
    """pyclaw examples/shallow_2d/radial_dam_break.py initial data (h=2 inside r<0.5, else 1)."""
    xc, dx = centers(lower, upper, mx)
    yc, dy = centers(lower, upper, my)
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2)
    q = np.zeros((3, mx, my), order="F")
    q[0] = np.where(r < r0, h_in, h_out)
    return q, dx, dy


def gaussian_humps(mx, my, seed, n_humps=16, lower=-2.5, upper=2.5, base=1.0,
                   amp=(0.05, 0.5), width=(0.02, 0.2)):
    """h = base + sum_k A_k exp(-|x-c_k|^2/(2 w_k^2)), hu = hv = 0 (distances without wrap)."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(amp[0], amp[1], n_humps)
    W = rng.uniform(width[0], width[1], n_humps)
    cx = rng.uniform(lower, upper, n_humps)
    cy = rng.uniform(lower, upper, n_humps)
    xc, dx = centers(lower, upper, mx)
    yc, dy = centers(lower, upper, my)
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    h = np.full((mx, my), base)
    for k in range(n_humps):
        h += A[k] * np.exp(-((X - cx[k]) ** 2 + (Y - cy[k]) ** 2) / (2.0 * W[k] ** 2))
    q = np.zeros((3, mx, my), order="F")
    q[0] = h
    return q, dx, dy


def euler_quadrants(mx, my, seed, gamma=1.4, split=(0.5, 0.5)):
    """Four-quadrant Riemann problem on [0,1]^2: rho~U[0.1,1.5], p~U[0.03,1.5], u,v~U[-1,1]."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.1, 1.5, 4)
    p = rng.uniform(0.03, 1.5, 4)
    u = rng.uniform(-1.0, 1.0, 4)
    v = rng.uniform(-1.0, 1.0, 4)
    xc, dx = centers(0.0, 1.0, mx)
    yc, dy = centers(0.0, 1.0, my)
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    quad = (X >= split[0]).astype(int) + 2 * (Y >= split[1]).astype(int)
    q = np.zeros((4, mx, my), order="F")
    R, U, V, P = rho[quad], u[quad], v[quad], p[quad]
    q[0] = R
    q[1] = R * U
    q[2] = R * V
    q[3] = P / (gamma - 1.0) + 0.5 * R * (U * U + V * V)
    return q, dx, dy


def random_smooth_swe(mx, my, seed, amp=0.3):
    rng = np.random.default_rng(seed)
    q = np.zeros((3, mx, my), order="F")
    q[0] = 1.0 + 0.5 * rng.random((mx, my))
    q[1] = amp * rng.standard_normal((mx, my))
    q[2] = amp * rng.standard_normal((mx, my))
    return q


def random_euler(mx, my, seed, gamma=1.4):
    rng = np.random.default_rng(seed)
    rho = 0.5 + rng.random((mx, my))
    u = 0.5 * rng.standard_normal((mx, my))
    v = 0.5 * rng.standard_normal((mx, my))
    p = 0.5 + rng.random((mx, my))
    q = np.zeros((4, mx, my), order="F")
    q[0], q[1], q[2] = rho, rho * u, rho * v
    q[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return q


def batch_workload(ngrids, seed, nmin=64, nmax=512, lower=-2.5, upper=2.5):
    """Independent grids with side lengths ~ U{nmin..nmax} (non-square allowed), each carrying
    16-hump data on its own [lower, upper]^2 (dx = (upper-lower)/mx, dy = (upper-lower)/my);
    grid g's humps use seed + 1 + g."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=(ngrids, 2))
    out = []
    for g, (mx, my) in enumerate(sizes):
        q, dx, dy = gaussian_humps(int(mx), int(my), seed + 1 + g, lower=lower, upper=upper)
        out.append((q, dx, dy))
    return out


def inversion_workload(mx=512, my=512, seed=4, n_humps=8, n_gauges=64, lower=-2.5, upper=2.5):
    """Initial-surface inversion: true eta0 = 8 seeded Gaussian humps (the 16-hump recipe with
    n_humps=8, base 0), gauges = n_gauges distinct seeded cells.  Returns (eta_true as (my, mx)
    planes, gauge flat indices y*mx + x, dx, dy)."""
    q, dx, dy = gaussian_humps(mx, my, seed, n_humps=n_humps, lower=lower, upper=upper, base=0.0)
    eta = np.ascontiguousarray(q[0].T)
    rng = np.random.default_rng(seed + 1000)
    gauges = np.sort(rng.choice(mx * my, size=n_gauges, replace=False))
    return eta, gauges, dx, dy


def sill(mx=200, my=200, lower=-1.0, upper=1.0, eta_amp=0.1, sill_amp=0.8, depth=1.0):
    """Shallow water over a sill, in the style of pyclaw's shallow_2d bathymetry example, made
    two-dimensional: bathymetry b = sill_amp exp(-(x^2 + (y/2)^2)/0.2^2) - depth, a surface hump
    eta = eta_amp exp(-((x+0.4)^2 + (y-0.2)^2)/0.2^2) on sea level 0, at rest; h = eta - b.
    Returns (q (3, mx, my), aux (1, mx, my), dx, dy)."""
    xc, dx = centers(lower, upper, mx)
    yc, dy = centers(lower, upper, my)
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    b = sill_amp * np.exp(-(X ** 2 + (Y / 2.0) ** 2) / 0.04) - depth
    eta = eta_amp * np.exp(-((X + 0.4) ** 2 + (Y - 0.2) ** 2) / 0.04)
    q = np.zeros((3, mx, my), order="F")
    q[0] = np.maximum(eta - b, 0.0)
    aux = np.asfortranarray(b[None])
    return q, aux, dx, dy


def random_bathymetry(mx, my, seed, dry_patch=False, amp=0.3):
    """Random moving water over smooth random bathymetry (wet everywhere unless dry_patch, which
    empties a block of cells and sets one cell below the usual dry tolerance 1e-3).  Returns
    (q (3, mx, my), aux (1, mx, my))."""
    rng = np.random.default_rng(seed)
    xc = (np.arange(mx) + 0.5) / mx
    yc = (np.arange(my) + 0.5) / my
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    b = -1.0 + 0.3 * np.exp(-30.0 * ((X - rng.uniform(0.3, 0.7)) ** 2 + (Y - rng.uniform(0.3, 0.7)) ** 2)) \
        + 0.1 * np.sin(2 * np.pi * (X + 2 * Y + rng.uniform()))
    q = np.zeros((3, mx, my), order="F")
    q[0] = 0.1 * rng.random((mx, my)) - b
    q[1] = amp * rng.standard_normal((mx, my)) * q[0]
    q[2] = amp * rng.standard_normal((mx, my)) * q[0]
    if dry_patch:
        i0, j0 = mx // 4, my // 3
        q[:, i0:i0 + 3, j0:j0 + 4] = 0.0
        q[0, mx // 2, my // 2] = 5e-4
    return q, np.asfortranarray(b[None])
