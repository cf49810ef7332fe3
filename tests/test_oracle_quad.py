"""Pins of the oracle on general quadrilateral grids (CPU only).

The reference's flux path reads per-cell volumes, edge lengths and edge normals (quad_cell.f90:69-229,
flux_solver.f90:254-367 and 398-422).  On grids that are not rectilinear the oracle's geometry and the
normal-dependent terms of its fluxes are pinned by what the mathematics fixes:

* geometry: unit normals perpendicular to their edges and pointing out of the cell, edge lengths equal to
  the node distances, the parallelogram volume of quad_cell.f90:160-175 equal to the exact (shoelace)
  area on an affine grid, and closed cells (sum of length x normal over the four edges = 0);
* rotation: on a uniform grid rotated by theta, with the edge states rotated alike, the fluxes whose
  normal velocity is the projection u nx + v ny (SLAU family, Roe, AUSM+, HLLC) are the rotated fluxes
  of the unrotated grid (AUSMPW+ and M-AUSMPW+ take u or v by the edge's index direction,
  ausmpw_plus_solver.fypp:343-351, so they are not rotation invariant and are left out);
* free stream on a skewed periodic grid (the same projecting fluxes) and conservation on it (every flux).
"""
import numpy as np
import pytest

from fvleg_2d_amd.problems import wavy_nodes

PROJECTING = ["SLAU", "SLAU2", "SD-SLAU", "SD-SLAU2", "Roe", "AUSM+", "HLLC"]
ALL = PROJECTING + ["AUSMPW+", "M-AUSMPW+"]


def rotated_nodes(ni, nj, nh, theta, shear=0.0):
    """Uniform nodes on [0,1]^2 (ghosts included), sheared by x += shear * y, then rotated by theta."""
    dx, dy = 1.0 / ni, 1.0 / nj
    x = np.arange(-nh, ni + nh + 1) * dx
    y = np.arange(-nh, nj + nh + 1) * dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    X = X + shear * Y
    c, s = np.cos(theta), np.sin(theta)
    return np.asfortranarray(c * X - s * Y), np.asfortranarray(s * X + c * Y)


def node_quads(nx, ny):
    x = np.stack([nx[:-1, :-1], nx[1:, :-1], nx[1:, 1:], nx[:-1, 1:]])
    y = np.stack([ny[:-1, :-1], ny[1:, :-1], ny[1:, 1:], ny[:-1, 1:]])
    return x, y


@pytest.mark.parametrize("grid", ["wavy", "rotated_sheared"])
def test_quad_geometry_is_closed_and_normals_are_outward(oracle, grid):
    ni, nj = 14, 11
    cfg = oracle.Config(ni, nj)
    if grid == "wavy":
        nx, ny = wavy_nodes(ni, nj, cfg.nh, amp=0.06)
    else:
        nx, ny = rotated_nodes(ni, nj, cfg.nh, 0.7, shear=0.4)
    g = oracle.Geometry(cfg, nx, ny)
    x, y = node_quads(nx, ny)
    for e in range(4):
        h = (e + 1) % 4
        ex, ey = x[h] - x[e], y[h] - y[e]
        length = np.hypot(ex, ey)
        n = g.edge_norm[:, e]
        assert np.allclose(np.hypot(n[0], n[1]), 1.0, rtol=0, atol=1e-15)
        # perpendicular, to the rounding of the midpoint form (mid + dy) - mid: eps x |coordinates|
        assert np.abs(n[0] * ex + n[1] * ey).max() <= 1e-15
        assert (n[0] * ey - n[1] * ex).min() > 0  # the edge turns left round the cell: (ey, -ex) is outward
        # node distance, up to the reference's snapping of coordinate differences below epsilon
        assert np.allclose(g.edge_lengths[e], length, rtol=1e-15, atol=2.3e-16)
    closure = sum(g.edge_lengths[e] * g.edge_norm[:, e] for e in range(4))
    assert np.abs(closure).max() <= 1e-15
    if grid == "rotated_sheared":  # parallelograms: the volume is the exact area
        area = 0.5 * np.abs((x[2] - x[0]) * (y[3] - y[1]) - (x[3] - x[1]) * (y[2] - y[0]))
        assert np.allclose(g.volume, area, rtol=1e-14, atol=0)
        assert np.allclose(g.volume, (1.0 / ni) * (1.0 / nj), rtol=1e-13, atol=0)


def random_edges(ni, nj, seed):
    rng = np.random.default_rng(seed)
    shape = (2, ni + 1, nj + 1)
    rho = [rng.uniform(0.3, 2.0, shape) for _ in range(2)]
    u = [rng.uniform(-1.5, 1.5, shape) for _ in range(2)]
    v = [rng.uniform(-1.5, 1.5, shape) for _ in range(2)]
    p = [rng.uniform(0.2, 3.0, shape) for _ in range(2)]
    return rho, u, v, p


@pytest.mark.parametrize("flux", PROJECTING)
@pytest.mark.parametrize("theta", [0.3, 2.2])
def test_projecting_fluxes_rotate_with_the_grid(oracle, flux, theta):
    """F(R s; R e_x) = R F(s; e_x): edge states (random, sub- and supersonic) rotated with the grid."""
    ni, nj = 9, 7
    cfg = oracle.Config(ni, nj, flux=flux, limiter="minmod")
    nh = cfg.nh
    g0 = oracle.Geometry(cfg, *rotated_nodes(ni, nj, nh, 0.0))
    g1 = oracle.Geometry(cfg, *rotated_nodes(ni, nj, nh, theta))
    c, s = np.cos(theta), np.sin(theta)
    rho, u, v, p = random_edges(ni, nj, 17)
    ur = [c * a - s * b for a, b in zip(u, v)]
    vr = [s * a + c * b for a, b in zip(u, v)]
    cell = [cfg.field() for _ in range(4)]
    cell[0][:] = 1.0
    rng = np.random.default_rng(3)
    cell[3][:] = rng.uniform(0.5, 1.5, cell[3].shape)  # sensor pressures (index-based, unrotated)
    e0 = [rho[0], rho[1], u[0], u[1], v[0], v[1], p[0], p[1]]
    e1 = [rho[0], rho[1], ur[0], ur[1], vr[0], vr[1], p[0], p[1]]
    i0, j0 = oracle.edge_fluxes(cfg, g0, *cell, e0)
    i1, j1 = oracle.edge_fluxes(cfg, g1, *cell, e1)
    for f0, f1 in ((i0, i1), (j0, j1)):
        scale = np.abs(f0).max()
        want = [f0[0], c * f0[1] - s * f0[2], s * f0[1] + c * f0[2], f0[3]]
        for k in range(4):
            assert np.abs(f1[k] - want[k]).max() <= 1e-13 * scale, (flux, k)


@pytest.mark.parametrize("flux", PROJECTING)
@pytest.mark.parametrize("lim", ["minmod", "MLP5"])
def test_free_stream_on_a_skewed_periodic_grid(oracle, flux, lim):
    ni, nj = 16, 12
    cfg = oracle.Config(ni, nj, flux=flux, limiter=lim, bcs=("periodic",) * 4)
    g = oracle.Geometry(cfg, *wavy_nodes(ni, nj, cfg.nh, amp=0.06))
    rho, u, v, p = (cfg.field() for _ in range(4))
    rho[:], u[:], v[:], p[:] = 1.3, 0.45, -0.3, 0.9
    d = oracle.solve(cfg, g, rho, u, v, p)
    flux_scale = 1.3 * 0.45 * 0.45 + 0.9
    for x in d:
        assert np.abs(x).max() <= 1e-12 * flux_scale * ni


@pytest.mark.parametrize("flux", ALL)
def test_conservation_on_a_skewed_periodic_grid(oracle, flux):
    """Sum of volume x d/dt(U) over a periodic wavy grid vanishes: every edge flux leaves one cell and
    enters its neighbour with the same length (the normals of both sides are the one edge's)."""
    from test_oracle import random_periodic

    ni, nj = 16, 12
    cfg = oracle.Config(ni, nj, flux=flux, limiter="minmod", bcs=("periodic",) * 4)
    g = oracle.Geometry(cfg, *wavy_nodes(ni, nj, cfg.nh, amp=0.06))
    rho, u, v, p = random_periodic(cfg, 9)
    d = oracle.solve(cfg, g, rho, u, v, p)
    nh = cfg.nh
    inner = (slice(nh, -nh), slice(nh, -nh))
    for x in d:
        tot = np.sum(x[inner] * g.volume[inner])
        mag = np.sum(np.abs(x[inner] * g.volume[inner]))
        assert abs(tot) <= 1e-13 * mag + 1e-300
