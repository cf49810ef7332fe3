"""Host-side logic that needs no GPU: deck round trip, serving-queue arithmetic, tile cache keys."""
import os

import numpy as np

from conftest import ROOT  # noqa: F401


def test_deck_round_trip(tmp_path):
    """write_deck -> read_initial_conditions: the pycato layout (transposed datasets, ghost nodes/cells)."""
    from fvleg_2d_amd import h5io
    from fvleg_2d_amd.deck import write_deck
    from fvleg_2d_amd.problems import riemann2d

    pb = riemann2d(n=24, nj=16, seed=3, flux="HLLC", limiter="MLP5", max_time=0.1)
    ini = write_deck(pb, str(tmp_path), "rp")
    text = open(ini).read()
    for line in ("flux_solver = 'HLLC'", "limiter = 'MLP5'", "integration_strategy = 'ssp_rk3'",
                 "plus_x = 'zero_gradient'", "read_from_file = true"):
        assert line in text
    d = h5io.read_initial_conditions(os.path.join(str(tmp_path), "rp.h5"))
    nh = pb.cfg.n_ghost
    assert d["n_ghost_layers"] == nh == 3
    assert d["x"].shape == (24 + 2 * nh + 1, 16 + 2 * nh + 1)
    assert np.array_equal(d["x"], pb.node_x) and np.array_equal(d["y"], pb.node_y)
    assert d["density"].shape == (24 + 2 * nh, 16 + 2 * nh)
    assert np.array_equal(d["density"][nh:-nh, nh:-nh], pb.rho)
    assert np.array_equal(d["pressure"][nh:-nh, nh:-nh], pb.p)
    # ghost cells are edge copies
    assert np.array_equal(d["density"][:nh, nh:-nh], np.repeat(pb.rho[:1, :], nh, axis=0))


def test_fifo_queue_time_to_solution():
    """Single-server FIFO: start = max(arrival, previous end)."""
    from fvleg_2d_amd.benchforms import fifo_tts

    tts, end = fifo_tts([0.0, 1.0, 1.5, 10.0], [2.0, 1.0, 1.0, 0.5])
    assert np.allclose(tts, [2.0, 2.0, 2.5, 0.5]) and end == 10.5


def test_uniform_nodes_match_linspace_closed_form():
    """problems.uniform_nodes follows pycato's np.linspace construction (generate_initial_grids.py:43-112)."""
    from fvleg_2d_amd.problems import uniform_nodes

    nx, ny = uniform_nodes(8, 4, 2, (0.0, 1.0), (0.0, 0.5))
    assert nx.shape == (8 + 1 + 4, 4 + 1 + 4)
    assert np.array_equal(nx[:, 0], np.linspace(-2 / 8, 1 + 2 / 8, 13))
    assert np.array_equal(ny[0, :], np.linspace(-2 * 0.125, 0.5 + 2 * 0.125, 9))
    assert np.all(nx[:, 1:] == nx[:, :1]) and np.all(ny[1:, :] == ny[:1, :])


def test_batch_problem_distribution_is_seeded():
    from fvleg_2d_amd.problems import BATCH_FLUXES, BATCH_LIMITERS, batch_riemann

    a = batch_riemann(n_solves=12, seed=5, side_range=(16, 40))
    b = batch_riemann(n_solves=12, seed=5, side_range=(16, 40))
    assert [p.name for p in a] == [p.name for p in b]
    assert all(16 <= p.ni <= 40 and p.ni == p.nj for p in a)
    assert all(p.cfg.flux in BATCH_FLUXES and p.cfg.limiter in BATCH_LIMITERS for p in a)
    assert all(np.array_equal(x.rho, y.rho) for x, y in zip(a, b))
