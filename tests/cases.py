"""Small seeded inputs shared by the CPU and GPU tests (edge cases the
reference's lists must survive: empty cells, one particle, one crowded cell,
points on cell boundaries, distinct target/source sets, 3D)."""

import numpy as np

from petrbf_amd import problems


def lattice(n_side, seed=11, buffer_cells=8, **kw):
    """Jittered lattice; buffer of `buffer_cells` lattice spacings per side
    (8 = two 4-spacing cells, the solver's default)."""
    return problems.lattice_2d_problem(n_side, seed=seed, buffer_cells=buffer_cells, **kw)


def cases_2d():
    out = {}
    p = lattice(24)
    out["lattice24"] = (2, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, [p.x, p.y], None)
    p = lattice(20, buffer_cells=2)
    out["lattice20_thin_buffer"] = (2, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, [p.x, p.y], None)
    rng = np.random.default_rng(3)
    # clustered random points: many empty cells, some crowded ones
    x = np.concatenate([rng.normal(0.3, 0.05, 300), rng.uniform(-1, 1, 100)])
    y = np.concatenate([rng.normal(-0.2, 0.05, 300), rng.uniform(-1, 1, 100)])
    out["clustered"] = (2, 0.02, 3.2, 16.0, 13.2, [x, y], None)
    # one particle
    out["single"] = (2, 0.1, 3.2, 16.0, 13.2, [np.array([0.25]), np.array([-0.5])], None)
    # everything inside one cell
    out["one_cell"] = (2, 0.5, 3.2, 16.0, 13.2, [rng.uniform(0, 0.1, 50), rng.uniform(0, 0.1, 50)], None)
    # exact lattice (no jitter): many points exactly on cell and box edges
    xs, ys, h = problems.jittered_lattice_2d(17, seed=1, jitter=0.0)
    out["exact_lattice"] = (2, h / 0.8, 3.2, 16.0, 13.2, [xs, ys], None)
    # distinct targets (h/2 grid) and sources (lattice), as 2d/main2.cxx evaluates
    p = lattice(16)
    xt, yt, _ = problems.jittered_lattice_2d(31, seed=2, jitter=0.0)
    out["targets_sources"] = (2, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, [xt, yt], [p.x, p.y])
    return out


def cases_3d():
    out = {}
    p = problems.lattice_3d_problem(9, seed=5)
    out["lattice3d_9"] = (3, p.sigma, p.nsigma_box, p.sigma_buffer, p.sigma_trunc, [p.x, p.y, p.z], None)
    rng = np.random.default_rng(7)
    pts = [rng.uniform(-1, 1, 400) for _ in range(3)]
    out["random3d"] = (3, 0.08, 2.4, 4.0, 10.4, pts, None)
    return out


def all_cases():
    d = cases_2d()
    d.update(cases_3d())
    return d
