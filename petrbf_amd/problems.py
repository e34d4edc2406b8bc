"""Seeded problem generators for the Gaussian-RBF interpolation workloads.

These play the role of the reference harness (2d/main.cxx:66-111,
3d/main.cxx:90-111): they lay particles on a lattice, attach the field to be
interpolated, and fix sigma and the cluster parameters.  Every random draw comes
from SplitMix64 (Steele, Lea & Flood 2014) turned into 53-bit doubles, so the
same seed gives bit-identical inputs on any host.

Cluster parameters follow the reference's units: ``nsigma_box`` is the cell
side, ``sigma_buffer`` and ``sigma_trunc`` the full widths of the buffer and
truncation boxes, all in units of sigma (2d/par.h:72-74, 2d/get_buffer.h:11,
2d/get_trunc.h:14).  The reference takes ints; here they are reals so that a
cell of 4x4 lattice spacings (4h = 3.2 sigma at h/sigma = 0.8) is expressible.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Outputs ``start .. start+count-1`` of the SplitMix64 stream of ``seed``."""
    k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform01(seed: int, count: int, start: int = 0) -> np.ndarray:
    """53-bit doubles in [0, 1) from the SplitMix64 stream."""
    return (splitmix64(seed, count, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclasses.dataclass
class Problem:
    """A particle set, its right-hand side(s) and the cluster parameters."""

    dim: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray | None
    w: np.ndarray            # (nrhs, n) field values at the particles (the RHS)
    sigma: float
    nsigma_box: float
    sigma_buffer: float
    sigma_trunc: float
    h: float
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def nrhs(self) -> int:
        return int(self.w.shape[0])


def lamb_oseen(x: np.ndarray, y: np.ndarray, nu_t: float = 0.01) -> np.ndarray:
    """Lamb-Oseen vortex vorticity exp(-r^2/(4 nu t)) / (4 pi nu t).

    The form of the commented line 2d/main.cxx:103 with nu*t as one parameter.
    """
    return np.exp(-(x * x + y * y) / (4 * nu_t)) / (math.pi * 4 * nu_t)


def jittered_lattice_2d(n_side: int, seed: int, jitter: float = 0.1,
                        lo: float = -1.0, hi: float = 1.0):
    """n_side^2 points, endpoint-inclusive lattice x = lo + k h (2d/main.cxx:97-98,
    particle i at (i // ny, i % ny)), each coordinate moved by U(-jitter h, jitter h)."""
    h = (hi - lo) / (n_side - 1)
    i = np.arange(n_side * n_side, dtype=np.int64)
    u = uniform01(seed, 2 * n_side * n_side).reshape(-1, 2)
    x = lo + (i // n_side).astype(np.float64) * h + (2.0 * u[:, 0] - 1.0) * jitter * h
    y = lo + (i % n_side).astype(np.float64) * h + (2.0 * u[:, 1] - 1.0) * jitter * h
    return x, y, h


def jittered_lattice_3d(n_side: int, seed: int, jitter: float = 0.1,
                        lo: float = -1.0, hi: float = 1.0):
    """n_side^3 points; particle i at (x, y, z) lattice indices
    ((i % (n n)) // n, i % n, i // (n n)) as 3d/main.cxx:100-102."""
    h = (hi - lo) / (n_side - 1)
    n = n_side
    i = np.arange(n ** 3, dtype=np.int64)
    u = uniform01(seed, 3 * n ** 3).reshape(-1, 3)
    x = lo + ((i % (n * n)) // n).astype(np.float64) * h + (2.0 * u[:, 0] - 1.0) * jitter * h
    y = lo + (i % n).astype(np.float64) * h + (2.0 * u[:, 1] - 1.0) * jitter * h
    z = lo + (i // (n * n)).astype(np.float64) * h + (2.0 * u[:, 2] - 1.0) * jitter * h
    return x, y, z, h


def lattice_2d_problem(n_side: int, seed: int = 0x5EED0001, overlap: float = 0.8,
                       cells_per_box: int = 4, buffer_cells: int = 8,
                       cutoff_sigma: float = 5.0, nu_t: float = 0.01,
                       jitter: float = 0.1) -> Problem:
    """The 2D lattice workload: n_side^2 jittered particles on [-1,1]^2,
    sigma = h/overlap, cells of ``cells_per_box`` lattice spacings, a buffer of
    ``buffer_cells`` spacings on each side, truncation box reaching
    ``cutoff_sigma`` sigma beyond the cell, Lamb-Oseen RHS.

    The default buffer is two whole cells (8 spacings) on each side, so each
    RASM block is the 5x5-cell patch around its cluster (about 20x20 = 400
    particles) and covers the cluster's whole truncation box (6.6 sigma from
    the centre vs 8 sigma).  With a 2-spacing buffer (8x8 blocks) GMRES(30)
    needs 169 iterations at N = 10^4 and stalls above 1e-9 at N = 4*10^4."""
    x, y, h = jittered_lattice_2d(n_side, seed, jitter)
    sigma = h / overlap
    box = cells_per_box * overlap                    # cell side in sigma
    buf = (cells_per_box + 2 * buffer_cells) * overlap
    trunc = box + 2.0 * cutoff_sigma
    w = lamb_oseen(x, y, nu_t)[None, :]
    return Problem(2, x, y, None, w, sigma, box, buf, trunc, h,
                   name=f"lattice2d-{n_side}x{n_side}")


def vortex_ring_vorticity(x, y, z, radius=0.5, core=0.1, circulation=1.0):
    """Vorticity of a Gaussian-core vortex ring about the z axis:
    omega = Gamma/(pi a^2) exp(-((rho-R)^2 + z^2)/a^2) * e_phi, returned as
    its three Cartesian components (3, n)."""
    rho = np.sqrt(x * x + y * y)
    mag = circulation / (math.pi * core * core) * np.exp(-((rho - radius) ** 2 + z * z) / (core * core))
    safe = np.where(rho > 0, rho, 1.0)
    ex = np.where(rho > 0, -y / safe, 0.0)
    ey = np.where(rho > 0, x / safe, 0.0)
    return np.stack([mag * ex, mag * ey, np.zeros_like(mag)])


def lattice_3d_problem(n_side: int, seed: int = 0x5EED0005, overlap: float = 0.8,
                       cells_per_box: int = 3, buffer_cells: int = 3,
                       cutoff_sigma: float = 4.0, jitter: float = 0.1) -> Problem:
    """The 3D vector-field workload: n_side^3 jittered particles on [-1,1]^3,
    3x3x3-spacing cells, a buffer of one whole cell (3 spacings) per side, cutoff ``cutoff_sigma`` sigma beyond
    the cell, vortex-ring vorticity as three right-hand sides."""
    x, y, z, h = jittered_lattice_3d(n_side, seed, jitter)
    sigma = h / overlap
    box = cells_per_box * overlap
    buf = (cells_per_box + 2 * buffer_cells) * overlap
    trunc = box + 2.0 * cutoff_sigma
    w = vortex_ring_vorticity(x, y, z)
    return Problem(3, x, y, z, w, sigma, box, buf, trunc, h,
                   name=f"lattice3d-{n_side}^3")


def mixture_2d_problem(n: int, seed: int = 0x5EED0003, overlap: float = 0.8,
                       cutoff_sigma: float = 5.0, nu_t: float = 0.01,
                       cells_per_box: float = 4.0, buffer_cells: float = 8.0) -> Problem:
    """Non-uniform 2D mixture: 70% uniform on [-1,1]^2, 30% split between two
    isotropic Gaussians at (+-0.4, 0) with std 0.15 (draws outside the square
    are redrawn from the same stream).  sigma is fixed from the mean spacing
    h = 2/sqrt(n) as sigma = h/overlap."""
    n_uni = (7 * n) // 10
    n_g = n - n_uni
    u = uniform01(seed, 2 * n_uni)
    xs = [2.0 * u[0::2] - 1.0]
    ys = [2.0 * u[1::2] - 1.0]
    # Box-Muller from the same stream, rejection outside the square.
    need = n_g
    start = 2 * n_uni
    got_x, got_y = [], []
    k = 0
    while need > 0:
        m = max(1024, int(need * 1.2))
        v = uniform01(seed, 3 * m, start)
        start += 3 * m
        r = np.sqrt(-2.0 * np.log1p(-v[0::3]))
        t = 2.0 * math.pi * v[1::3]
        side = np.where(v[2::3] < 0.5, -0.4, 0.4)
        gx = side + 0.15 * r * np.cos(t)
        gy = 0.15 * r * np.sin(t)
        ok = (np.abs(gx) <= 1.0) & (np.abs(gy) <= 1.0)
        gx, gy = gx[ok][:need], gy[ok][:need]
        got_x.append(gx)
        got_y.append(gy)
        need -= gx.shape[0]
        k += 1
    x = np.concatenate(xs + got_x)
    y = np.concatenate(ys + got_y)
    h = 2.0 / math.sqrt(n)
    sigma = h / overlap
    box = cells_per_box * overlap
    buf = (cells_per_box + 2 * buffer_cells) * overlap
    trunc = box + 2.0 * cutoff_sigma
    w = lamb_oseen(x, y, nu_t)[None, :]
    return Problem(2, x, y, None, w, sigma, box, buf, trunc, h, name=f"mixture2d-{n}")


def config(name: str) -> Problem:
    """Named workloads: 'c1' N=10^4 lattice, 'c2' N=4*10^6 lattice,
    'c3' N=4*10^6 mixture, 'c5' 160^3 lattice with three right-hand sides."""
    if name == "c1":
        return lattice_2d_problem(100)
    if name == "c2":
        return lattice_2d_problem(2000)
    if name == "c3":
        return mixture_2d_problem(4_000_000)
    if name == "c5":
        return lattice_3d_problem(160)
    raise KeyError(name)
