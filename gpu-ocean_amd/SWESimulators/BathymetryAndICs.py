"""Seeded synthetic bathymetry and initial conditions.

Helpers in the spirit of gpu_ocean/SWESimulators/BathymetryAndICs.py
(Gaussian bumps, geostrophically balanced velocities, :56-74) and the input
manifest of the five benchmark workloads of this build:

  C1  plain 256x256, dx=dy=200 m, flat H=60 m, walls, f=1.2e-4, Gaussian bump
  C2  plain 2048x2048, dx=dy=500 m, Gaussian-random-field depth, periodic E/W,
      walls N/S, 20 balanced eddies, uniform wind 0.1 N/m^2 at 30 degrees
  C3  100 members of 512x512 sharing one random depth, walls, member-seeded
      random-field perturbations of eta
  C4  one 512x512 domain for the adjoint (walls, random depth)
  C5  32 domains with sides drawn from {256, 384, 512, 768}

Every array is (ny+4, nx+4) for state and (ny+5, nx+5) for corner depth.
"""

from __future__ import annotations

import dataclasses
from typing import List

import numpy as np

G_DEFAULT = 9.81
RHO_WATER = 1025.0


def cell_coords(nx, ny, dx, dy):
    """Cell-centre coordinates of the padded grid, (i+0.5)*dx for interior
    index i, ghosts at i=-2,-1 and nx, nx+1."""
    x = (np.arange(-2, nx + 2) + 0.5) * dx
    y = (np.arange(-2, ny + 2) + 0.5) * dy
    return np.meshgrid(x, y)


def gaussian_random_field(shape, corr_cells, rng, std=1.0, mean=0.0):
    """Periodic Gaussian random field with Gaussian correlation of length
    ``corr_cells`` (FFT filter of white noise), scaled to ``std``/``mean``."""
    ny, nx = shape
    noise = rng.standard_normal(shape)
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.fftfreq(nx)[None, :]
    k2 = (2 * np.pi) ** 2 * (kx ** 2 + ky ** 2)
    spec = np.exp(-0.5 * k2 * corr_cells ** 2)
    field = np.real(np.fft.ifft2(np.fft.fft2(noise) * np.sqrt(spec)))
    s = field.std()
    field = (field - field.mean()) / (s if s > 0 else 1.0)
    return mean + std * field


def geostrophic_momentum(eta, H_m, f, g, dx, dy):
    """hu = -(g/f) h d(eta)/dy, hv = (g/f) h d(eta)/dx with central
    differences (BathymetryAndICs.initializeBalancedVelocityField :56-74)."""
    h = eta + H_m
    hu = np.zeros_like(eta)
    hv = np.zeros_like(eta)
    hu[1:-1, 1:-1] = -(g / f) * h[1:-1, 1:-1] * (eta[2:, 1:-1] - eta[:-2, 1:-1]) / (2.0 * dy)
    hv[1:-1, 1:-1] = (g / f) * h[1:-1, 1:-1] * (eta[1:-1, 2:] - eta[1:-1, :-2]) / (2.0 * dx)
    return hu, hv


def corner_to_cell_depth(Hi):
    return 0.25 * (Hi[:-1, :-1] + Hi[1:, :-1] + Hi[:-1, 1:] + Hi[1:, 1:])


def random_depth(nx, ny, rng, mean=100.0, std=30.0, corr_cells=32.0, min_depth=10.0):
    Hi = gaussian_random_field((ny + 5, nx + 5), corr_cells, rng, std=std, mean=mean)
    return np.maximum(Hi, min_depth).astype(np.float32)


def eddies(nx, ny, dx, dy, rng, n=20, amp=0.3, rmin=15, rmax=40):
    """Sum of ``n`` Gaussian eddies of amplitude +-amp and radius rmin..rmax
    cells at seeded positions."""
    X, Y = cell_coords(nx, ny, dx, dy)
    eta = np.zeros_like(X)
    for _ in range(n):
        xc = rng.uniform(0, nx * dx)
        yc = rng.uniform(0, ny * dy)
        R = rng.uniform(rmin, rmax) * dx
        a = amp * (1.0 if rng.random() < 0.5 else -1.0)
        eta += a * np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / (2.0 * R * R))
    return eta


@dataclasses.dataclass
class DomainInput:
    """Inputs of one domain: padded state, corner depths and parameters."""
    nx: int
    ny: int
    dx: float
    dy: float
    eta: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    Hi: np.ndarray
    f: float = 1.2e-4
    g: float = G_DEFAULT
    r: float = 0.0
    bc: tuple = (1, 1, 1, 1)          # north, east, south, west
    wind_x: float = 0.0               # m^2/s^2
    wind_y: float = 0.0
    courant: float = 0.7
    steps: int = 1
    name: str = ""


def config1(seed=1, nx=256, ny=256):
    """C1: plain 256^2, flat 60 m, walls, centred Gaussian bump (0.5 m,
    sigma 10 cells) plus N(0, 1e-3) noise, hu = hv = 0, 200 steps."""
    rng = np.random.default_rng(seed)
    dx = dy = 200.0
    X, Y = cell_coords(nx, ny, dx, dy)
    xc, yc = nx * dx / 2.0, ny * dy / 2.0
    s = 10.0 * dx
    eta = 0.5 * np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / (2.0 * s * s))
    eta = eta + rng.normal(0.0, 1e-3, eta.shape)
    z = np.zeros_like(eta, dtype=np.float32)
    Hi = np.full((ny + 5, nx + 5), 60.0, dtype=np.float32)
    return DomainInput(nx, ny, dx, dy, eta.astype(np.float32), z.copy(), z.copy(), Hi,
                       f=1.2e-4, bc=(1, 1, 1, 1), steps=200, name="C1 plain 256x256 walls")


def config2(seed=2, nx=2048, ny=2048):
    """C2: plain 2048^2, dx=dy=500 m, random depth (mean 100, std 30, corr 32
    cells, >= 10 m), periodic E/W + walls N/S, 20 balanced eddies, wind
    0.1 N/m^2 at 30 degrees, f = 1.2e-4, 1000 steps."""
    rng = np.random.default_rng(seed)
    dx = dy = 500.0
    Hi = random_depth(nx, ny, rng)
    eta = eddies(nx, ny, dx, dy, rng, n=20, amp=0.3, rmin=15, rmax=40)
    Hm = corner_to_cell_depth(Hi.astype(np.float64))
    hu, hv = geostrophic_momentum(eta, Hm, 1.2e-4, G_DEFAULT, dx, dy)
    a = np.deg2rad(30.0)
    return DomainInput(nx, ny, dx, dy, eta.astype(np.float32), hu.astype(np.float32),
                       hv.astype(np.float32), Hi, f=1.2e-4, bc=(1, 2, 1, 2),
                       wind_x=float(0.1 * np.cos(a) / RHO_WATER),
                       wind_y=float(0.1 * np.sin(a) / RHO_WATER),
                       steps=1000, name="C2 plain 2048x2048 periodic-EW")


def config3(seed=3, members=100, n=512):
    """C3: ``members`` domains of n^2 (dx=dy=500 m) sharing one random depth,
    walls; member m's eta = base eddies + GRF_m (std 0.05 m, corr 16 cells,
    seed + 1000 + m), momenta balanced geostrophically, 500 steps."""
    rng = np.random.default_rng(seed)
    dx = dy = 500.0
    Hi = random_depth(n, n, rng)
    base = eddies(n, n, dx, dy, rng, n=8, amp=0.3, rmin=15, rmax=40)
    Hm = corner_to_cell_depth(Hi.astype(np.float64))
    out: List[DomainInput] = []
    for m in range(members):
        r = np.random.default_rng(seed * 100003 + 1000 + m)
        eta = base + gaussian_random_field(base.shape, 16.0, r, std=0.05)
        hu, hv = geostrophic_momentum(eta, Hm, 1.2e-4, G_DEFAULT, dx, dy)
        out.append(DomainInput(n, n, dx, dy, eta.astype(np.float32), hu.astype(np.float32),
                               hv.astype(np.float32), Hi, f=1.2e-4, bc=(1, 1, 1, 1),
                               steps=500, name=f"C3 member {m}"))
    return out


def config4(seed=4, n=512):
    """C4: one n^2 domain (dx=dy=500 m, walls, random depth) for the adjoint."""
    rng = np.random.default_rng(seed)
    dx = dy = 500.0
    Hi = random_depth(n, n, rng)
    eta = eddies(n, n, dx, dy, rng, n=8, amp=0.3, rmin=15, rmax=40)
    Hm = corner_to_cell_depth(Hi.astype(np.float64))
    hu, hv = geostrophic_momentum(eta, Hm, 1.2e-4, G_DEFAULT, dx, dy)
    return DomainInput(n, n, dx, dy, eta.astype(np.float32), hu.astype(np.float32),
                       hv.astype(np.float32), Hi, f=1.2e-4, bc=(1, 1, 1, 1), steps=500,
                       name="C4 adjoint 512x512")


def config5(seed=5, domains=32, sides=(256, 384, 512, 768)):
    """C5: ``domains`` square domains with sides drawn from ``sides``, own
    random depth (mean 80, std 25, >= 10 m), walls, 0.05 m eta perturbation."""
    rng = np.random.default_rng(seed)
    out = []
    dx = dy = 500.0
    for k in range(domains):
        n = int(rng.choice(sides))
        r = np.random.default_rng(seed * 7919 + k)
        Hi = random_depth(n, n, r, mean=80.0, std=25.0, corr_cells=32.0, min_depth=10.0)
        eta = gaussian_random_field((n + 4, n + 4), 16.0, r, std=0.05)
        Hm = corner_to_cell_depth(Hi.astype(np.float64))
        hu, hv = geostrophic_momentum(eta, Hm, 1.2e-4, G_DEFAULT, dx, dy)
        out.append(DomainInput(n, n, dx, dy, eta.astype(np.float32), hu.astype(np.float32),
                               hv.astype(np.float32), Hi, f=1.2e-4, bc=(1, 1, 1, 1), steps=200,
                               name=f"C5 domain {k} ({n}x{n})"))
    return out
