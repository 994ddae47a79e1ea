"""NumPy restatement of gpu-ocean's drifter advection and observation
operator.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP drifter
kernels: only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` may import it.  The product path never calls it.

* ``drift`` follows ``passiveDrifterKernel``
  (gpu_ocean/SWESimulators/gpu_kernels/driftKernels.cu:108-210) in float32
  operation by operation, including its cell lookup: the cell index comes
  from ceil(x/dx) while the interpolation fraction comes from floor(x/dx),
  so away from cell faces the velocity is sampled one cell east and one cell
  north of the drifter (driftKernels.cu:137-146).  Every thread, the
  observation position included, is moved (``ti < num_drifters + 1``).
* ``enforce_bc`` follows ``enforceBoundaryConditions`` (:214-250).
* ``observe_flow`` follows ``observeUnderlyingFlow``
  (gpu_kernels/observationKernels.cu:29-82): hu, hv of the cell containing
  each drifter (floor-based index).
* ``sample_point`` is the textbook bilinear interpolation of cell-centred
  values, used by the tests as an independent statement of what ``drift``
  samples.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32


def _vel(q, eta, Hm, j, i):
    h = f32(Hm[j, i]) + f32(eta[j, i])
    return f32(q[j, i]) / h


def drift(positions, eta, hu, hv, Hm, nx, ny, dx, dy, dt, periodic_ns, periodic_ew,
          sensitivity=1.0, wind=None, wind_t=0.0, wind_drift_factor=0.0,
          x_zero=2, y_zero=2):
    """New (n+1, 2) float32 positions after one drift step.

    positions: (num_drifters + 1, 2) float32 (drifters, then the observation).
    wind: (x0, x1, y0, y1) uniform wind at the two bracketing times.
    """
    P = np.array(positions, dtype=f32, copy=True)
    dx, dy, dt, sens = f32(dx), f32(dy), f32(dt), f32(sensitivity)
    wt, wf = f32(wind_t), f32(wind_drift_factor)
    for k in range(P.shape[0]):
        x, y = P[k, 0], P[k, 1]
        qx, qy = f32(x / dx), f32(y / dy)
        cx = int(f32(np.ceil(qx)) + f32(x_zero))
        cy = int(f32(np.ceil(qy)) + f32(y_zero))
        fx = f32(qx - f32(np.floor(qx)))
        fy = f32(qy - f32(np.floor(qy)))
        x0 = cx - 1 if fx < f32(0.5) else cx
        xf = f32(fx + f32(0.5)) if fx < f32(0.5) else f32(fx - f32(0.5))
        y0 = cy - 1 if fy < f32(0.5) else cy
        yf = f32(fy + f32(0.5)) if fy < f32(0.5) else f32(fy - f32(0.5))
        x1, y1 = x0 + 1, y0 + 1
        u00, u10 = _vel(hu, eta, Hm, y0, x0), _vel(hu, eta, Hm, y0, x1)
        u01, u11 = _vel(hu, eta, Hm, y1, x0), _vel(hu, eta, Hm, y1, x1)
        v00, v10 = _vel(hv, eta, Hm, y0, x0), _vel(hv, eta, Hm, y0, x1)
        v01, v11 = _vel(hv, eta, Hm, y1, x0), _vel(hv, eta, Hm, y1, x1)
        one = f32(1.0)
        uy0 = f32(f32((one - xf) * u00) + f32(xf * u10))
        uy1 = f32(f32((one - xf) * u01) + f32(xf * u11))
        vy0 = f32(f32((one - xf) * v00) + f32(xf * v10))
        vy1 = f32(f32((one - xf) * v01) + f32(xf * v11))
        u = f32(f32((one - yf) * uy0) + f32(yf * uy1))
        v = f32(f32((one - yf) * vy0) + f32(yf * vy1))
        if wf != 0:
            wx = f32(f32(wt * f32(wind[1])) + f32(f32(one - wt) * f32(wind[0])))
            wy = f32(f32(wt * f32(wind[3])) + f32(f32(one - wt) * f32(wind[2])))
            u = f32(u + f32(wx * wf))
            v = f32(v + f32(wy * wf))
        x = f32(x + f32(f32(sens * u) * dt))
        y = f32(y + f32(f32(sens * v) * dt))
        Lx = f32(f32(nx) * dx)
        Ly = f32(f32(ny) * dy)
        if periodic_ew and x < 0:
            x = f32(x + Lx)
        if periodic_ew and x > Lx:
            x = f32(x - Lx)
        if periodic_ns and y < 0:
            y = f32(y + Ly)
        if periodic_ns and y > Ly:
            y = f32(y - Ly)
        P[k, 0], P[k, 1] = x, y
    return P


def enforce_bc(positions, domain_x, domain_y, periodic_ns, periodic_ew):
    P = np.array(positions, dtype=f32, copy=True)
    Lx, Ly = f32(domain_x), f32(domain_y)
    for k in range(P.shape[0]):
        x, y = P[k]
        if periodic_ew and x < 0:
            x = f32(x + Lx)
        if periodic_ew and x > Lx:
            x = f32(x - Lx)
        if periodic_ns and y < 0:
            y = f32(y + Ly)
        if periodic_ns and y > Ly:
            y = f32(y - Ly)
        P[k] = x, y
    return P


def observe_flow(positions, hu, hv, dx, dy, num_drifters, x_zero=2, y_zero=2):
    """(num_drifters, 2) float32: [hu, hv] at the floor-based cell."""
    out = np.zeros((num_drifters, 2), dtype=f32)
    for k in range(num_drifters):
        x, y = f32(positions[k, 0]), f32(positions[k, 1])
        i = int(f32(np.floor(f32(x / f32(dx)))) + f32(x_zero))
        j = int(f32(np.floor(f32(y / f32(dy)))) + f32(y_zero))
        out[k] = hu[j, i], hv[j, i]
    return out


def sample_point(field, x, y, dx, dy, x_zero=2, y_zero=2):
    """Bilinear interpolation in float64 of a cell-centred field (ghosted
    array, cell (j, i) centred at ((i - x_zero + 1/2) dx, (j - y_zero + 1/2) dy))."""
    fx = x / dx - 0.5 + x_zero
    fy = y / dy - 0.5 + y_zero
    i0, j0 = int(np.floor(fx)), int(np.floor(fy))
    ax, ay = fx - i0, fy - j0
    F = np.asarray(field, dtype=np.float64)
    return ((1 - ax) * (1 - ay) * F[j0, i0] + ax * (1 - ay) * F[j0, i0 + 1]
            + (1 - ax) * ay * F[j0 + 1, i0] + ax * ay * F[j0 + 1, i0 + 1])
