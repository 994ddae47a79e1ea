"""NumPy restatement of gpu-ocean's model error (OceanStateNoise).

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP model-error
kernels: only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` may import it.  The product path never calls it.

It restates two things from the reference:

* the device kernels (``gpu_ocean/SWESimulators/gpu_kernels/ocean_noise.cu``)
  as the functions without a ``cpu_twin_`` prefix: the LCG / Box-Muller
  generator, the SOAR covariance on the coarse grid, the geostrophic balance
  for interpolation factor 1 and the bicubic interpolation + balance for odd
  factors > 1, the norms (``reductions.cu``) and ``makePerpendicular``;
* the CPU twins the reference tests the device kernels against
  (``gpu_ocean/SWESimulators/OceanStateNoise.py:572-937``) as the
  ``cpu_twin_*`` functions.

The two differ where the reference's device kernels and CPU twins differ
(the kernels are what a simulator runs):

1. ``geostrophicBalance`` (ocean_noise.cu:482) adds the current eta to the
   water depth h_mid; the CPU twin (OceanStateNoise.py:766) uses H_mid + d_eta.
2. The kernels evaluate Coriolis as f + beta*((i+0.5)dx n_x + (j+0.5)dy n_y)
   on ghosted indices (ocean_noise.cu:473), or read the Coriolis field when
   f == 0; the CPU twin uses f + beta*(j+2)*dy (OceanStateNoise.py:761).
3. The kernels treat land: geostrophicBalance sets eta, hu, hv of a cell with
   a land corner to zero (:498-503); bicubicInterpolation adds no
   perturbation there (:758-763).  The CPU twin has no land.
4. bicubicInterpolation accepts alignment offsets (:617-618); the CPU twin's
   interpolation has none (OceanStateNoise.py:784-852).
5. The device generator rounds the uniform numbers to float32 before
   Box-Muller (ocean_noise.cu:87-91); the CPU twin keeps them in float64
   (OceanStateNoise.py:658-671).
6. The CPU twin reads the corner depths of interior cell (j, i) at H[j, i]
   (OceanStateNoise.py:747-749), the kernels at the ghosted H[j+2, i+2]
   (ocean_noise.cu:416-431); the reference tests use a constant H.

With eta = 0, beta = 0, no land and no offsets both restatements must agree,
which is how the reference's own tests (tests/stochastic/
OceanStateNoise_test.py) compare its kernels with its twins.

Floating-point values are computed in float64 except where the reference's
float32 arithmetic decides an index (the coarse cell of a fine cell in the
bicubic interpolation), which is reproduced in float32.
"""

from __future__ import annotations

import numpy as np

SOAR_CUTOFF = 2          # config.soar_cutoff (gpu_ocean/SWESimulators/config.py:31)
LCG_A = 1103515245
LCG_C = 12345
LCG_M = 2147483647       # 2^31 - 1
LCG_DENOM = 2147483648.0


# --------------------------------------------------------------------------
# sizes (OceanStateNoise.py:97-135)
# --------------------------------------------------------------------------
def noise_sizes(nx, ny, factor=1, periodic_ns=True, periodic_ew=True, cutoff=SOAR_CUTOFF):
    """(coarse_nx, coarse_ny, rand_nx, rand_ny, seed_nx, seed_ny, ghost_x, ghost_y)."""
    assert factor > 0 and factor % 2 == 1, "interpolation_factor must be a positive odd integer"
    assert nx % factor == 0 and ny % factor == 0
    cnx, cny = nx // factor, ny // factor
    gx = 0 if periodic_ew else 2 + cutoff
    gy = 0 if periodic_ns else 2 + cutoff
    rnx, rny = cnx + 2 * gx, cny + 2 * gy
    return cnx, cny, rnx, rny, (rnx + 1) // 2, rny, gx, gy


# --------------------------------------------------------------------------
# LCG + Box-Muller (ocean_noise.cu:80-195; OceanStateNoise.py:652-703)
# --------------------------------------------------------------------------
def lcg_next(seed):
    """One ANSI C LCG step modulo 2^31-1 on uint64 seeds (ocean_noise.cu:85-86)."""
    s = np.asarray(seed, dtype=np.uint64)
    return (s * np.uint64(LCG_A) + np.uint64(LCG_C)) % np.uint64(LCG_M)


def lcg_generate(seed, rand_nx, normal=True, uniform_float32=True):
    """Advance every seed twice and return (new_seed, random field).

    seed: (seed_ny, seed_nx) uint64.  Seed (y, x) writes columns 2x and 2x+1
    of row y when 2x+1 < rand_nx, only column 2x when 2x == rand_nx
    (ocean_noise.cu:186-192).  The second case never occurs, so for odd
    rand_nx the last column keeps what the buffer held (zeros here), as in
    the reference.

    uniform_float32: round u to float32 before Box-Muller as the device
    kernel does (``float u1 = seed / denum``); False follows the CPU twin.
    """
    s1 = lcg_next(seed)
    s2 = lcg_next(s1)
    u1 = s1.astype(np.float64) / LCG_DENOM
    u2 = s2.astype(np.float64) / LCG_DENOM
    if uniform_float32:
        u1 = u1.astype(np.float32).astype(np.float64)
        u2 = u2.astype(np.float32).astype(np.float64)
    if normal:
        with np.errstate(divide="ignore"):
            r = np.sqrt(-2.0 * np.log(u1))
        n1 = r * np.cos(2.0 * np.pi * u2)
        n2 = r * np.sin(2.0 * np.pi * u2)
    else:
        n1, n2 = u1, u2
    seed_ny, seed_nx = seed.shape
    out = np.zeros((seed_ny, rand_nx), dtype=np.float64)
    x = np.arange(seed_nx)
    full = 2 * x + 1 < rand_nx
    out[:, 2 * x[full]] = n1[:, full]
    out[:, 2 * x[full] + 1] = n2[:, full]
    return s2, out


def lcg_jump(seed, k):
    """seed after k LCG steps in closed form: a^k s + c (a^k - 1)/(a - 1) mod m.

    m is prime, so (a - 1) has an inverse; an independent check of lcg_next.
    """
    m = LCG_M
    ak = pow(LCG_A, k, m)
    geo = (ak - 1) * pow(LCG_A - 1, -1, m) % m
    return (ak * int(seed) + LCG_C * geo) % m


def cpu_twin_lcg_normal(seed, rand_nx):
    """OceanStateNoise._CPUUpdateRandom(True) (OceanStateNoise.py:674-703):
    float64 Box-Muller, stored into a float32 buffer."""
    s, out = lcg_generate(seed, rand_nx, normal=True, uniform_float32=False)
    return s, out.astype(np.float32)


# --------------------------------------------------------------------------
# Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
# easy as 1, 2, 3", SC'11).  The device generator used when use_lcg is False,
# in place of pycuda's XORWOW stream (OceanStateNoise.py:146-148).
# --------------------------------------------------------------------------
PHILOX_M0, PHILOX_M1 = 0xD2511F53, 0xCD9E8D57
PHILOX_W0, PHILOX_W1 = 0x9E3779B9, 0xBB67AE85


def philox4x32_10(ctr, key):
    """ctr: (4, N) uint32; key: two uint32.  Returns (4, N) uint32."""
    c = [np.asarray(ctr[i], dtype=np.uint64) for i in range(4)]
    k0, k1 = np.uint64(key[0]), np.uint64(key[1])
    mask = np.uint64(0xFFFFFFFF)
    for r in range(10):
        if r:
            k0 = (k0 + np.uint64(PHILOX_W0)) & mask
            k1 = (k1 + np.uint64(PHILOX_W1)) & mask
        p0 = np.uint64(PHILOX_M0) * c[0]
        p1 = np.uint64(PHILOX_M1) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & mask,
             (p0 >> np.uint64(32)) ^ c[3] ^ k1, p0 & mask]
    return np.stack(c).astype(np.uint32)


def philox_generate(rand_nx, rand_ny, key, counter, member=0, normal=True):
    """Random field of the Philox kinds: thread t makes elements 4t..4t+3 of
    the row-major field from counter (t, counter_lo, counter_hi, member) and
    key (key_lo, key_hi); u = ((x >> 9) + 0.5) 2^-23; normals by Box-Muller of
    (u0, u1) and (u2, u3)."""
    n = rand_nx * rand_ny
    nt = (n + 3) // 4
    t = np.arange(nt, dtype=np.uint64)
    ctr = np.stack([t, np.full(nt, counter & 0xFFFFFFFF, np.uint64),
                    np.full(nt, (counter >> 32) & 0xFFFFFFFF, np.uint64),
                    np.full(nt, member, np.uint64)])
    x = philox4x32_10(ctr, (key & 0xFFFFFFFF, (key >> 32) & 0xFFFFFFFF))
    u = ((x >> 9).astype(np.float64) + 0.5) * 2.0 ** -23
    if normal:
        r0 = np.sqrt(-2.0 * np.log(u[0]))
        r1 = np.sqrt(-2.0 * np.log(u[2]))
        v = np.stack([r0 * np.cos(2 * np.pi * u[1]), r0 * np.sin(2 * np.pi * u[1]),
                      r1 * np.cos(2 * np.pi * u[3]), r1 * np.sin(2 * np.pi * u[3])])
    else:
        v = u
    return v.T.reshape(-1)[:n].reshape(rand_ny, rand_nx)


# --------------------------------------------------------------------------
# SOAR (ocean_noise.cu:248-370; OceanStateNoise.py:705-760)
# --------------------------------------------------------------------------
def soar_weights(cdx, cdy, L, cutoff=SOAR_CUTOFF):
    """(1 + d/L) exp(-d/L) for offsets (dy, dx) in [-cutoff, cutoff]^2."""
    o = np.arange(-cutoff, cutoff + 1, dtype=np.float64)
    d = np.sqrt((cdx * o[None, :]) ** 2 + (cdy * o[:, None]) ** 2)
    return (1.0 + d / L) * np.exp(-d / L)


def soar_coarse(random, cnx, cny, cdx, cdy, q0, L, scale=1.0, periodic_ns=True,
                periodic_ew=True, coarse=None, additive=False):
    """Coarse buffer ((cny+4, cnx+4)) = scale * sum_b Q(a, b) xi(b).

    Coarse index a maps to random index a+2 without periodicity (the random
    field has 2+cutoff ghost cells) and to (a-2) mod n with it
    (ocean_noise.cu:305-327)."""
    c = SOAR_CUTOFF
    w = q0 * soar_weights(cdx, cdy, L)
    xi = np.asarray(random, dtype=np.float64)
    ay = np.arange(cny + 4)
    ax = np.arange(cnx + 4)
    out = np.zeros((cny + 4, cnx + 4))
    for oy in range(-c, c + 1):
        ry = (ay - 2 + oy) % cny if periodic_ns else ay + 2 + oy
        for ox in range(-c, c + 1):
            rx = (ax - 2 + ox) % cnx if periodic_ew else ax + 2 + ox
            out += w[oy + c, ox + c] * xi[np.ix_(ry, rx)]
    out *= scale
    if additive:
        return np.asarray(coarse, dtype=np.float64) + out
    return out


def cpu_twin_applyQ(random, cnx, cny, cdx, cdy, q0, L, periodic_ns, periodic_ew):
    """OceanStateNoise._applyQ_CPU (OceanStateNoise.py:715-760), as loops."""
    c = SOAR_CUTOFF
    rny, rnx = random.shape
    nyh, nxh = cny + (2 + c) * 2, cnx + (2 + c) * 2
    local = np.zeros((nyh, nxh))
    for j in range(nyh):
        gj = (j - c - 2) % rny if periodic_ns else j
        for i in range(nxh):
            gi = (i - c - 2) % rnx if periodic_ew else i
            local[j, i] = random[gj, gi]
    out = np.zeros((cny + 4, cnx + 4))
    for a_y in range(cny + 4):
        for a_x in range(cnx + 4):
            la_x, la_y = a_x + c, a_y + c
            acc = 0.0
            for b_y in range(la_y - c, la_y + c + 1):
                for b_x in range(la_x - c, la_x + c + 1):
                    dist = np.sqrt(cdx * cdx * (la_x - b_x) ** 2 + cdy * cdy * (la_y - b_y) ** 2)
                    acc += q0 * (1.0 + dist / L) * np.exp(-dist / L) * local[b_y, b_x]
            out[a_y, a_x] = acc
    return out


# --------------------------------------------------------------------------
# geostrophic balance, factor 1 (ocean_noise.cu:379-505)
# --------------------------------------------------------------------------
def _coriolis(nx, ny, dx, dy, f, beta, north, f_field):
    tj, ti = np.meshgrid(np.arange(2, ny + 2), np.arange(2, nx + 2), indexing="ij")
    if f == 0.0:
        return np.asarray(f_field, dtype=np.float64)[2:ny + 2, 2:nx + 2]
    if np.ndim(north[0]) == 0:
        n_x, n_y = float(north[0]), float(north[1])
    else:
        n_x = np.asarray(north[0], dtype=np.float64)[2:ny + 2, 2:nx + 2]
        n_y = np.asarray(north[1], dtype=np.float64)[2:ny + 2, 2:nx + 2]
    return f + beta * ((ti + 0.5) * dx * n_x + (tj + 0.5) * dy * n_y)


def _land_and_hmid(Hi, nx, ny, land):
    H = np.asarray(Hi, dtype=np.float64)
    c00 = H[2:ny + 2, 2:nx + 2]
    c01 = H[2:ny + 2, 3:nx + 3]
    c10 = H[3:ny + 3, 2:nx + 2]
    c11 = H[3:ny + 3, 3:nx + 3]
    Hf = np.asarray(Hi, dtype=np.float32)
    dry = ((Hf[2:ny + 2, 2:nx + 2] == np.float32(land)) | (Hf[2:ny + 2, 3:nx + 3] == np.float32(land))
           | (Hf[3:ny + 3, 2:nx + 2] == np.float32(land)) | (Hf[3:ny + 3, 3:nx + 3] == np.float32(land)))
    with np.errstate(over="ignore", invalid="ignore"):
        hmid = 0.25 * (c00 + c01 + c10 + c11)
    return dry, hmid


def geostrophic_balance(coarse, eta, hu, hv, Hi, nx, ny, dx, dy, g, f, beta=0.0,
                        north=(0.0, 1.0), f_field=None, land=1e20):
    """eta, hu, hv (ny+4, nx+4) after adding d_eta = coarse and its balanced
    (d_hu, d_hv) on the interior (ocean_noise.cu:443-504)."""
    d = np.asarray(coarse, dtype=np.float64)
    e = np.array(eta, dtype=np.float64)
    u = np.array(hu, dtype=np.float64)
    v = np.array(hv, dtype=np.float64)
    dry, hmid = _land_and_hmid(Hi, nx, ny, land)
    cor = _coriolis(nx, ny, dx, dy, f, beta, north, f_field)
    I = (slice(2, ny + 2), slice(2, nx + 2))
    dc = d[I]
    ex = (d[2:ny + 2, 3:nx + 3] - d[2:ny + 2, 1:nx + 1]) / (2.0 * dx)
    ey = (d[3:ny + 3, 2:nx + 2] - d[1:ny + 1, 2:nx + 2]) / (2.0 * dy)
    with np.errstate(over="ignore", invalid="ignore"):
        hm = dc + hmid + e[I]
        dhu = -(g / cor) * hm * ey
        dhv = (g / cor) * hm * ex
        e[I] += dc
        u[I] += dhu
        v[I] += dhv
    e[I][dry] = 0.0
    u[I][dry] = 0.0
    v[I][dry] = 0.0
    return e, u, v


# --------------------------------------------------------------------------
# bicubic interpolation + balance, odd factor > 1 (ocean_noise.cu:522-772)
# --------------------------------------------------------------------------
_B = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [-3, 3, -2, -1], [2, -2, 1, 1]], dtype=np.float64)


def coarse_cell(loop_t, ghost, off, d, cd, cn):
    """Coarse cell and relative position of fine ghosted index loop_t, in the
    float32 arithmetic of ocean_noise.cu:614-637."""
    f32 = np.float32
    x = (np.asarray(loop_t - ghost).astype(f32) + f32(0.5) + f32(off)) * f32(d)
    ci = np.floor((x / f32(cd)) + f32(ghost) - f32(0.5)).astype(np.int64)
    ci = np.where(ci > cn + 1, ci - 1, ci)
    cx = ((ci - ghost).astype(f32) + f32(0.5)) * f32(cd)
    rel = (x - cx) / f32(cd)
    return ci, rel.astype(np.float64)


def bicubic_eta(coarse, nx, ny, dx, dy, cnx, cny, cdx, cdy, offset_i=0, offset_j=0):
    """d_eta on fine ghosted cells 1..nx+2 x 1..ny+2 (the interior and one
    ring), shape (ny+2, nx+2); coarse reads clamp to the coarse buffer as the
    reference's shared-memory load does (:598-606)."""
    C = np.asarray(coarse, dtype=np.float64)
    ci, rx = coarse_cell(np.arange(1, nx + 3), 2, offset_i, dx, cdx, cnx)
    cj, ry = coarse_cell(np.arange(1, ny + 3), 2, offset_j, dy, cdy, cny)
    CI = ci[None, :]
    CJ = cj[:, None]

    def at(dj, di):
        return C[np.clip(CJ + dj, 0, cny + 3), np.clip(CI + di, 0, cnx + 3)]

    f00, f01, f10, f11 = at(0, 0), at(1, 0), at(0, 1), at(1, 1)
    fx00 = (at(0, 1) - at(0, -1)) / 2
    fx01 = (at(1, 1) - at(1, -1)) / 2
    fx10 = (at(0, 2) - at(0, 0)) / 2
    fx11 = (at(1, 2) - at(1, 0)) / 2
    fy00 = (at(1, 0) - at(-1, 0)) / 2
    fy01 = (at(2, 0) - at(0, 0)) / 2
    fy10 = (at(1, 1) - at(-1, 1)) / 2
    fy11 = (at(2, 1) - at(0, 1)) / 2
    fy_10 = (at(1, -1) - at(-1, -1)) / 2
    fy_11 = (at(2, -1) - at(0, -1)) / 2
    fy20 = (at(1, 2) - at(-1, 2)) / 2
    fy21 = (at(2, 2) - at(0, 2)) / 2
    fxy00 = (fy10 - fy_10) / 2
    fxy01 = (fy11 - fy_11) / 2
    fxy10 = (fy20 - fy00) / 2
    fxy11 = (fy21 - fy01) / 2
    F = np.array([[f00, f01, fy00, fy01],
                  [f10, f11, fy10, fy11],
                  [fx00, fx01, fxy00, fxy01],
                  [fx10, fx11, fxy10, fxy11]])            # (4, 4, ny+2, nx+2)
    A = np.einsum("ik,kl...,jl->ij...", _B, F, _B)         # B F B^T
    X = np.stack([np.ones_like(rx), rx, rx * rx, rx * rx * rx])[:, None, :]
    Y = np.stack([np.ones_like(ry), ry, ry * ry, ry * ry * ry])[:, :, None]
    return np.einsum("i...,ij...,j...->...", X, A, Y)


def bicubic_balance(coarse, eta, hu, hv, Hi, nx, ny, dx, dy, g, f, beta=0.0,
                    north=(0.0, 1.0), f_field=None, land=1e20, factor=3,
                    offset_i=0, offset_j=0):
    """eta, hu, hv after adding the interpolated d_eta and its balanced
    (d_hu, d_hv) on the interior; zero perturbation on cells with a land
    corner (ocean_noise.cu:720-770)."""
    cnx, cny = nx // factor, ny // factor
    cdx, cdy = float(np.float32(dx * factor)), float(np.float32(dy * factor))
    d = bicubic_eta(coarse, nx, ny, dx, dy, cnx, cny, cdx, cdy, offset_i, offset_j)
    e = np.array(eta, dtype=np.float64)
    u = np.array(hu, dtype=np.float64)
    v = np.array(hv, dtype=np.float64)
    dry, hmid = _land_and_hmid(Hi, nx, ny, land)
    cor = _coriolis(nx, ny, dx, dy, f, beta, north, f_field)
    dc = d[1:ny + 1, 1:nx + 1]
    ex = (d[1:ny + 1, 2:nx + 2] - d[1:ny + 1, 0:nx]) / (2.0 * dx)
    ey = (d[2:ny + 2, 1:nx + 1] - d[0:ny, 1:nx + 1]) / (2.0 * dy)
    with np.errstate(over="ignore", invalid="ignore"):
        dhu = -(g / cor) * (hmid + dc) * ey
        dhv = (g / cor) * (hmid + dc) * ex
    dc = np.where(dry, 0.0, dc)
    dhu = np.where(dry, 0.0, dhu)
    dhv = np.where(dry, 0.0, dhv)
    I = (slice(2, ny + 2), slice(2, nx + 2))
    e[I] += dc
    u[I] += dhu
    v[I] += dhv
    return e, u, v


def cpu_twin_interpolate(coarse_eta, nx, ny, dx, dy, factor):
    """OceanStateNoise._interpolate_CPU + _bicubic_interpolation_inner
    (OceanStateNoise.py:784-890): d_eta of shape (ny+4, nx+4), filled on the
    interior and the first ghost ring, in float64."""
    cdx, cdy = dx * factor, dy * factor
    d_eta = np.zeros((ny + 4, nx + 4))
    for loc_j in range(ny + 2):
        for loc_i in range(nx + 2):
            i, j = loc_i + 1, loc_j + 1
            x = (i - 2 + 0.5) * dx
            y = (j - 2 + 0.5) * dy
            ci = int(np.floor(x / cdx + 2 - 0.5))
            cj = int(np.floor(y / cdy + 2 - 0.5))
            rel_x = (x - (ci - 2 + 0.5) * cdx) / cdx
            rel_y = (y - (cj - 2 + 0.5) * cdy) / cdy
            c = coarse_eta
            fy = lambda jj, ii: (c[jj + 1, ii] - c[jj - 1, ii]) / 2
            F = np.array([
                [c[cj, ci], c[cj + 1, ci], fy(cj, ci), fy(cj + 1, ci)],
                [c[cj, ci + 1], c[cj + 1, ci + 1], fy(cj, ci + 1), fy(cj + 1, ci + 1)],
                [(c[cj, ci + 1] - c[cj, ci - 1]) / 2, (c[cj + 1, ci + 1] - c[cj + 1, ci - 1]) / 2,
                 (fy(cj, ci + 1) - fy(cj, ci - 1)) / 2, (fy(cj + 1, ci + 1) - fy(cj + 1, ci - 1)) / 2],
                [(c[cj, ci + 2] - c[cj, ci]) / 2, (c[cj + 1, ci + 2] - c[cj + 1, ci]) / 2,
                 (fy(cj, ci + 2) - fy(cj, ci)) / 2, (fy(cj + 1, ci + 2) - fy(cj + 1, ci)) / 2]])
            A = _B @ F @ _B.T
            xv = np.array([1.0, rel_x, rel_x ** 2, rel_x ** 3])
            yv = np.array([1.0, rel_y, rel_y ** 2, rel_y ** 3])
            d_eta[j, i] = xv @ A @ yv
    return d_eta


def cpu_twin_perturbation(random, H, nx, ny, dx, dy, factor, q0, L, f, beta, g,
                          periodic_ns, periodic_ew):
    """OceanStateNoise._obtainOceanPerturbations_CPU (OceanStateNoise.py:763-781):
    (d_eta (ny+4, nx+4), d_hu (ny, nx), d_hv (ny, nx)); H is (ny+5, nx+5)."""
    cnx, cny = nx // factor, ny // factor
    d_eta = cpu_twin_applyQ(random, cnx, cny, dx * factor, dy * factor, q0, L,
                            periodic_ns, periodic_ew)
    if factor > 1:
        d_eta = cpu_twin_interpolate(d_eta, nx, ny, dx, dy, factor)
    H = np.asarray(H, dtype=np.float64)
    d_hu = np.zeros((ny, nx))
    d_hv = np.zeros((ny, nx))
    for j in range(ny):
        lj = j + 2
        cor = f + beta * lj * dy
        for i in range(nx):
            li = i + 2
            hmid = d_eta[lj, li] + 0.25 * (H[j, i] + H[j + 1, i] + H[j, i + 1] + H[j + 1, i + 1])
            d_hu[j, i] = -(g / cor) * hmid * (d_eta[lj + 1, li] - d_eta[lj - 1, li]) / (2.0 * dy)
            d_hv[j, i] = (g / cor) * hmid * (d_eta[lj, li + 1] - d_eta[lj, li - 1]) / (2.0 * dx)
    return d_eta, d_hu, d_hv


# --------------------------------------------------------------------------
# norms and perpendicular pair (reductions.cu:27-185; ocean_noise.cu:206-243)
# --------------------------------------------------------------------------
def norms(xi, nu=None):
    xi = np.asarray(xi, dtype=np.float64)
    if nu is None:
        return np.array([np.sum(xi * xi)])
    nu = np.asarray(nu, dtype=np.float64)
    return np.array([np.sum(xi * xi), np.sum(nu * nu), np.sum(xi * nu)])


def make_perpendicular(xi, nu, red):
    xi = np.asarray(xi, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    par = red[2] / red[0]
    perp = red[1] - par * red[2]
    return np.sqrt(red[1] / perp) * (nu - par * xi)


def coarse_grid_offset(factor, fine_i=None, fine_j=None):
    """OceanStateNoise._obtain_coarse_grid_offset (OceanStateNoise.py:498-512)."""
    default = factor // 2
    oi = 0 if fine_i is None else -int(fine_i % factor - default)
    oj = 0 if fine_j is None else -int(fine_j % factor - default)
    return oi, oj
