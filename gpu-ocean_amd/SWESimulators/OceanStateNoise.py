"""Model error for the CDKLM16 state on MI355X (OceanStateNoise).

Keeps the surface of the reference class
(gpu_ocean/SWESimulators/OceanStateNoise.py:40-570): the constructor and
``fromsim``, ``perturbSim`` / ``perturbOceanState``, the generators
(``generateNormalDistribution``, ``generateUniformDistribution``,
``generateNormalDistributionPerpendicular``,
``generatePerpendicularNormalDistributions``), the norms
(``getRandomNorm``, ``findDoubleNormAndDot``), the seed handling and the
buffer getters.  Every device operation is a call into
``lib/libgocean_hip.so`` (``gocean_noise_*``).

Differences from the reference, by design:
  * ``use_lcg=False`` draws from a Philox4x32-10 counter-based generator
    (keyed by ``xorwow_seed`` when given) instead of pycuda's XORWOW stream;
    ``use_lcg=True`` reproduces the reference's LCG + Box-Muller bit for bit
    in the seeds.
  * ``angle`` / ``coriolis_f`` are per-cell fields at full resolution (or
    constants), not textures.
  * ``ghost_cells_x/y`` and ``block_width/height`` only shaped the
    reference's thread grid; every interior cell is perturbed here.
  * The ``*CPU`` twins of the reference are test helpers; their
    restatement lives with the tests' oracle, not in this class.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import Common
from . import _gocean as G


class OceanStateNoise(object):
    """SOAR-correlated eta perturbations with geostrophically balanced
    (hu, hv) (OceanStateNoise.py:40-47)."""

    def __init__(self, gpu_ctx, gpu_stream, nx, ny, dx, dy, boundaryConditions, staggered,
                 soar_q0=None, soar_L=None, interpolation_factor=1, use_lcg=False,
                 xorwow_seed=None, angle=np.array([[0]], dtype=np.float32),
                 coriolis_f=np.array([[0]], dtype=np.float32), block_width=16, block_height=16):
        torch = Common._torch()
        self.lib = G.load()
        self.use_lcg = use_lcg
        self.random_state = np.random.RandomState()
        self.random_numbers = None
        self.rng = None
        self.seed = None
        self.host_seed = None
        self.gpu_ctx = gpu_ctx
        self.gpu_stream = gpu_stream
        self.device = gpu_ctx.torch_device if gpu_ctx is not None else torch.device("cuda")

        self.nx = np.int32(nx)
        self.ny = np.int32(ny)
        self.dx = np.float32(dx)
        self.dy = np.float32(dy)
        self.staggered = 1 if staggered else 0
        self.cutoff = np.int32(2)          # config.soar_cutoff
        assert interpolation_factor > 0 and interpolation_factor % 2 == 1, \
            'interpolation_factor must be a positive odd integer'
        assert nx % interpolation_factor == 0, 'nx must be divisible by the interpolation factor'
        assert ny % interpolation_factor == 0, 'ny must be divisible by the interpolation factor'
        self.interpolation_factor = np.int32(interpolation_factor)
        self.coarse_nx = np.int32(nx // interpolation_factor)
        self.coarse_ny = np.int32(ny // interpolation_factor)
        self.coarse_dx = np.float32(dx * self.interpolation_factor)
        self.coarse_dy = np.float32(dy * self.interpolation_factor)
        self.periodicNorthSouth = np.int32(boundaryConditions.isPeriodicNorthSouth())
        self.periodicEastWest = np.int32(boundaryConditions.isPeriodicEastWest())

        self.soar_q0 = np.float32(self.dx / 100000)
        if soar_q0 is not None:
            self.soar_q0 = np.float32(soar_q0)
        self.soar_L = np.float32(0.75 * self.coarse_dx)
        if soar_L is not None:
            self.soar_L = np.float32(soar_L)

        self.desc = G.Noise()
        d = self.desc
        d.nx, d.ny = int(nx), int(ny)
        d.pitch, d.hi_pitch = int(nx) + 4, int(nx) + 5
        d.dx, d.dy = float(self.dx), float(self.dy)
        d.interpolation_factor = int(interpolation_factor)
        d.periodic_ns, d.periodic_ew = int(self.periodicNorthSouth), int(self.periodicEastWest)
        d.soar_q0, d.soar_L = float(self.soar_q0), float(self.soar_L)
        d.north_x, d.north_y = 0.0, 1.0
        d.land_value = float(np.float32(1.0e20))
        sizes = (ctypes.c_int32 * 8)()
        G.check(self.lib.gocean_noise_sizes(d, sizes), "gocean_noise_sizes")
        (_, _, rnx, rny, snx, sny, gx, gy) = list(sizes)
        self.rand_ghost_cells_x = np.int32(gx)
        self.rand_ghost_cells_y = np.int32(gy)
        self.rand_nx = np.int32(rnx)
        self.rand_ny = np.int32(rny)
        self.seed_nx = np.int32(snx)
        self.seed_ny = np.int32(sny)

        self.floatMax = 2147483648.0
        if self.use_lcg:
            self.host_seed = (self.random_state.rand(self.seed_ny, self.seed_nx) * self.floatMax).astype(np.uint64)
            with torch.cuda.stream(gpu_stream):
                self.seed = torch.from_numpy(self.host_seed.astype(np.int64)).to(self.device)
        else:
            key = xorwow_seed if xorwow_seed is not None else int(self.random_state.randint(0, 2 ** 31)) \
                | (int(self.random_state.randint(0, 2 ** 31)) << 31)
            self.rng = _Philox(int(key) & 0xFFFFFFFFFFFFFFFF, self.device, gpu_stream)

        z = np.zeros((self.rand_ny, self.rand_nx), dtype=np.float32)
        self.random_numbers_host = z
        self.random_numbers = Common.DeviceArray2D(gpu_stream, self.rand_nx, self.rand_ny, 0, 0, z,
                                                   device=self.device)
        self.perpendicular_random_numbers = Common.DeviceArray2D(gpu_stream, self.rand_nx, self.rand_ny,
                                                                 0, 0, z, device=self.device)
        cz = np.zeros((self.coarse_ny + 4, self.coarse_nx + 4), dtype=np.float32)
        self.coarse_buffer = Common.DeviceArray2D(gpu_stream, self.coarse_nx, self.coarse_ny, 2, 2, cz,
                                                  device=self.device)
        self.reduction_buffer = Common.DeviceArray2D(gpu_stream, 3, 1, 0, 0,
                                                     np.zeros((1, 3), dtype=np.float32), device=self.device)

        self.north_const, self.north_fields = _resolve_north(angle)
        self.coriolis_field = _resolve_field(coriolis_f, self.device, gpu_stream)
        self.fields = G.Fields()
        if self.north_fields is not None:
            self.fields.north_x, self.fields.north_y = (t.data_ptr() for t in self.north_fields)
        if self.coriolis_field is not None:
            self.fields.f = self.coriolis_field.data_ptr()

    def __del__(self):
        self.cleanUp()

    def cleanUp(self):
        self.rng = None
        for name in ("random_numbers", "perpendicular_random_numbers", "coarse_buffer", "reduction_buffer"):
            b = getattr(self, name, None)
            if b is not None:
                b.release()
        self.seed = None
        self.gpu_ctx = None

    @classmethod
    def fromsim(cls, sim, soar_q0=None, soar_L=None, interpolation_factor=1, use_lcg=False,
                xorwow_seed=None, block_width=16, block_height=16):
        """OceanStateNoise.py:289-302 with the simulator's north vector and
        Coriolis field."""
        angle = sim.angle_texref if sim.angle_texref is not None else tuple(sim.north_const)
        return cls(sim.gpu_ctx, sim.gpu_stream, sim.nx, sim.ny, sim.dx, sim.dy,
                   sim.boundary_conditions, False, soar_q0=soar_q0, soar_L=soar_L,
                   interpolation_factor=interpolation_factor, angle=angle,
                   coriolis_f=sim.coriolis_texref, use_lcg=use_lcg, xorwow_seed=xorwow_seed,
                   block_width=block_width, block_height=block_height)

    # ------------------------------------------------------------------
    def _s(self, stream=None):
        return G.stream_handle(self.gpu_stream if stream is None else stream)

    def getSeed(self):
        assert self.use_lcg, "getSeed is only valid if LCG is used as pseudo-random generator."
        torch = Common._torch()
        with torch.cuda.stream(self.gpu_stream):
            return self.seed.to("cpu").numpy().astype(np.uint64)

    def resetSeed(self):
        assert self.use_lcg, "resetSeed is only valid if LCG is used as pseudo-random generator."
        torch = Common._torch()
        self.host_seed = (self.random_state.rand(self.seed_ny, self.seed_nx) * self.floatMax).astype(np.uint64)
        with torch.cuda.stream(self.gpu_stream):
            self.seed.copy_(torch.from_numpy(self.host_seed.astype(np.int64)))

    def setSeed(self, seed):
        """Upload LCG seeds (seed_ny x seed_nx, values below 2^31)."""
        assert self.use_lcg, "setSeed is only valid if LCG is used as pseudo-random generator."
        torch = Common._torch()
        s = np.asarray(seed, dtype=np.uint64)
        assert s.shape == (self.seed_ny, self.seed_nx)
        self.host_seed = s.copy()
        with torch.cuda.stream(self.gpu_stream):
            self.seed.copy_(torch.from_numpy(s.astype(np.int64)))

    def getRandomNumbers(self):
        return self.random_numbers.download(self.gpu_stream)

    def getPerpendicularRandomNumbers(self):
        return self.perpendicular_random_numbers.download(self.gpu_stream)

    def getCoarseBuffer(self):
        return self.coarse_buffer.download(self.gpu_stream)

    def getReductionBuffer(self):
        return self.reduction_buffer.download(self.gpu_stream)

    def _generate(self, buf, normal):
        if self.use_lcg:
            kind = G.RNG_LCG_NORMAL if normal else G.RNG_LCG_UNIFORM
            G.check(self.lib.gocean_noise_generate(self.desc, kind, self.seed.data_ptr(), buf.data.gpudata,
                                                   0, None, self._s()), "gocean_noise_generate")
        else:
            kind = G.RNG_PHILOX_NORMAL if normal else G.RNG_PHILOX_UNIFORM
            G.check(self.lib.gocean_noise_generate(self.desc, kind, None, buf.data.gpudata,
                                                   self.rng.key, self.rng.counter.data_ptr(), self._s()),
                    "gocean_noise_generate")

    def generateNormalDistribution(self):
        self._generate(self.random_numbers, True)

    def generateNormalDistributionPerpendicular(self):
        self._generate(self.perpendicular_random_numbers, True)

    def generateUniformDistribution(self):
        self._generate(self.random_numbers, False)

    # ------------------------------------------------------------------
    def perturbSim(self, sim, q0_scale=1.0, update_random_field=True, perturbation_scale=1.0,
                   perpendicular_scale=0.0, align_with_cell_i=None, align_with_cell_j=None,
                   stream=None):
        """Add a perturbation to sim's current state (OceanStateNoise.py:376-396)."""
        self.perturbOceanState(sim.gpu_data.h0, sim.gpu_data.hu0, sim.gpu_data.hv0,
                               sim.bathymetry.Bi, sim.f, beta=sim.coriolis_beta, g=sim.g,
                               y0_reference_cell=sim.y_zero_reference_cell,
                               ghost_cells_x=sim.ghost_cells_x, ghost_cells_y=sim.ghost_cells_y,
                               q0_scale=q0_scale, update_random_field=update_random_field,
                               perturbation_scale=perturbation_scale,
                               perpendicular_scale=perpendicular_scale,
                               align_with_cell_i=align_with_cell_i,
                               align_with_cell_j=align_with_cell_j,
                               land_mask_value=sim.bathymetry.mask_value, stream=stream)

    def perturbOceanState(self, eta, hu, hv, H, f, beta=0.0, g=9.81, y0_reference_cell=0,
                          ghost_cells_x=0, ghost_cells_y=0, q0_scale=1.0, update_random_field=True,
                          perturbation_scale=1.0, perpendicular_scale=0.0,
                          align_with_cell_i=None, align_with_cell_j=None,
                          land_mask_value=np.float32(1.0e20), stream=None):
        """Add the SOAR perturbation of eta and its geostrophic (hu, hv) to the
        interior of eta, hu, hv (OceanStateNoise.py:399-496).  eta/hu/hv are
        (ny+4, nx+4) device arrays, H the (ny+5, nx+5) corner depths."""
        s = self._s(stream)
        if update_random_field:
            self.generateNormalDistribution()
        soar_q0 = np.float32(self.soar_q0 * q0_scale)
        offset_i, offset_j = self._obtain_coarse_grid_offset(align_with_cell_i, align_with_cell_j)
        d = self.desc
        d.pitch = _pitch(eta)
        d.hi_pitch = _pitch(H)
        d.g = float(np.float32(g))
        d.f = float(np.float32(np.asarray(f).flat[0])) if np.size(f) == 1 else 0.0
        d.beta = float(np.float32(beta))
        d.north_x, d.north_y = self.north_const
        d.land_value = float(np.float32(land_mask_value))
        scale = float(np.float32(perturbation_scale))
        if self.interpolation_factor == 1 and not perpendicular_scale > 0:
            # SOAR and geostrophicBalance in one launch
            G.check(self.lib.gocean_noise_soar_balance(
                d, float(soar_q0), scale, self.random_numbers.data.gpudata, self.coarse_buffer.data.gpudata,
                _addr(eta), _addr(hu), _addr(hv), _addr(H), self.fields, s), "gocean_noise_soar_balance")
            return
        G.check(self.lib.gocean_noise_soar(d, float(soar_q0), scale, self.random_numbers.data.gpudata,
                                           self.coarse_buffer.data.gpudata, 0, s), "gocean_noise_soar")
        if perpendicular_scale > 0:
            G.check(self.lib.gocean_noise_soar(d, float(soar_q0), float(np.float32(perpendicular_scale)),
                                               self.perpendicular_random_numbers.data.gpudata,
                                               self.coarse_buffer.data.gpudata, 1, s), "gocean_noise_soar")
        G.check(self.lib.gocean_noise_balance(d, self.coarse_buffer.data.gpudata, _addr(eta), _addr(hu),
                                              _addr(hv), _addr(H), self.fields, int(offset_i),
                                              int(offset_j), s), "gocean_noise_balance")

    def _obtain_coarse_grid_offset(self, fine_index_i, fine_index_j):
        """OceanStateNoise.py:498-512."""
        default_offset = self.interpolation_factor // 2
        offset_i, offset_j = 0, 0
        if fine_index_i is not None:
            offset_i = -int(fine_index_i % self.interpolation_factor - default_offset)
        if fine_index_j is not None:
            offset_j = -int(fine_index_j % self.interpolation_factor - default_offset)
        return offset_i, offset_j

    # ------------------------------------------------------------------
    def getRandomNorm(self):
        """sum(xi^2) of the random buffer (OceanStateNoise.py:515-527)."""
        G.check(self.lib.gocean_noise_norms(self.desc, self.random_numbers.data.gpudata, None,
                                            self.reduction_buffer.data.gpudata, self._s()),
                "gocean_noise_norms")
        return self.getReductionBuffer()[0, 0]

    def findDoubleNormAndDot(self):
        """[xi.xi, nu.nu, xi.nu] into the reduction buffer (OceanStateNoise.py:530-541)."""
        G.check(self.lib.gocean_noise_norms(self.desc, self.random_numbers.data.gpudata,
                                            self.perpendicular_random_numbers.data.gpudata,
                                            self.reduction_buffer.data.gpudata, self._s()),
                "gocean_noise_norms")

    def _makePerpendicular(self):
        """nu <- perpendicular to xi, same norm (OceanStateNoise.py:543-556)."""
        G.check(self.lib.gocean_noise_make_perpendicular(
            self.desc, self.random_numbers.data.gpudata, self.perpendicular_random_numbers.data.gpudata,
            self.reduction_buffer.data.gpudata, self._s()), "gocean_noise_make_perpendicular")

    def generatePerpendicularNormalDistributions(self):
        """xi, nu ~ N(0, I) with nu perpendicular to xi (OceanStateNoise.py:558-570)."""
        self.generateNormalDistribution()
        self.generateNormalDistributionPerpendicular()
        self.findDoubleNormAndDot()
        self._makePerpendicular()


class _Philox(object):
    """Key and device counter of the counter-based generator."""

    def __init__(self, key, device, stream):
        torch = Common._torch()
        self.key = key
        with torch.cuda.stream(stream):
            self.counter = torch.zeros(1, dtype=torch.int64, device=device)


def _addr(a):
    if hasattr(a, "data") and hasattr(a.data, "gpudata"):
        return a.data.gpudata
    return G.ptr(a)


def _pitch(a):
    if hasattr(a, "pitch_elems"):
        return int(a.pitch_elems)
    return int(a.stride(0))


def _resolve_north(angle):
    """(constant north, per-cell north fields or None) from an angle array,
    a (north_x, north_y) pair of constants or a pair of device fields."""
    if isinstance(angle, tuple) and len(angle) == 2:
        if hasattr(angle[0], "data_ptr"):
            return (0.0, 1.0), (angle[0], angle[1])
        return (float(np.float32(angle[0])), float(np.float32(angle[1]))), None
    a = np.asarray(angle)
    if a.size == 1:
        a32 = float(np.float32(a.flat[0]))
        return (float(np.float32(np.sin(a32))), float(np.float32(np.cos(a32)))), None
    torch = Common._torch()
    ang = a.astype(np.float32).astype(np.float64)
    nx_f = torch.from_numpy(np.sin(ang).astype(np.float32)).cuda()
    ny_f = torch.from_numpy(np.cos(ang).astype(np.float32)).cuda()
    return (0.0, 1.0), (nx_f, ny_f)


def _resolve_field(field, device, stream):
    if field is None:
        return None
    if hasattr(field, "data_ptr"):
        return field
    a = np.asarray(field, dtype=np.float32)
    if a.size == 1:
        return None
    torch = Common._torch()
    with torch.cuda.stream(stream):
        return torch.from_numpy(np.ascontiguousarray(a)).to(device)
