"""Device context, device arrays, boundary conditions and bathymetry.

Same classes and call signatures as the reference's
``gpu_ocean/SWESimulators/Common.py`` (CUDAContext :285-495, CUDAArray2D
:503-631, SWEDataArakawaA :634-689, BoundaryConditions :824-943, Bathymetry
:1313-1459), backed by PyTorch-ROCm device tensors and the HIP library in
``lib/libgocean_hip.so``.  PyTorch only allocates memory and owns streams;
every kernel is the library's.
"""

from __future__ import annotations

import logging
import time

import numpy as np

from . import _gocean as G

WALL, PERIODIC, FRS, LINEAR = 1, 2, 3, 4


def _torch():
    import torch
    return torch


class Timer(object):
    """Context manager logging the elapsed milliseconds (Common.py:53-70)."""

    def __init__(self, tag, log_level=logging.DEBUG):
        self.logger = logging.getLogger(__name__)
        self.tag = tag
        self.log_level = log_level

    def __enter__(self):
        self.start = time.time()
        return self

    def __exit__(self, *args):
        self.end = time.time()
        self.secs = self.end - self.start
        self.msecs = self.secs * 1000
        self.logger.log(self.log_level, "%s: %f ms", self.tag, self.msecs)


class HIPContext(object):
    """Device selection and synchronisation (CUDAContext, Common.py:285-340).

    There is no run-time kernel compilation: the kernels are built ahead of
    time into ``libgocean_hip.so``.  ``device`` is taken modulo the device
    count, as the reference does (Common.py:304-306)."""

    def __init__(self, device=0, blocking=False, use_cache=True):
        torch = _torch()
        self.logger = logging.getLogger(__name__)
        self.blocking = blocking
        self.use_cache = use_cache
        if not torch.cuda.is_available():
            raise RuntimeError("No HIP device is visible; the CDKLM16 HIP kernels need one.")
        n = torch.cuda.device_count()
        if device >= n:
            print("Changing device ID from " + str(device) + " to " + str(device % n))
            device = device % n
        self.device = device
        self.torch_device = torch.device("cuda", device)
        self.lib = G.load()
        self.logger.info("Using '%s' (%s)", torch.cuda.get_device_name(device),
                         self.lib.gocean_version().decode())

    def synchronize(self):
        _torch().cuda.synchronize(self.torch_device)

    def new_stream(self):
        return _torch().cuda.Stream(device=self.torch_device)

    def get_kernel(self, *args, **kwargs):
        raise NotImplementedError(
            "kernels are compiled ahead of time into libgocean_hip.so; there is no JIT")

    def __str__(self):
        return "HIPContext id " + str(id(self))


# drop-in name used by the reference's callers
CUDAContext = HIPContext


class _DevicePointer(object):
    """Stands in for a pycuda GPUArray where callers read ``.data.gpudata``."""

    def __init__(self, tensor):
        self.tensor = tensor

    @property
    def gpudata(self):
        return self.tensor.data_ptr()

    @property
    def ptr(self):
        return self.tensor.data_ptr()

    def get(self):
        return self.tensor.cpu().numpy()


class DeviceArray2D:
    """Pitched 2-D float32 device array (CUDAArray2D, Common.py:503-631).

    ``pitch`` is in bytes and equals ``nx_halo*4`` (rows are unpadded, as in
    the reference).  ``data.gpudata`` is the device address."""

    def __init__(self, gpu_stream, nx, ny, halo_x, halo_y, data,
                 asymHalo=None, double_precision=False, integers=False, device=None):
        torch = _torch()
        self.double_precision = double_precision
        self.integers = integers
        host_data = data if double_precision else self.convert_to_float32(data)
        self.nx = nx
        self.ny = ny
        self.nx_halo = nx + 2 * halo_x
        self.ny_halo = ny + 2 * halo_y
        if asymHalo is not None and len(asymHalo) == 4:
            self.nx_halo = nx + asymHalo[1] + asymHalo[3]
            self.ny_halo = ny + asymHalo[0] + asymHalo[2]
        assert host_data.shape[1] == self.nx_halo, str(host_data.shape[1]) + " vs " + str(self.nx_halo)
        assert host_data.shape[0] == self.ny_halo, str(host_data.shape[0]) + " vs " + str(self.ny_halo)
        self.stream = gpu_stream
        self.device = device if device is not None else (
            gpu_stream.device if gpu_stream is not None else torch.device("cuda"))
        self.mask = None
        if np.ma.is_masked(data):
            self.mask = data.mask
        arr = np.ascontiguousarray(np.ma.getdata(host_data))
        dtype = torch.float64 if double_precision else torch.float32
        with torch.cuda.stream(gpu_stream) if gpu_stream is not None else _nullcontext():
            self.tensor = torch.empty((self.ny_halo, self.nx_halo), dtype=dtype, device=self.device)
            self.tensor.copy_(torch.from_numpy(arr), non_blocking=False)
        self.data = _DevicePointer(self.tensor)
        self.holds_data = True
        self.bytes_per_float = arr.itemsize
        self.pitch = np.int32(self.nx_halo * self.bytes_per_float)

    @property
    def pitch_elems(self):
        return int(self.pitch) // self.bytes_per_float

    def upload(self, gpu_stream, data):
        """Fill the allocated buffer with new data (Common.py:560-582)."""
        if not self.holds_data:
            raise RuntimeError('The buffer has been freed before upload is called')
        torch = _torch()
        if np.ma.is_masked(data):
            self.mask = data.mask
        host_data = data if self.double_precision else self.convert_to_float32(data)
        assert host_data.shape == (self.ny_halo, self.nx_halo), \
            str(host_data.shape) + " vs " + str((self.ny_halo, self.nx_halo))
        arr = np.ascontiguousarray(np.ma.getdata(host_data))
        with torch.cuda.stream(gpu_stream) if gpu_stream is not None else _nullcontext():
            self.tensor.copy_(torch.from_numpy(arr))

    def copyBuffer(self, gpu_stream, buffer):
        """Device-to-device copy of another array of the same shape
        (Common.py:585-603)."""
        if not self.holds_data:
            raise RuntimeError('The buffer has been freed before copying buffer')
        if not buffer.holds_data:
            raise RuntimeError('The provided buffer is either not allocated, or has been freed before copying buffer')
        assert buffer.nx_halo == self.nx_halo, str(buffer.nx_halo) + " vs " + str(self.nx_halo)
        assert buffer.ny_halo == self.ny_halo, str(buffer.ny_halo) + " vs " + str(self.ny_halo)
        assert buffer.bytes_per_float == self.bytes_per_float
        torch = _torch()
        with torch.cuda.stream(gpu_stream) if gpu_stream is not None else _nullcontext():
            self.tensor.copy_(buffer.tensor)

    def download(self, gpu_stream):
        """Copy to a host numpy array (masked where land, Common.py:606-620)."""
        if not self.holds_data:
            raise RuntimeError('CUDA buffer has been freed')
        torch = _torch()
        with torch.cuda.stream(gpu_stream) if gpu_stream is not None else _nullcontext():
            host = self.tensor.to("cpu").numpy()
        if self.mask is not None:
            host = np.ma.array(host, mask=self.mask)
        return host

    def release(self):
        if self.holds_data:
            self.tensor = None
            self.data = None
            self.holds_data = False

    @staticmethod
    def convert_to_float32(data):
        """C-ordered float32 view/copy (Common.py:624-631)."""
        if not np.issubdtype(data.dtype, np.float32) or np.isfortran(data):
            return data.astype(np.float32, order='C')
        return data


CUDAArray2D = DeviceArray2D


class _nullcontext:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


class SWEDataArakawaA:
    """Double-buffered A-grid state (Common.py:634-689)."""

    def __init__(self, gpu_stream, nx, ny, halo_x, halo_y, h0, hu0, hv0, device=None):
        self.h0 = DeviceArray2D(gpu_stream, nx, ny, halo_x, halo_y, h0, device=device)
        self.hu0 = DeviceArray2D(gpu_stream, nx, ny, halo_x, halo_y, hu0, device=device)
        self.hv0 = DeviceArray2D(gpu_stream, nx, ny, halo_x, halo_y, hv0, device=device)
        self.h1 = DeviceArray2D(gpu_stream, nx, ny, halo_x, halo_y, h0, device=device)
        self.hu1 = DeviceArray2D(gpu_stream, nx, ny, halo_x, halo_y, hu0, device=device)
        self.hv1 = DeviceArray2D(gpu_stream, nx, ny, halo_x, halo_y, hv0, device=device)

    def swap(self):
        self.h1, self.h0 = self.h0, self.h1
        self.hu1, self.hu0 = self.hu0, self.hu1
        self.hv1, self.hv0 = self.hv0, self.hv1

    def download(self, gpu_stream):
        return (self.h0.download(gpu_stream), self.hu0.download(gpu_stream),
                self.hv0.download(gpu_stream))

    def downloadPrevTimestep(self, gpu_stream):
        return (self.h1.download(gpu_stream), self.hu1.download(gpu_stream),
                self.hv1.download(gpu_stream))

    def release(self):
        for a in (self.h0, self.hu0, self.hv0, self.h1, self.hu1, self.hv1):
            a.release()


class BoundaryConditions:
    """Boundary condition codes per side (Common.py:824-943): 1 wall,
    2 periodic (must be paired), 3 flow relaxation, 4 linear interpolation.
    The HIP step implements 1 and 2; 3 and 4 are rejected by the simulator."""

    def __init__(self, north=1, east=1, south=1, west=1,
                 spongeCells={'north': 0, 'south': 0, 'east': 0, 'west': 0}):
        self.north = np.int32(north)
        self.east = np.int32(east)
        self.south = np.int32(south)
        self.west = np.int32(west)
        self.spongeCells = dict(spongeCells)
        for key in self.spongeCells.keys():
            self.spongeCells[key] = np.int32(self.spongeCells[key])
        assert not ((self.north == 2 or self.south == 2) and (self.north != self.south)), \
            'The given periodic boundary conditions are not periodically (north/south)'
        assert not ((self.east == 2 or self.west == 2) and (self.east != self.west)), \
            'The given periodic boundary conditions are not periodically (east/west)'

    def get(self):
        raise NotImplementedError("get() implementation changed! Change calling code")

    def getSponge(self):
        return self.spongeCells

    def isDefault(self):
        return (self.north == 1 and self.east == 1 and self.south == 1 and self.east == 1)

    def isSponge(self):
        return any(v in (3, 4) for v in (self.north, self.east, self.south, self.west))

    def isPeriodicNorthSouth(self):
        return self.north == 2 and self.south == 2

    def isPeriodicEastWest(self):
        return self.east == 2 and self.west == 2

    def _toString(self, cond):
        return {1: "Wall", 2: "Periodic", 3: "Flow_Relaxation_Scheme",
                4: "Open_Linear_Interpolation"}.get(int(cond), "Invalid_:|")

    def __str__(self):
        msg = "north: " + self._toString(self.north) + \
              ", east: " + self._toString(self.east) + \
              ", south: " + self._toString(self.south) + \
              ", west: " + self._toString(self.west)
        # plain ints, as numpy 1.x printed them (numpy 2 prints np.int32(...))
        return msg + ", spongeCells: " + str({k: int(v) for k, v in self.spongeCells.items()})

    @classmethod
    def fromstring(cls, bc_string):
        def keyword_to_cond(key):
            return {'Wall': 1, 'Periodic': 2, 'Flow_Relaxation_Scheme': 3,
                    'Open_Linear_Interpolation': 4}.get(key, -1)
        clean = bc_string
        for ch in [',', '}', '{', '[', ']', ':', "'"]:
            clean = clean.replace(ch, '')
        a = clean.split(' ')
        north, east, south, west = (keyword_to_cond(a[1]), keyword_to_cond(a[3]),
                                    keyword_to_cond(a[5]), keyword_to_cond(a[7]))
        if len(a) > 13:
            sponge = {a[9]: int(a[10]), a[11]: int(a[12]), a[13]: int(a[14]), a[15]: int(a[16])}
        else:
            sponge = {'north': int(a[9]), 'east': int(a[10]), 'south': int(a[11]), 'west': int(a[12])}
        return cls(north=north, east=east, south=south, west=west, spongeCells=sponge)

    def codes(self):
        return (int(self.north), int(self.east), int(self.south), int(self.west))


class BoundaryConditionsData:
    """Time series for open boundaries (Common.py:945-1016).  Only the empty
    default is accepted: open boundaries are not part of this build."""

    def __init__(self, t=None, north=None, south=None, east=None, west=None):
        self.t = [0] if t is None else t
        self.north, self.south, self.east, self.west = north, south, east, west
        self.numSteps = len(self.t)


def make_domain(nx, ny, dx, dy, g, r, theta, rk_order, bc, pitch=None, hi_pitch=None,
                f=0.0, north_x=0.0, north_y=1.0, desing_eps=0.1, flux_slope_eps=0.1,
                depth_cutoff=1e-5, land_value=1.0e20):
    """Fill a ``gocean_domain_t`` (the reference's #define set,
    CDKLM16.py:179-191, plus the per-domain run-time values)."""
    d = G.Domain()
    d.nx, d.ny = int(nx), int(ny)
    d.pitch = int(pitch if pitch is not None else nx + 4)
    d.hi_pitch = int(hi_pitch if hi_pitch is not None else nx + 5)
    d.dx, d.dy, d.g, d.r = float(np.float32(dx)), float(np.float32(dy)), float(np.float32(g)), float(np.float32(r))
    d.theta = float(np.float32(theta))
    d.desing_eps = float(np.float32(desing_eps))
    d.flux_slope_eps = float(np.float32(flux_slope_eps))
    d.depth_cutoff = float(np.float32(depth_cutoff))
    d.land_value = float(np.float32(land_value))
    d.f = float(np.float32(f))
    d.north_x = float(np.float32(north_x))
    d.north_y = float(np.float32(north_y))
    d.bc_north, d.bc_east, d.bc_south, d.bc_west = (int(b) for b in bc)
    d.rk_order = int(rk_order)
    return d


class Bathymetry:
    """Corner depths Bi and cell depths Bm on the device (Common.py:1313-1459).

    Masked input marks land with ``mask_value`` (1e20).  The halo of Bi is
    completed (walls mirror, periodic wraps) and Bm computed by the HIP
    library, which replaces the reference's boundary and initBm kernels."""

    def __init__(self, gpu_ctx, gpu_stream, nx, ny, halo_x, halo_y, Bi_host,
                 boundary_conditions=BoundaryConditions(), block_width=16, block_height=16):
        self.gpu_stream = gpu_stream
        self.nx = np.int32(nx)
        self.ny = np.int32(ny)
        self.halo_x = np.int32(halo_x)
        self.halo_y = np.int32(halo_y)
        self.halo_nx = np.int32(nx + 2 * halo_x)
        self.halo_ny = np.int32(ny + 2 * halo_y)
        self.boundary_conditions = boundary_conditions
        self.mask_value = np.float32(1.0e20)
        self.use_mask = False
        if np.ma.is_masked(Bi_host):
            Bi_host = Bi_host.copy().filled(self.mask_value).astype(np.float32)
            self.use_mask = True
        BiShapeY, BiShapeX = Bi_host.shape
        assert BiShapeX == nx + 1 + 2 * halo_x and BiShapeY == ny + 1 + 2 * halo_y, \
            "Wrong size of bottom bathymetry, should be defined on cell intersections, not cell centers. " + \
            str((BiShapeX, BiShapeY)) + " vs " + str((nx + 1 + 2 * halo_x, ny + 1 + 2 * halo_y))
        device = gpu_ctx.torch_device if gpu_ctx is not None else None
        self.Bi = DeviceArray2D(gpu_stream, nx + 1, ny + 1, halo_x, halo_y,
                                np.ma.getdata(Bi_host), device=device)
        Bm_host = np.zeros((self.halo_ny, self.halo_nx), dtype=np.float32, order='C')
        self.Bm = DeviceArray2D(gpu_stream, nx, ny, halo_x, halo_y, Bm_host, device=device)
        bc = boundary_conditions.codes()
        # the halo routines treat every non-periodic side as closed
        bc = tuple(b if b == PERIODIC else WALL for b in bc)
        dom = make_domain(nx, ny, 1.0, 1.0, 9.81, 0.0, 1.3, 2, bc,
                          pitch=self.Bm.pitch_elems, hi_pitch=self.Bi.pitch_elems)
        G.check(G.load().gocean_bathymetry_init(dom, G.ptr(self.Bi.tensor), G.ptr(self.Bm.tensor),
                                                G.stream_handle(gpu_stream)),
                "gocean_bathymetry_init")

    def download(self, gpu_stream):
        Bm_cpu = self.Bm.download(gpu_stream)
        Bi_cpu = self.Bi.download(gpu_stream)
        if self.use_mask:
            Bi_cpu = np.ma.array(data=Bi_cpu, mask=(Bi_cpu == self.mask_value), fill_value=0.0)
            Bm_cpu = np.ma.array(data=Bm_cpu, mask=(Bm_cpu == self.mask_value), fill_value=0.0)
        return Bi_cpu, Bm_cpu

    def release(self):
        self.Bm.release()
        self.Bi.release()
