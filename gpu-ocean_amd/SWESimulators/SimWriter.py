"""NetCDF output of a simulator run (gpu_ocean/SWESimulators/SimWriter.py:35-471).

Same file layout as the reference's writer for the non-staggered (CDKLM16)
grid: the global attributes the hot start reads back, the dimensions
time (unlimited), x, y, x_Hi, y_Hi, the axis variables, the bogus
polar-stereographic projection, land mask, Hm, Hi and one eta/hu/hv record
per writeTimestep.  netCDF4 (HDF5) is not available in this environment, so
the file is NetCDF-3 with 64-bit offsets, streamed record by record
(netcdf3.NC3Writer), which netCDF4, xarray, scipy and the NetCDF tools
read; variables are therefore not zlib-compressed.
"""

from __future__ import annotations

import datetime
import os
import subprocess

import numpy as np

from . import netcdf3


class SimNetCDFWriter:

    def __init__(self, sim, super_dir_name=None, filename=None, num_layers=1, staggered_grid=False,
                 ignore_ghostcells=False, offset_x=0, offset_y=0):
        if staggered_grid:
            raise NotImplementedError("staggered grids belong to the FBL/CTCS schemes, not part of this build")
        self.gpu_stream = sim.gpu_stream
        self.ignore_ghostcells = ignore_ghostcells
        self.staggered_grid = staggered_grid
        self.num_layers = num_layers
        now = datetime.datetime.now()
        self.timestamp = now.strftime("%Y_%m_%d-%H_%M_%S")
        self.timestamp_short = now.strftime("%Y_%m_%d")
        try:
            self.git_hash = subprocess.check_output(['git', 'rev-parse', 'HEAD'],
                                                    stderr=subprocess.DEVNULL).decode().strip()
        except Exception:
            self.git_hash = "git info missing..."
        self.simulator_short = str(sim.__class__.__name__)
        self.super_dir_name = super_dir_name if super_dir_name is not None else os.getcwd()
        self.dir_name = os.path.join(self.super_dir_name, "netcdf_" + self.timestamp_short + "/")
        if not filename:
            self.output_file_name = self.dir_name + self.simulator_short + "_" + self.timestamp + ".nc"
        elif '/' not in filename:
            self.output_file_name = self.dir_name + self.simulator_short + "_" + filename + ".nc"
        else:
            self.output_file_name = filename + ".nc"
            self.dir_name = filename.rsplit('/', 1)[0]
        self.output_file_name = self.output_file_name.replace('.nc.nc', '.nc')
        os.makedirs(self.dir_name, exist_ok=True)

        Hi, Hm = sim.bathymetry.download(self.gpu_stream)
        nx, ny, dx, dy = int(sim.nx), int(sim.ny), float(sim.dx), float(sim.dy)
        gx, gy = int(sim.ghost_cells_x), int(sim.ghost_cells_y)
        self.dx, self.dy = dx, dy
        attrs = {
            "Conventions": "CF-1.4",
            "git_hash": self.git_hash,
            "ignore_ghostcells": str(ignore_ghostcells),
            "num_layers": num_layers,
            "staggered_grid": str(staggered_grid),
            "simulator_short": self.simulator_short,
            "boundary_conditions": str(sim.boundary_conditions),
            "boundary_conditions_sponge_mr": str(sim.boundary_conditions.getSponge()),
            "time_integrator": int(sim.rk_order),
            "minmod_theta": float(sim.theta),
            "coriolis_force": float(np.asarray(sim.f).flat[0]) if np.size(sim.f) == 1 else 0.0,
            "coriolis_beta": float(sim.coriolis_beta),
            "y_zero_reference_cell": float(sim.y_zero_reference_cell) - 2,
            "eddy_viscosity_coefficient": sim.A if isinstance(sim.A, str) else float(sim.A),
            "g": float(sim.g), "nx": nx, "ny": ny, "dx": dx, "dy": dy, "dt": float(sim.dt),
            "auto_dt": "False", "bottom_friction_r": float(sim.r),
            "ghost_cells_north": gy, "ghost_cells_east": gx, "ghost_cells_south": gy, "ghost_cells_west": gx,
        }
        if sim.wind_stress.source_filename is not None:
            attrs["wind_stress_source"] = str(sim.wind_stress.source_filename)
        if self.simulator_short == "CDKLM16":
            attrs["model_time_step"] = float(sim.model_time_step)
            attrs["small_scale_perturbation"] = str(bool(sim.small_scale_perturbation))
            if sim.small_scale_perturbation:
                attrs["small_scale_perturbation_amplitude"] = float(sim.small_scale_model_error.soar_q0)
                attrs["small_scale_perturbation_interpolation_factor"] = int(
                    sim.small_scale_model_error.interpolation_factor)

        tot_x, tot_y = 2 * gx, 2 * gy
        sx, sy = (nx, ny) if ignore_ghostcells else (nx + tot_x, ny + tot_y)
        dims = [('time', None), ('x', sx), ('y', sy), ('x_Hi', nx + tot_x + 1), ('y_Hi', ny + tot_y + 1)]
        if not ignore_ghostcells:
            xs = np.linspace(-gx * dx + dx / 2.0, (nx + gx) * dx - dx / 2.0, nx + tot_x)
            ys = np.linspace(-gy * dy + dy / 2.0, (ny + gy) * dy - dy / 2.0, ny + tot_y)
        else:
            xs = np.linspace(offset_x, nx * dx, nx)
            ys = np.linspace(offset_y, ny * dy, ny)
        axis = lambda a: {"standard_name": "projection_%s_coordinate" % a.lower(), "axis": a, "units": "meter"}
        depth = {"standard_name": "water_surface_reference_datum_altitude", "grid_mapping": "projection_stere",
                 "coordinates": "y x", "units": "meter"}
        proj = {"grid_mapping_name": "polar_stereographic",
                "scale_factor_at_projection_origin": 0.9330127018922193,
                "straight_vertical_longitude_from_pole": 70.0, "latitude_of_projection_origin": 90.0,
                "earth_radius": 6371000.0,
                "proj4": "+proj=stere +lat_0=90 +lon_0=70 +lat_ts=60 +units=m +a=6.371e+06 +e=0 +no_defs"}
        field = lambda std, coords, units: {"standard_name": std, "grid_mapping": "projection_stere",
                                            "coordinates": coords, "units": units}
        variables = [
            ('time', 'float32', ('time',), {"units": "seconds since 1970-01-01 00:00:00"}),
            ('x', 'float32', ('x',), axis("X")), ('y', 'float32', ('y',), axis("Y")),
            ('x_Hi', 'float32', ('x_Hi',), axis("X")), ('y_Hi', 'float32', ('y_Hi',), axis("Y")),
            ('projection_stere', 'int32', (), proj),
            ('land_binary_mask', 'float32', ('y', 'x'), {"standard_name": "land_binary_mask", "units": "1"}),
            ('Hm', 'float32', ('y', 'x'), depth), ('Hi', 'float32', ('y_Hi', 'x_Hi'), depth),
            ('eta', 'float32', ('time', 'y', 'x'),
             field('water_surface_height_above_reference_datum', 'y x', 'meter')),
            ('hu', 'float32', ('time', 'y', 'x'), field('x_sea_water_velocity', 'y_hu x_hu', 'meter second-1')),
            ('hv', 'float32', ('time', 'y', 'x'), field('y_sea_water_velocity', 'y_hv x_hv', 'meter second-1')),
        ]
        fixed = {'x': xs, 'y': ys,
                 'x_Hi': np.linspace(-gx * dx, (nx + gx) * dx, nx + tot_x + 1),
                 'y_Hi': np.linspace(-gy * dy, (ny + gy) * dy, ny + tot_y + 1),
                 'projection_stere': np.zeros((), np.int32),
                 'land_binary_mask': np.zeros((sy, sx), np.float32),
                 'Hm': np.ma.getdata(Hm) if not ignore_ghostcells else np.ma.getdata(Hm)[gy:-gy, gx:-gx],
                 'Hi': np.ma.getdata(Hi)}
        self.ncfile = netcdf3.NC3Writer(self.output_file_name, dims, attrs, variables, fixed)
        self.i = 0
        self.writeTimestep(sim)

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_value, traceback):
        self.close()

    def close(self):
        self.ncfile = None

    def writeTimestep(self, sim):
        """Append the current state as one record (SimWriter.py:443-471)."""
        eta, hu, hv = sim.download()
        eta, hu, hv = (np.ma.getdata(a) for a in (eta, hu, hv))
        if self.ignore_ghostcells:
            # interior cells only; the reference's slices here assume one
            # ghost layer (the staggered schemes) and do not fit the x, y
            # dimensions of a two-ghost-cell grid
            eta, hu, hv = eta[2:-2, 2:-2], hu[2:-2, 2:-2], hv[2:-2, 2:-2]
        self.ncfile.append({'time': np.float32(sim.t), 'eta': eta, 'hu': hu, 'hv': hv})
        self.i += 1
