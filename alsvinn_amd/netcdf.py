"""NetCDF ensemble writer with the reference's file contents
(alsfvm/src/io/NetCDFWriter.cpp:35-184).

Each file holds dimensions member, time, x, y (and z in 3D), coordinate
variables member (int), time, x, y[, z] (cell midpoints) and one variable per
conserved quantity with dimensions (member, time, x, y[, z]): rho, E, mx, my
(mz) for the Euler equations, as the reference names them. A snapshot of one
member at one save is written as the reference writes it: the interior cells
of each variable in memory order (x fastest) filling the (x, y[, z]) slab.
The time variable holds the save fractions (k+1)/numberOfSaves (0 first when
the initial state is written), as the reference stores them.

The file format is netCDF classic with 64-bit offsets (CDF-2), written by this
module directly: member is the record dimension, so a member's slabs are
contiguous and any number of members and saves fits (CDF-2 limits only the
per-record size of a variable to 4 GiB). Slabs are written in place as the
samples reach their saves; nothing is held in memory. netCDF-C, xarray and
scipy.io.netcdf_file read these files.
"""
import os
import struct

import numpy as np

NC_INT = 4
NC_FLOAT = 5
NC_DOUBLE = 6
_ABSENT = b"\x00" * 8
_NC_DIMENSION = 10
_NC_VARIABLE = 11
_NC_ATTRIBUTE = 12


def variable_names(equation_nvar, dim):
    """Output variables (name, index into the conserved volume). Euler 2D/3D
    follow the reference: rho = 0, E = last, mx = 1, my = 2 (mz = 3)."""
    if equation_nvar == 1:
        return [("u", 0)]
    if equation_nvar == 3:
        return [("rho", 0), ("E", 2), ("mx", 1)]
    names = [("rho", 0), ("E", equation_nvar - 1), ("mx", 1), ("my", 2)]
    if equation_nvar == 5:
        names.append(("mz", 3))
    return names


def _name(s):
    b = s.encode()
    return struct.pack(">i", len(b)) + b + b"\x00" * (-len(b) % 4)


def _attrs(attrs):
    if not attrs:
        return _ABSENT
    out = struct.pack(">ii", _NC_ATTRIBUTE, len(attrs))
    for k, v in attrs.items():
        b = str(v).encode()
        out += _name(k) + struct.pack(">ii", 2, len(b)) + b + b"\x00" * (-len(b) % 4)
    return out


class EnsembleWriter:
    """One file of num_members members and len(times) saves on a grid of
    dims (nx, ny, nz) with cell midpoints xs, ys, zs (zs ignored in 2D)."""

    def __init__(self, path, dims, midpoints, num_members, times, variables, dtype=np.float64,
                 attributes=None):
        nx, ny, nz = dims
        self.path = path
        self.dim = 3 if nz > 1 else 2
        self.shape = (nx, ny, nz) if self.dim == 3 else (nx, ny)
        self.cells = nx * ny * nz
        self.ntimes = len(times)
        self.nmembers = num_members
        self.variables = list(variables)
        self.dtype = np.dtype(dtype).newbyteorder(">")
        nctype = NC_DOUBLE if np.dtype(dtype) == np.float64 else NC_FLOAT
        esize = self.dtype.itemsize

        dims_list = [("member", 0), ("time", self.ntimes), ("x", nx), ("y", ny)]
        if self.dim == 3:
            dims_list.append(("z", nz))
        dimid = {n: i for i, (n, _) in enumerate(dims_list)}
        coord = [("time", [dimid["time"]], nctype, np.asarray(times, dtype=float)),
                 ("x", [dimid["x"]], nctype, np.asarray(midpoints[0], dtype=float)),
                 ("y", [dimid["y"]], nctype, np.asarray(midpoints[1], dtype=float))]
        if self.dim == 3:
            coord.append(("z", [dimid["z"]], nctype, np.asarray(midpoints[2], dtype=float)))
        slab_dims = [dimid["member"], dimid["time"], dimid["x"], dimid["y"]] + (
            [dimid["z"]] if self.dim == 3 else [])
        rec_size = self.ntimes * self.cells * esize  # per member, per variable
        rec_vsize = rec_size + (-rec_size % 4)
        self._rec_vsize = rec_vsize

        def header(begins):
            h = b"CDF\x02" + struct.pack(">i", num_members)
            h += struct.pack(">ii", _NC_DIMENSION, len(dims_list))
            for n, size in dims_list:
                h += _name(n) + struct.pack(">i", size)
            h += _attrs(attributes or {})
            allv = [(n, d, t, None) for n, d, t, _ in coord] + \
                   [(n, slab_dims, nctype, None) for n, _ in self.variables] + \
                   [("member", [dimid["member"]], NC_INT, None)]
            h += struct.pack(">ii", _NC_VARIABLE, len(allv))
            for (n, d, t, _), b in zip(allv, begins):
                if n == "member":
                    vsize = 4
                elif d == slab_dims:
                    vsize = rec_vsize
                else:
                    sz = len(coord[[c[0] for c in coord].index(n)][3]) * esize
                    vsize = sz + (-sz % 4)
                h += _name(n) + struct.pack(">i", len(d)) + b"".join(struct.pack(">i", x) for x in d)
                h += _ABSENT + struct.pack(">iiq", t, vsize, b)
            return h

        nvars = len(coord) + len(self.variables) + 1
        hsize = len(header([0] * nvars))
        begins = []
        pos = hsize
        for _, _, _, data in coord:
            begins.append(pos)
            sz = len(data) * esize
            pos += sz + (-sz % 4)
        self._recstart = pos
        for _ in self.variables:
            begins.append(pos)
            pos += rec_vsize
        begins.append(pos)  # member (int) record variable
        self._begins = begins
        self._recsize = rec_vsize * len(self.variables) + 4
        with open(path, "wb") as f:
            f.write(header(begins))
            for (_, _, _, data), b in zip(coord, begins):
                f.seek(b)
                f.write(np.asarray(data, dtype=self.dtype).tobytes())
            # member ids in every record
            for m in range(num_members):
                f.seek(begins[-1] + m * self._recsize)
                f.write(struct.pack(">i", m))
            end = self._recstart + num_members * self._recsize
            f.truncate(end)

    def write(self, member, time_index, volume, ghosts):
        """Writes member's snapshot time_index from a conserved volume
        (nvar, tz, ty, tx) with ghost layers ghosts = (gx, gy, gz)."""
        if not (0 <= member < self.nmembers and 0 <= time_index < self.ntimes):
            raise IndexError("member %d / save %d outside the file" % (member, time_index))
        gx, gy, gz = ghosts
        _, tz, ty, tx = volume.shape
        with open(self.path, "r+b") as f:
            for k, (_, var) in enumerate(self.variables):
                inner = volume[var, gz:tz - gz, gy:ty - gy, gx:tx - gx]
                data = np.ascontiguousarray(inner).astype(self.dtype).tobytes()  # x fastest
                off = self._begins[len(self._begins) - 1 - len(self.variables) + k] \
                    + member * self._recsize + time_index * self.cells * self.dtype.itemsize
                f.seek(off)
                f.write(data)


def save_fractions(number_of_saves, write_initial=False):
    """The reference's time axis: (k+1)/numberOfSaves, with 0 first when the
    initial state is written (SimulatorSetup.cpp createWriter)."""
    t = [0.0] if write_initial else []
    return t + [float(k + 1) / number_of_saves for k in range(number_of_saves)]


def midpoints(lower, upper, dims):
    """Cell midpoints per axis (Grid::getCellMidpoints, Grid.cpp:148-171)."""
    from .initial_data import midpoints as mids
    return [mids(n, a, b) for a, b, n in zip(lower, upper, dims)]


def remove(path):
    if os.path.exists(path):
        os.remove(path)
