"""Streaming writer of NetCDF-3 files with 64-bit offsets (the classic data
model, "CDF-2"), following the published NetCDF classic format
specification: header, fixed-size variables, then records appended one at
a time with the record count updated in place.  Written because the netCDF4
library is not available and scipy.io.netcdf_file keeps every record in
memory until the file is closed."""

from __future__ import annotations

import struct

import numpy as np

_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 0x0A, 0x0B, 0x0C
_TYPES = {np.dtype('int8'): 1, np.dtype('S1'): 2, np.dtype('int16'): 3, np.dtype('int32'): 4,
          np.dtype('float32'): 5, np.dtype('float64'): 6}


def _pad(n):
    return (4 - n % 4) % 4


def _name(s):
    b = s.encode()
    return struct.pack('>i', len(b)) + b + b'\0' * _pad(len(b))


def _attr_values(v):
    if isinstance(v, str):
        b = v.encode()
        return 2, len(b), b + b'\0' * _pad(len(b))
    a = np.atleast_1d(np.asarray(v))
    if a.dtype.kind == 'b':
        a = a.astype(np.int32)
    if a.dtype.kind in 'iu':
        a = a.astype(np.int32)
    elif a.dtype.kind == 'f' and a.dtype != np.float32:
        a = a.astype(np.float64)
    b = a.astype(a.dtype.newbyteorder('>')).tobytes()
    return _TYPES[np.dtype(a.dtype.name)], a.size, b + b'\0' * _pad(len(b))


def _attrs(d):
    if not d:
        return struct.pack('>ii', 0, 0)
    out = struct.pack('>ii', _NC_ATTRIBUTE, len(d))
    for k, v in d.items():
        t, n, b = _attr_values(v)
        out += _name(k) + struct.pack('>ii', t, n) + b
    return out


class NC3Writer:
    """dims: ordered (name, length) pairs, length None for the record
    dimension (at most one, first in each record variable); variables:
    ordered (name, dtype, dim names, attributes) tuples; fixed: data of the
    non-record variables, written with the header."""

    def __init__(self, path, dims, gattrs, variables, fixed):
        self.path = path
        self.dims = list(dims)
        dim_index = {n: i for i, (n, _) in enumerate(self.dims)}
        dim_len = dict(self.dims)
        self.vars = []
        for name, dtype, vdims, vattrs in variables:
            dtype = np.dtype(dtype)
            is_rec = len(vdims) > 0 and dim_len[vdims[0]] is None
            shape = [dim_len[d] for d in vdims[1:]] if is_rec else [dim_len[d] for d in vdims]
            nbytes = int(np.prod(shape, dtype=np.int64)) * dtype.itemsize
            self.vars.append(dict(name=name, dtype=dtype, dims=vdims, attrs=vattrs, rec=is_rec,
                                  shape=shape, nbytes=nbytes, vsize=nbytes + _pad(nbytes)))
        self.rec_vars = [v for v in self.vars if v['rec']]
        # a single record variable is stored without padding between records
        if len(self.rec_vars) == 1:
            self.rec_vars[0]['vsize'] = self.rec_vars[0]['nbytes']
        self.numrecs = 0

        def header(begins):
            h = b'CDF\x02' + struct.pack('>i', 0)
            if self.dims:
                h += struct.pack('>ii', _NC_DIMENSION, len(self.dims))
                for n, L in self.dims:
                    h += _name(n) + struct.pack('>i', 0 if L is None else L)
            else:
                h += struct.pack('>ii', 0, 0)
            h += _attrs(gattrs)
            if self.vars:
                h += struct.pack('>ii', _NC_VARIABLE, len(self.vars))
                for v, b in zip(self.vars, begins):
                    h += _name(v['name']) + struct.pack('>i', len(v['dims']))
                    h += b''.join(struct.pack('>i', dim_index[d]) for d in v['dims'])
                    h += _attrs(v['attrs'])
                    h += struct.pack('>iiq', _TYPES[v['dtype']], min(v['vsize'], 2 ** 31 - 1), b)
            else:
                h += struct.pack('>ii', 0, 0)
            return h

        hlen = len(header([0] * len(self.vars)))
        begins, off = [], hlen
        for v in self.vars:
            if not v['rec']:
                begins.append(off)
                off += v['vsize']
            else:
                begins.append(0)
        rec_off = off
        for v, i in zip(self.vars, range(len(self.vars))):
            if v['rec']:
                begins[i] = rec_off
                rec_off += v['vsize']
        self.record_bytes = sum(v['vsize'] for v in self.rec_vars)
        self.data_start = off
        with open(path, 'wb') as f:
            f.write(header(begins))
            for v in self.vars:
                if not v['rec']:
                    a = np.asarray(fixed.get(v['name'], np.zeros(v['shape'], v['dtype'])), dtype=v['dtype'])
                    assert list(a.shape) == list(v['shape']) or a.size == int(np.prod(v['shape'])), v['name']
                    b = a.astype(v['dtype'].newbyteorder('>')).tobytes()
                    f.write(b + b'\0' * _pad(len(b)))

    def append(self, values):
        """Append one record: values[name] for every record variable."""
        with open(self.path, 'r+b') as f:
            f.seek(self.data_start + self.numrecs * self.record_bytes)
            for v in self.rec_vars:
                a = np.asarray(values[v['name']], dtype=v['dtype'])
                assert a.size == int(np.prod(v['shape'], dtype=np.int64)), v['name']
                b = a.astype(v['dtype'].newbyteorder('>')).tobytes()
                f.write(b + b'\0' * (v['vsize'] - len(b)))
            self.numrecs += 1
            f.seek(4)
            f.write(struct.pack('>i', self.numrecs))
