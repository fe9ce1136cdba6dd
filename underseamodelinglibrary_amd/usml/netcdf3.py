"""A minimal writer of netCDF classic files in the 64-bit offset variant
(CDF-2), for fixed-size variables only.

Layout (the netCDF classic format specification, all integers big-endian):
header = 'CDF' 0x02, numrecs, dimension list, global attribute list,
variable list; then each variable's values at the offset its header entry
gives, padded to 4 bytes.  Lists are a 4-byte tag (0x0A dimensions,
0x0C attributes, 0x0B variables) and a 4-byte count, or 8 zero bytes when
empty; names are a count and the bytes padded to 4; a variable entry holds
its dimension ids, attributes, type, padded size and 8-byte offset.
"""
import struct

import numpy as np

NC_BYTE, NC_CHAR, NC_SHORT, NC_INT, NC_FLOAT, NC_DOUBLE = 1, 2, 3, 4, 5, 6
_DTYPE = {NC_SHORT: ">i2", NC_INT: ">i4", NC_FLOAT: ">f4", NC_DOUBLE: ">f8", NC_BYTE: ">i1"}
_TAG_DIM, _TAG_VAR, _TAG_ATT = 0x0A, 0x0B, 0x0C


def _pad4(b):
    return b + b"\x00" * ((4 - len(b) % 4) % 4)


def _name(s):
    b = s.encode("utf-8")
    return struct.pack(">i", len(b)) + _pad4(b)


def _att_list(atts):
    if not atts:
        return b"\x00" * 8
    out = struct.pack(">ii", _TAG_ATT, len(atts))
    for key, value in atts.items():
        out += _name(key)
        if isinstance(value, str):
            b = value.encode("utf-8")
            out += struct.pack(">ii", NC_CHAR, len(b)) + _pad4(b)
        else:
            arr = np.atleast_1d(np.asarray(value, dtype=">f8"))
            out += struct.pack(">ii", NC_DOUBLE, arr.size) + arr.tobytes()
    return out


class Writer:
    """Collects dimensions, global attributes and variables, then writes
    them in definition order."""

    def __init__(self):
        self.dims = []      # (name, length)
        self.atts = {}
        self.vars = []      # (name, nc_type, dim ids, attributes, values)

    def add_dim(self, name, length):
        self.dims.append((name, int(length)))
        return len(self.dims) - 1

    def add_var(self, name, nc_type, dims, values, atts=None):
        shape = tuple(self.dims[d][1] for d in dims)
        arr = np.asarray(values, dtype=_DTYPE[nc_type]).reshape(shape)
        self.vars.append((name, nc_type, list(dims), dict(atts or {}), arr))

    def _header(self, begins):
        out = b"CDF\x02" + struct.pack(">i", 0)
        if self.dims:
            out += struct.pack(">ii", _TAG_DIM, len(self.dims))
            for name, length in self.dims:
                out += _name(name) + struct.pack(">i", length)
        else:
            out += b"\x00" * 8
        out += _att_list(self.atts)
        if not self.vars:
            return out + b"\x00" * 8
        out += struct.pack(">ii", _TAG_VAR, len(self.vars))
        for (name, nc_type, dims, atts, arr), begin in zip(self.vars, begins):
            out += _name(name) + struct.pack(">i", len(dims))
            out += b"".join(struct.pack(">i", d) for d in dims)
            out += _att_list(atts)
            vsize = len(_pad4(arr.tobytes()))
            out += struct.pack(">iiq", nc_type, vsize, begin)
        return out

    def write(self, filename):
        sizes = [len(_pad4(v[4].tobytes())) for v in self.vars]
        # the header's length does not depend on the offsets' values
        start = len(self._header([0] * len(self.vars)))
        begins, pos = [], start
        for s in sizes:
            begins.append(pos)
            pos += s
        with open(filename, "wb") as fh:
            fh.write(self._header(begins))
            for v in self.vars:
                fh.write(_pad4(v[4].tobytes()))
