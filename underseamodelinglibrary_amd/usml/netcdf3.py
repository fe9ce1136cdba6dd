"""A minimal writer of netCDF classic files in the 64-bit offset variant
(CDF-2): fixed-size variables plus record variables along one unlimited
dimension, appended one record at a time.

Layout (the netCDF classic format specification, all integers big-endian):
header = 'CDF' 0x02, numrecs, dimension list, global attribute list,
variable list; then each fixed-size variable's values at the offset its
header entry gives, padded to 4 bytes; then the records, each holding one
slice of every record variable in definition order.  Lists are a 4-byte tag
(0x0A dimensions, 0x0C attributes, 0x0B variables) and a 4-byte count, or 8
zero bytes when empty; names are a count and the bytes padded to 4; the
unlimited dimension has length 0; a variable entry holds its dimension ids,
attributes, type, size (of one record, for a record variable) and 8-byte
offset (of its slice in the first record).  Record slices are padded to 4
bytes unless there is a single record variable.  numrecs, at byte 4, is
rewritten after every appended record, so the file is complete after each.
"""
import struct

import numpy as np

NC_BYTE, NC_CHAR, NC_SHORT, NC_INT, NC_FLOAT, NC_DOUBLE = 1, 2, 3, 4, 5, 6
_DTYPE = {NC_SHORT: ">i2", NC_INT: ">i4", NC_FLOAT: ">f4", NC_DOUBLE: ">f8", NC_BYTE: ">i1"}
_TAG_DIM, _TAG_VAR, _TAG_ATT = 0x0A, 0x0B, 0x0C
UNLIMITED = None


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


class _Var:
    def __init__(self, name, nc_type, dims, atts, shape, values, is_record):
        self.name, self.nc_type, self.dims, self.atts = name, nc_type, dims, atts
        self.shape = shape          # whole shape, or one record's shape
        self.values = values        # big-endian array (fixed variables only)
        self.is_record = is_record
        self.nbytes = int(np.prod(shape, dtype=np.int64)) * np.dtype(_DTYPE[nc_type]).itemsize
        self.vsize = 0
        self.begin = 0


class Writer:
    """Collects dimensions, global attributes and variables, then writes
    them in definition order: write() for a file of fixed-size variables,
    or create() / put_record() / close() when there are record variables."""

    def __init__(self):
        self.dims = []      # (name, length); length 0 is the unlimited dimension
        self.atts = {}
        self.vars = []
        self._fh = None
        self._numrecs = 0
        self._rec_dim = None

    def add_dim(self, name, length):
        """length UNLIMITED (None) makes the record dimension (at most one)."""
        if length is UNLIMITED:
            if self._rec_dim is not None:
                raise ValueError("only one unlimited dimension")
            self._rec_dim = len(self.dims)
            self.dims.append((name, 0))
        else:
            self.dims.append((name, int(length)))
        return len(self.dims) - 1

    def add_var(self, name, nc_type, dims, values=None, atts=None):
        """A record variable has the unlimited dimension first and takes its
        values through put_record(); a fixed variable takes them here."""
        if self._fh is not None:
            raise RuntimeError("variables must be defined before create()")
        dims = list(dims)
        if self._rec_dim is not None and self._rec_dim in dims[1:]:
            raise ValueError("the unlimited dimension must be a variable's first")
        is_record = bool(dims) and dims[0] == self._rec_dim
        shape = tuple(self.dims[d][1] for d in (dims[1:] if is_record else dims))
        arr = None
        if not is_record:
            arr = np.asarray(values, dtype=_DTYPE[nc_type]).reshape(shape)
        self.vars.append(_Var(name, nc_type, dims, dict(atts or {}), shape, arr, is_record))

    def _header(self):
        out = b"CDF\x02" + struct.pack(">i", self._numrecs)
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
        for v in self.vars:
            out += _name(v.name) + struct.pack(">i", len(v.dims))
            out += b"".join(struct.pack(">i", d) for d in v.dims)
            out += _att_list(v.atts)
            out += struct.pack(">iiq", v.nc_type, v.vsize, v.begin)
        return out

    def _layout(self):
        rec = [v for v in self.vars if v.is_record]
        for v in self.vars:
            pad = not v.is_record or len(rec) > 1
            v.vsize = v.nbytes + ((4 - v.nbytes % 4) % 4 if pad else 0)
        # the header's length does not depend on the offsets' values
        pos = len(self._header())
        for v in self.vars:
            if not v.is_record:
                v.begin = pos
                pos += v.vsize
        for v in rec:
            v.begin = pos
            pos += v.vsize
        self._recsize = sum(v.vsize for v in rec)
        self._rec_start = rec[0].begin if rec else pos
        self._rec_vars = rec

    def create(self, filename):
        self._numrecs = 0
        self._layout()
        self._fh = open(filename, "wb")
        self._fh.write(self._header())
        for v in self.vars:
            if not v.is_record:
                self._fh.write(_pad4(v.values.tobytes()))
        self._fh.flush()

    def put_record(self, values):
        """Append one record: values maps each record variable's name to its
        slice (one record's shape)."""
        if self._fh is None:
            raise RuntimeError("file not open")
        chunks = []
        for v in self._rec_vars:
            arr = np.asarray(values[v.name], dtype=_DTYPE[v.nc_type]).reshape(v.shape)
            b = arr.tobytes()
            chunks.append(b + b"\x00" * (v.vsize - len(b)))
        self._fh.seek(self._rec_start + self._numrecs * self._recsize)
        self._fh.write(b"".join(chunks))
        self._numrecs += 1
        self._fh.seek(4)
        self._fh.write(struct.pack(">i", self._numrecs))
        self._fh.flush()

    @property
    def numrecs(self):
        return self._numrecs

    def close(self):
        if self._fh is not None:
            self._fh.close()
            self._fh = None

    def write(self, filename):
        self.create(filename)
        self.close()
