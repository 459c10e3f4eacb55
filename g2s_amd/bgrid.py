"""G2S .bgrid files (serialized DataImage) through the qsmi C-ABI.

The format G2S's qs reads its training images and destination grids from
and writes its results to (include/DataImage.hpp:143-254,
src/DataImage.cpp:35-163): a header (total size, dimensions, variable
types, encoding) followed by float32 data in DataImage order (variable
fastest, then dimension 0). Files may be gzip-compressed.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib

CONTINUOUS, CATEGORICAL = 0, 1
FLOAT, INTEGER, UINTEGER = 0, 1, 2


def serialize(data, dims, var_types, encoding=FLOAT):
    """Bytes of a grid: data has prod(dims) * len(var_types) values in
    DataImage order; dims lists dimension 0 first."""
    dims = np.ascontiguousarray(dims, np.uint32)
    types = np.ascontiguousarray(var_types, np.int32)
    data = np.ascontiguousarray(data, np.float32).reshape(-1)
    n = C.c_size_t(0)
    check(lib().qsmi_bgrid_bytes(int(dims.size), dims, int(types.size), C.byref(n)))
    if data.size != int(np.prod(dims.astype(np.int64))) * types.size:
        raise ValueError("data size does not match dims x variables")
    out = np.empty(n.value, np.uint8)
    check(lib().qsmi_bgrid_serialize(int(dims.size), dims, int(types.size), types, int(encoding), data,
                                     out.ctypes.data, out.size))
    return out.tobytes()


def valid(raw):
    """True when raw is a well-formed serialized grid."""
    buf = np.frombuffer(raw, np.uint8)
    return bool(lib().qsmi_bgrid_valid(buf.ctypes.data if buf.size else None, buf.size))


def parse(raw):
    """(data [cells, nVar] float32, dims, var_types, encoding) of a buffer."""
    buf = np.frombuffer(raw, np.uint8)
    nd, nv, enc, off = C.c_int(0), C.c_int(0), C.c_int(0), C.c_size_t(0)
    dims = np.zeros(16, np.uint32)
    types = np.zeros(4096, np.int32)
    check(lib().qsmi_bgrid_header(buf.ctypes.data, buf.size, C.byref(nd), dims, C.byref(nv), types, types.size,
                                  C.byref(enc), C.byref(off)))
    d = tuple(int(x) for x in dims[: nd.value])
    data = np.frombuffer(raw, np.float32, offset=off.value).reshape(-1, nv.value).copy()
    return data, d, [int(t) for t in types[: nv.value]], enc.value


def write(path, data, dims, var_types, encoding=FLOAT, gzip=True):
    raw = np.frombuffer(serialize(data, dims, var_types, encoding), np.uint8)
    check(lib().qsmi_bgrid_write_file(str(path).encode(), raw.ctypes.data, raw.size, int(bool(gzip))))


def read(path):
    """(data [cells, nVar], dims, var_types, encoding) of a plain or gzip file."""
    n = C.c_size_t(0)
    check(lib().qsmi_bgrid_read_file(str(path).encode(), None, 0, C.byref(n)))
    buf = np.empty(n.value, np.uint8)
    check(lib().qsmi_bgrid_read_file(str(path).encode(), buf.ctypes.data, buf.size, C.byref(n)))
    return parse(buf.tobytes())
