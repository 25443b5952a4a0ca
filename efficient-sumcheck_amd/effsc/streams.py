"""Streamed input (src/streams/): FileStream files of serialized field
elements, their claims, and provers whose tables are uploaded from streams
with the transfer overlapping the first pass over the data."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native as N
from .field import CANONICAL, W_FROM_RAW7
from .provers import InnerProductProver, MultilinearProver, NativeProver, _config


class FileStream:
    """FileStream (src/streams/file/core.rs:13-76): 2^v serialized elements,
    each component a little-endian canonical u64, mapped read-only."""

    def __init__(self, path, ext=1):
        self.path, self.ext = str(path), ext
        self.h = C.c_void_p()
        N.check(N.lib().effsc_hip_stream_open(os.fsencode(self.path), ext, C.byref(self.h)))

    @staticmethod
    def write_to_file(path, data, ext=1):
        """FileStream::write_to_file (file/core.rs:51-61) for canonical words
        (ext words per element, element-major)."""
        np.ascontiguousarray(data, dtype="<u8").tofile(str(path))

    @staticmethod
    def delete_file(path):
        os.remove(str(path))

    def num_variables(self):
        v = C.c_size_t()
        N.check(N.lib().effsc_hip_stream_num_variables(self.h, C.byref(v)))
        return v.value

    def read(self, first, count):
        out = np.zeros(max(count, 1) * self.ext, dtype=np.uint64)
        N.check(N.lib().effsc_hip_stream_read(self.h, first, count, out.ctypes.data_as(N.u64p)))
        return out[:count * self.ext]

    def evaluation(self, point):
        w = self.read(point, 1)
        return int(w[0]) if self.ext == 1 else tuple(int(x) for x in w)

    def close(self):
        if self.h:
            N.lib().effsc_hip_stream_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _handles(streams):
    return (C.c_void_p * len(streams))(*[s.h.value for s in streams])


def multivariate_product_claim(streams, w=W_FROM_RAW7, device=0):
    """sum_i prod_j s_j[i] (src/streams/stream.rs:16-41), on the GPU."""
    ext = streams[0].ext
    out = np.zeros(ext, dtype=np.uint64)
    N.check(N.lib().effsc_hip_stream_claim(device, _handles(streams), len(streams), w, out.ctypes.data_as(N.u64p)))
    return int(out[0]) if ext == 1 else tuple(int(x) for x in out)


def multivariate_claim(stream, w=W_FROM_RAW7, device=0):
    """sum_i s[i] (src/streams/stream.rs:3-14), on the GPU."""
    return multivariate_product_claim([stream], w, device)


RESIDENT, OUT_OF_CORE, AUTO = 0, 1, 2


def _stream_prover(cls, kind, streams, ext, repr, w, device, mode=RESIDENT):
    p = cls.__new__(cls)
    NativeProver.__init__(p, ext, repr)
    cfg = _config(ext, 1, repr, w, device, 0)
    h = C.c_void_p()
    N.check(N.lib().effsc_hip_stream_prover_new_ex(C.byref(cfg), kind, _handles(streams), mode, C.byref(h)))
    p._finish(h)
    # an out-of-core prover reads the streams again during its proof
    p._streams = list(streams)
    return p


def multilinear_from_stream(stream, *, ext=1, repr=CANONICAL, w=W_FROM_RAW7, device=0, mode=RESIDENT):
    """MultilinearProver over a stream's table.  mode OUT_OF_CORE keeps the
    table in the stream (whole proofs only); AUTO picks by device memory."""
    return _stream_prover(MultilinearProver, 1, [stream], ext, repr, w, device, mode)


def inner_product_from_stream(a, b, *, ext=1, repr=CANONICAL, w=W_FROM_RAW7, device=0, mode=RESIDENT):
    """InnerProductProver over two streams' tables (mode as for
    multilinear_from_stream)."""
    return _stream_prover(InnerProductProver, 2, [a, b], ext, repr, w, device, mode)
