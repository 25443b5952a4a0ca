"""Served proofs (effsc_hip_server_*): a request loop with round-level
continuous batching on one GPU.  Each request's proof equals
runner::sumcheck on that request alone (src/runner.rs:29-69)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .proof import SumcheckProof, proof_from_words

MULTILINEAR = 1
INNER_PRODUCT = 2
PRODUCT3_EXT3 = 3

_SHAPE = {MULTILINEAR: (1, 1), INNER_PRODUCT: (2, 1), PRODUCT3_EXT3: (3, 3)}  # (values per round, ext)


@dataclass
class Served:
    """A completed request: its proof and server-clock times (seconds)."""

    proof: SumcheckProof
    submitted: float
    admitted: float
    completed: float

    @property
    def latency(self):
        return self.completed - self.submitted


def _request(kind, v, seeds, tables, transcript, priority):
    if kind not in _SHAPE:
        raise ValueError("kind must be MULTILINEAR, INNER_PRODUCT or PRODUCT3_EXT3")
    q = N.Request()
    q.kind = kind
    q.v = v
    for i in range(3):
        q.seeds[i] = int(seeds[i]) if seeds is not None and i < len(seeds) else 0
        t = tables[i] if tables is not None and i < len(tables) else None
        q.tables[i] = C.cast(C.c_void_p(t), N.u64p) if t else N.u64p()
    q.transcript = transcript.h if transcript is not None else None
    q.priority = float(priority)
    return q


class Server:
    """hbm_budget: bytes of device memory for admitted requests;
    max_active: requests on the device at once."""

    def __init__(self, hbm_budget=1 << 30, max_active=256, device=0):
        self.h = C.c_void_p()
        N.check(N.lib().effsc_hip_server_new(device, hbm_budget, max_active, C.byref(self.h)))
        self._meta = {}

    @staticmethod
    def footprint(kind, v, tables=None):
        q = _request(kind, v, None, tables, None, 0.0)
        out = C.c_size_t()
        N.check(N.lib().effsc_hip_server_footprint(C.byref(q), C.byref(out)))
        return out.value

    def submit(self, kind, v, transcript, seeds=None, tables=None, priority=0.0):
        """Queues one proof; tables (device addresses) or seeds (tables made
        on the device) give its inputs.  Returns the request id."""
        q = _request(kind, v, seeds, tables, transcript, priority)
        rid = C.c_uint64()
        N.check(N.lib().effsc_hip_server_submit(self.h, C.byref(q), C.byref(rid)))
        self._meta[rid.value] = (kind, v)
        return rid.value

    def wait(self, rid):
        kind, v = self._meta.pop(rid)
        K, ext = _SHAPE[kind]
        polys = np.zeros(max(v * K * ext, 1), dtype=np.uint64)
        chals = np.zeros(max(v * ext, 1), dtype=np.uint64)
        fin = np.zeros(ext, dtype=np.uint64)
        times = (C.c_double * 3)()
        N.check(N.lib().effsc_hip_server_wait(self.h, rid, polys.ctypes.data_as(N.u64p), chals.ctypes.data_as(N.u64p),
                                              fin.ctypes.data_as(N.u64p), times))
        return Served(proof_from_words(polys, chals, fin, v, K, ext), times[0], times[1], times[2])

    def now(self):
        t = C.c_double()
        N.check(N.lib().effsc_hip_server_now(self.h, C.byref(t)))
        return t.value

    def stats(self):
        vals = [C.c_uint64() for _ in range(4)]
        N.check(N.lib().effsc_hip_server_stats(self.h, *[C.byref(x) for x in vals]))
        return dict(zip(("ticks", "peak_active", "peak_bytes", "completed"), (x.value for x in vals)))

    def close(self):
        if self.h:
            N.lib().effsc_hip_server_free(self.h)
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
