"""Batched proofs of many independent multilinear / inner-product instances
(effsc_hip_batch_*).  Each instance's proof equals runner::sumcheck on that
instance alone (src/runner.rs:29-69); the batch only shares launches."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .proof import SumcheckProof

NO_GRAPH = 0x2


class DeviceTable:
    """A canonical base-field table in device memory (owned)."""

    def __init__(self, words, device=0):
        w = np.ascontiguousarray(words, dtype=np.uint64)
        self.n = len(w)
        self.device = device
        self.ptr = N.u64p()
        N.check(N.lib().effsc_hip_device_alloc(device, max(self.n, 1), C.byref(self.ptr)))
        if self.n:
            N.check(N.lib().effsc_hip_copy(device, self.ptr, w.ctypes.data_as(N.u64p), self.n))

    @classmethod
    def uniform(cls, n, seed, device=0):
        t = cls.__new__(cls)
        t.n, t.device = n, device
        t.ptr = N.u64p()
        N.check(N.lib().effsc_hip_device_alloc(device, max(n, 1), C.byref(t.ptr)))
        N.check(N.lib().effsc_hip_fill_uniform(device, t.ptr, n, seed, 0))
        return t

    def free(self):
        if getattr(self, "ptr", None):
            N.lib().effsc_hip_device_free(self.device, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BatchProver:
    """kind 1: multilinear over tables a; kind 2: inner product of a and b.
    Tables are DeviceTable objects (kept alive by this object)."""

    def __init__(self, a, b=None, kind=2, device=0, graph=True):
        self.kind = kind
        self.count = len(a)
        self._keep = (a, b)
        lens = (C.c_size_t * self.count)(*[t.n for t in a])
        pa = (N.u64p * self.count)(*[t.ptr for t in a])
        pb = (N.u64p * self.count)(*[t.ptr for t in (b if b is not None else a)])
        self.h = C.c_void_p()
        N.check(N.lib().effsc_hip_batch_new(device, kind, self.count, lens, pa, pb, 0 if graph else NO_GRAPH,
                                            C.byref(self.h)))
        self.v, self.first = [], []
        for i in range(self.count):
            v, f = C.c_size_t(), C.c_size_t()
            N.check(N.lib().effsc_hip_batch_rounds(self.h, i, C.byref(v), C.byref(f)))
            self.v.append(v.value)
            self.first.append(f.value)
        self.total = sum(self.v)

    def autotune(self, candidates=(256, 512, 1024, 2048, 4096, 8192, 16384, 65536), reps=2):
        """Picks the units-per-workgroup tile by timing whole proofs;
        returns (chosen, {candidate: ms per proof})."""
        cand = (C.c_size_t * len(candidates))(*candidates)
        ms = (C.c_double * len(candidates))()
        chosen = C.c_size_t()
        N.check(N.lib().effsc_hip_batch_autotune(self.h, cand, len(candidates), reps, C.byref(chosen), ms))
        return chosen.value, dict(zip(candidates, list(ms)))

    def buffers(self):
        """Output words of one batched proof: (round polys, challenges, finals)."""
        return (np.zeros(max(self.total, 1) * self.kind, dtype=np.uint64),
                np.zeros(max(self.total, 1), dtype=np.uint64), np.zeros(self.count, dtype=np.uint64))

    def prove_raw(self, handles, polys, chals, fins):
        """One batched proof into preallocated word arrays; `handles` is a
        ctypes array of native transcript handles (states advanced)."""
        N.check(N.lib().effsc_hip_batch_prove(self.h, handles, polys.ctypes.data_as(N.u64p),
                                              chals.ctypes.data_as(N.u64p), fins.ctypes.data_as(N.u64p)))

    def prove(self, transcripts):
        """transcripts: one native transcript per instance (state advanced)."""
        K = self.kind
        ts = (C.c_void_p * self.count)(*[t.h for t in transcripts])
        polys, chals, fins = self.buffers()
        self.prove_raw(ts, polys, chals, fins)
        pw, cw, fw = polys.tolist(), chals.tolist(), fins.tolist()
        out = []
        for i in range(self.count):
            f, v = self.first[i], self.v[i]
            rp = [pw[(f + r) * K:(f + r + 1) * K] for r in range(v)]
            out.append(SumcheckProof(rp, cw[f:f + v], fw[i]))
        return out

    def close(self):
        if getattr(self, "h", None):
            N.lib().effsc_hip_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
