"""The served scheduler's learned service-time model.

A small network (csrc/learn.hip, trained on the GPU by the library's own
forward / backward / Adam kernels) maps a request's shape to the log of its
measured service time; Server admission then takes the shortest predicted
request first instead of the first to arrive.  Proofs are unaffected: the
model only orders admissions.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _native as N
from .serve import INNER_PRODUCT, MULTILINEAR, PRODUCT3_EXT3

INPUTS = 8
_WORK = {MULTILINEAR: 1, INNER_PRODUCT: 2, PRODUCT3_EXT3: 9}  # tables x words per element after round 0


def features(kind, v):
    """[ML?, IP?, P3?, v/26, (v/26)^2, log2(work)/40, work/2^30, 1]."""
    work = _WORK[kind] * (1 << v)
    return [float(kind == MULTILINEAR), float(kind == INNER_PRODUCT), float(kind == PRODUCT3_EXT3), v / 26.0,
            (v / 26.0) ** 2, math.log2(work) / 40.0, work / 2.0 ** 30, 1.0]


def work_priority(kind, v):
    """The non-learned ordering on the same features: log of the request's
    work, 2^v x tables x words per element (shortest work first)."""
    return math.log(_WORK[kind] * float(1 << v))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Network:
    """Thin handle over effsc_hip_mlp_*."""

    def __init__(self, inputs=INPUTS, hidden=32, seed=1, device=0):
        self.inputs, self.hidden = inputs, hidden
        self.h = C.c_void_p()
        N.check(N.lib().effsc_hip_mlp_new(device, inputs, hidden, seed, C.byref(self.h)))
        n = C.c_size_t()
        N.check(N.lib().effsc_hip_mlp_param_count(self.h, C.byref(n)))
        self.count = n.value

    def params(self):
        out = np.zeros(self.count, dtype=np.float32)
        N.check(N.lib().effsc_hip_mlp_get_params(self.h, _fp(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        N.check(N.lib().effsc_hip_mlp_set_params(self.h, _fp(p)))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(len(X), dtype=np.float32)
        N.check(N.lib().effsc_hip_mlp_forward(self.h, _fp(X), len(X), _fp(y)))
        return y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.count, dtype=np.float32)
        loss = C.c_float()
        N.check(N.lib().effsc_hip_mlp_loss_grad(self.h, _fp(X), _fp(T), len(T), C.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, X, T, steps=2000, lr=1e-2):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = C.c_float()
        N.check(N.lib().effsc_hip_mlp_train(self.h, _fp(X), _fp(T), len(T), steps, lr, C.byref(loss)))
        return loss.value

    def close(self):
        if self.h:
            N.lib().effsc_hip_mlp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServiceTimeModel:
    """Predicts log(service time in ms) of a request from (kind, v)."""

    def __init__(self, hidden=32, seed=1, device=0):
        self.net = Network(INPUTS, hidden, seed, device)

    def fit(self, samples, steps=3000, lr=1e-2):
        """samples: (kind, v, service_ms) triples measured on the device."""
        X = np.array([features(k, v) for k, v, _ in samples], dtype=np.float32)
        T = np.array([math.log(ms) for _, _, ms in samples], dtype=np.float32)
        return self.net.train(X, T, steps, lr)

    def predict_ms(self, reqs):
        X = np.array([features(k, v) for k, v in reqs], dtype=np.float32)
        return np.exp(self.net.forward(X).astype(np.float64))

    def priority(self, kind, v):
        """Admission priority: the predicted log service time."""
        return float(self.net.forward(np.array([features(kind, v)], dtype=np.float32))[0])


def measure_service_times(server, reqs, transcript_for, tables_for=None):
    """Runs each (kind, v) alone on `server` and returns (kind, v, ms):
    admitted -> completed on the server clock."""
    out = []
    for i, (kind, v) in enumerate(reqs):
        tables = tables_for(i) if tables_for else None
        rid = server.submit(kind, v, transcript_for(i, kind), seeds=[i + 1, i + 2, i + 3], tables=tables)
        r = server.wait(rid)
        out.append((kind, v, (r.completed - r.admitted) * 1e3))
    return out
