"""effsc_hip_sumcheck_noalloc: a whole multilinear / inner-product proof with
every buffer beyond the inputs and outputs in the library's static device
scratch.  Checked through the C ABI against the allocating path
(effsc_hip_*_new over the same borrowed device tables + effsc_hip_sumcheck):
round polynomials, challenges, final value and the transcript state after
the proof (the next challenge) bit for bit, on calls in a row with shrinking
inputs, at the size limit, and after a rejected oversize call.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from effsc import _native as N

pytestmark = pytest.mark.gpu
MAX_N = 1 << 20


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DevTable:
    def __init__(self, n, seed):
        self.n = n
        self.p = N.u64p()
        N.check(N.lib().effsc_hip_device_alloc(0, n, C.byref(self.p)))
        N.check(N.lib().effsc_hip_fill_uniform(0, self.p, n, seed, 0))

    def free(self):
        N.lib().effsc_hip_device_free(0, self.p)


def _transcript(ds):
    return N.lib().effsc_hip_transcript_new(ds, len(ds))


def _next_challenge(t):
    out = np.zeros(1, dtype=np.uint64)
    N.check(N.lib().effsc_hip_transcript_challenge(t, 1, out.ctypes.data_as(N.u64p)))
    return int(out[0])


def _reference_path(kind, a, b, n, ds):
    cfg = N.Config()
    N.lib().effsc_hip_config_default(C.byref(cfg))
    cfg.flags = 1  # EFFSC_HIP_BORROW_INPUT
    h = C.c_void_p()
    if kind == 1:
        N.check(N.lib().effsc_hip_multilinear_new(C.byref(cfg), a.p, n, C.byref(h)))
    else:
        N.check(N.lib().effsc_hip_inner_product_new(C.byref(cfg), a.p, b.p, n, C.byref(h)))
    v = n.bit_length() - 1
    polys = np.zeros(v * kind, dtype=np.uint64)
    chals = np.zeros(v, dtype=np.uint64)
    fin = np.zeros(1, dtype=np.uint64)
    t = _transcript(ds)
    try:
        N.check(N.lib().effsc_hip_sumcheck(h, v, t, C.cast(None, N.HOOK), None, polys.ctypes.data_as(N.u64p),
                                           chals.ctypes.data_as(N.u64p), fin.ctypes.data_as(N.u64p)))
        nxt = _next_challenge(t)
    finally:
        N.lib().effsc_hip_transcript_free(t)
        N.lib().effsc_hip_free(h)
    return polys, chals, fin, nxt


def _noalloc_path(kind, a, b, n, ds):
    v = n.bit_length() - 1
    polys = np.zeros(v * kind, dtype=np.uint64)
    chals = np.zeros(v, dtype=np.uint64)
    fin = np.zeros(1, dtype=np.uint64)
    t = _transcript(ds)
    try:
        N.check(N.lib().effsc_hip_sumcheck_noalloc(0, kind, a.p, b.p if b else None, n, t,
                                                   polys.ctypes.data_as(N.u64p), chals.ctypes.data_as(N.u64p),
                                                   fin.ctypes.data_as(N.u64p)))
        nxt = _next_challenge(t)
    finally:
        N.lib().effsc_hip_transcript_free(t)
    return polys, chals, fin, nxt


def _check(kind, n, seed):
    say(f"noalloc kind={kind} n={n} seed={seed:#x}")
    a = DevTable(n, seed)
    b = DevTable(n, seed + 1) if kind == 2 else None
    try:
        ds = f"noalloc/{kind}/{n}".encode()
        ref = _reference_path(kind, a, b, n, ds)
        got = _noalloc_path(kind, a, b, n, ds)
        for x, y, what in zip(got, ref, ("round polys", "challenges", "final value", "next challenge")):
            assert np.array_equal(np.asarray(x), np.asarray(y)), what
    finally:
        a.free()
        if b:
            b.free()


@pytest.mark.parametrize("kind", [2, 1])
def test_noalloc_matches_allocating_path_on_calls_in_a_row(kind):
    # a large call, then smaller ones: the scratch keeps the previous call's
    # contents, which must not leak into the next proof
    for n in (1 << 18, 1 << 11, 1 << 5, 2):
        _check(kind, n, 0x5EED7000 + n)


def test_noalloc_at_the_size_limit_and_rejects_larger():
    _check(2, MAX_N, 0x5EED7100)
    a = DevTable(2 * MAX_N, 0x5EED7101)
    b = DevTable(2 * MAX_N, 0x5EED7102)
    try:
        out = np.zeros(64, dtype=np.uint64)
        t = _transcript(b"x")
        rc = N.lib().effsc_hip_sumcheck_noalloc(0, 2, a.p, b.p, 2 * MAX_N, t, out.ctypes.data_as(N.u64p),
                                                out.ctypes.data_as(N.u64p), out.ctypes.data_as(N.u64p))
        N.lib().effsc_hip_transcript_free(t)
        assert rc == 1  # EFFSC_HIP_EINVAL
        assert b"scratch" in N.lib().effsc_hip_last_error()
    finally:
        a.free()
        b.free()
    _check(2, 1 << 12, 0x5EED7103)


def test_noalloc_rejects_bad_arguments():
    a = DevTable(16, 1)
    try:
        out = np.zeros(64, dtype=np.uint64)
        t = _transcript(b"x")
        for kind, n in ((3, 16), (1, 12), (1, 1), (2, 16)):
            rc = N.lib().effsc_hip_sumcheck_noalloc(0, kind, a.p, None, n, t, out.ctypes.data_as(N.u64p),
                                                    out.ctypes.data_as(N.u64p), out.ctypes.data_as(N.u64p))
            assert rc == 1, (kind, n)
        N.lib().effsc_hip_transcript_free(t)
    finally:
        a.free()
