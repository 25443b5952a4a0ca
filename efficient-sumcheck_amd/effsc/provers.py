"""GPU-backed provers behind the reference's SumcheckProver trait
(src/sumcheck_prover.rs:36-71): degree(), round(challenge|None), finalize(r),
final_value(), plus each prover's post-state accessors.

Every method forwards to the C ABI (include/effsc_hip.h); the tables live in
device memory for the prover's lifetime.

Inputs: numpy uint64 arrays of words (AoS, `input_ext` words per element) or
sequences of elements (ints, or ext-tuples for ext = 2, 3).  With borrow=True the
inputs are device addresses (ints) of canonical planes, see
EFFSC_HIP_BORROW_INPUT.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .field import CANONICAL, MONT64, P, W_FROM_RAW7, Field

BORROW_INPUT = 0x1


def _words(table, input_ext):
    if isinstance(table, np.ndarray):
        return np.ascontiguousarray(table, dtype=np.uint64)
    out = []
    for x in table:
        if isinstance(x, (tuple, list)):
            out.extend(int(c) % P for c in x)
        else:
            out.append(int(x) % P)
    return np.array(out, dtype=np.uint64)


def _config(ext, input_ext, repr, w, device, flags):
    cfg = N.Config()
    N.lib().effsc_hip_config_default(C.byref(cfg))
    cfg.ext = ext
    cfg.input_ext = input_ext
    cfg.repr = repr
    cfg.ext_w = w
    cfg.device = device
    cfg.flags = flags
    return cfg


class NativeProver:
    """Common forwarding to the C ABI."""

    def __init__(self, ext, repr):
        self.h = None
        self.ext = ext
        self.repr = repr
        self.field = Field(ext)

    def _finish(self, handle):
        self.h = handle
        self._deg = self._size(N.lib().effsc_hip_degree)

    def _size(self, fn, *args):
        out = C.c_size_t(0)
        N.check(fn(self.h, *args, C.byref(out)))
        return out.value

    def _elem(self, words):
        return int(words[0]) if self.ext == 1 else tuple(int(x) for x in words[:self.ext])

    def _chal_words(self, c):
        if self.ext == 1:
            return np.array([int(c)], dtype=np.uint64)
        return np.array([int(x) for x in c], dtype=np.uint64)

    # ---- trait
    def degree(self):
        return self._deg

    def round(self, challenge=None):
        nw = max(self._deg, 1)
        out = np.zeros(nw * self.ext, dtype=np.uint64)
        if challenge is None:
            N.check(N.lib().effsc_hip_round(self.h, None, out.ctypes.data_as(N.u64p)))
        else:
            cw = self._chal_words(challenge)
            N.check(N.lib().effsc_hip_round(self.h, cw.ctypes.data_as(N.u64p), out.ctypes.data_as(N.u64p)))
        return [self._elem(out[i * self.ext:(i + 1) * self.ext]) for i in range(nw)]

    def finalize(self, last_challenge):
        cw = self._chal_words(last_challenge)
        N.check(N.lib().effsc_hip_finalize(self.h, cw.ctypes.data_as(N.u64p)))

    def final_value(self):
        out = np.zeros(self.ext, dtype=np.uint64)
        N.check(N.lib().effsc_hip_final_value(self.h, out.ctypes.data_as(N.u64p)))
        return self._elem(out)

    # ---- post-state
    def table_words(self, i):
        n = self._size(N.lib().effsc_hip_table_len, i)
        out = np.zeros(max(n * self.ext, 1), dtype=np.uint64)
        got = C.c_size_t(0)
        N.check(N.lib().effsc_hip_read_table(self.h, i, out.ctypes.data_as(N.u64p), n * self.ext, C.byref(got)))
        return out[:n * self.ext]

    def table(self, i):
        w = self.table_words(i)
        return [self._elem(w[j * self.ext:(j + 1) * self.ext]) for j in range(len(w) // self.ext)]

    def reset(self):
        N.check(N.lib().effsc_hip_reset(self.h))

    def close(self):
        if getattr(self, "h", None):
            N.lib().effsc_hip_free(self.h)
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


class MultilinearProver(NativeProver):
    """MultilinearProver (src/provers/multilinear.rs:25-80), degree 1."""

    def __init__(self, evals, *, ext=1, input_ext=None, repr=CANONICAL, w=W_FROM_RAW7, device=0,
                 borrow=False, n=None):
        super().__init__(ext, repr)
        input_ext = ext if input_ext is None else input_ext
        cfg = _config(ext, input_ext, repr, w, device, BORROW_INPUT if borrow else 0)
        h = C.c_void_p()
        if borrow:
            N.check(N.lib().effsc_hip_multilinear_new(C.byref(cfg), C.cast(C.c_void_p(evals), N.u64p), n,
                                                       C.byref(h)))
        else:
            e = _words(evals, input_ext)
            N.check(N.lib().effsc_hip_multilinear_new(C.byref(cfg), e.ctypes.data_as(N.u64p), len(e) // input_ext,
                                                       C.byref(h)))
        self._finish(h)

    def evals(self):
        return self.table(0)

    def num_variables(self):
        n = self._size(N.lib().effsc_hip_table_len, 0)
        return 0 if n == 0 else (n - 1).bit_length()


class InnerProductProver(NativeProver):
    """InnerProductProver (src/provers/inner_product.rs:26-127), degree 2."""

    def __init__(self, a, b, *, ext=1, input_ext=None, repr=CANONICAL, w=W_FROM_RAW7, device=0, borrow=False,
                 n=None):
        super().__init__(ext, repr)
        input_ext = ext if input_ext is None else input_ext
        cfg = _config(ext, input_ext, repr, w, device, BORROW_INPUT if borrow else 0)
        h = C.c_void_p()
        if borrow:
            N.check(N.lib().effsc_hip_inner_product_new(C.byref(cfg), C.cast(C.c_void_p(a), N.u64p),
                                                         C.cast(C.c_void_p(b), N.u64p), n, C.byref(h)))
        else:
            A, B = _words(a, input_ext), _words(b, input_ext)
            if len(A) != len(B):
                raise ValueError("a and b must have equal length")  # src/provers/inner_product.rs:34
            N.check(N.lib().effsc_hip_inner_product_new(C.byref(cfg), A.ctypes.data_as(N.u64p),
                                                         B.ctypes.data_as(N.u64p), len(A) // input_ext, C.byref(h)))
        self._finish(h)

    def evaluations(self):
        return self.table(0), self.table(1)

    def final_evaluations(self):
        a, b = self.evaluations()
        if len(a) == 1:
            return a[0], b[0]
        return self.field.zero(), self.field.zero()


class CoefficientProver(NativeProver):
    """CoefficientProver with the built-in product-of-k evaluator
    (src/provers/coefficient.rs:18-330; the k = 3 evaluator is the
    reference's TripleProductEval, tests/adversarial_verifier.rs:444-465):
    proves sum_x prod_i f_i(x), degree k (1 <= k <= 4)."""

    def __init__(self, pairwise, *, ext=1, input_ext=None, repr=CANONICAL, w=W_FROM_RAW7, device=0,
                 borrow=False, n=None):
        super().__init__(ext, repr)
        input_ext = ext if input_ext is None else input_ext
        k = len(pairwise)
        cfg = _config(ext, input_ext, repr, w, device, BORROW_INPUT if borrow else 0)
        h = C.c_void_p()
        if borrow:
            ptrs = (N.u64p * k)(*[C.cast(C.c_void_p(p), N.u64p) for p in pairwise])
            N.check(N.lib().effsc_hip_product_new(C.byref(cfg), ptrs, k, n, C.byref(h)))
        else:
            arrs = [_words(t, input_ext) for t in pairwise]
            if len({len(a) for a in arrs}) != 1:
                raise ValueError("pairwise tables must have equal length")
            ptrs = (N.u64p * k)(*[a.ctypes.data_as(N.u64p) for a in arrs])
            N.check(N.lib().effsc_hip_product_new(C.byref(cfg), ptrs, k, len(arrs[0]) // input_ext, C.byref(h)))
        self._finish(h)

    def pairwise(self):
        return [self.table(i) for i in range(self._size(N.lib().effsc_hip_table_count))]


ProductProver = CoefficientProver


class EqFactoredProver(NativeProver):
    """EqFactoredProver (src/provers/eq_factored.rs:67-289): sum_x eq(w,x) p(x)."""

    def __init__(self, w_point, p_evals, *, ext=1, repr=CANONICAL, w=W_FROM_RAW7, device=0, borrow=False):
        super().__init__(ext, repr)
        cfg = _config(ext, ext, repr, w, device, BORROW_INPUT if borrow else 0)
        W = _words(w_point, ext)
        h = C.c_void_p()
        if borrow:
            # p_evals is a device address of a canonical 2^v-element table
            v = len(W) // ext
            N.check(N.lib().effsc_hip_eq_factored_new(C.byref(cfg), W.ctypes.data_as(N.u64p), v,
                                                       C.cast(C.c_void_p(p_evals), N.u64p), 1 << v, C.byref(h)))
        else:
            Pv = _words(p_evals, ext)
            N.check(N.lib().effsc_hip_eq_factored_new(C.byref(cfg), W.ctypes.data_as(N.u64p), len(W) // ext,
                                                       Pv.ctypes.data_as(N.u64p), len(Pv) // ext, C.byref(h)))
        self._finish(h)

    def final_factors(self):
        p = self.table(0)
        if len(p) == 1:
            l, r = self.table(1)[0], self.table(2)[0]
            return p[0], self.field.mul(l, r)
        return self.field.zero(), self.field.zero()


def _lsb(cls_fn, tables, ext, input_ext, repr, w, device, borrow, n):
    """Common constructor of the LSB provers: host word arrays or, with
    borrow, device addresses of canonical tables of n elements."""
    input_ext = ext if input_ext is None else input_ext
    cfg = _config(ext, input_ext, repr, w, device, BORROW_INPUT if borrow else 0)
    if borrow:
        ptrs = [C.cast(C.c_void_p(t), N.u64p) for t in tables]
        count = n
    else:
        arrs = [_words(t, input_ext) for t in tables]
        if len({len(a) for a in arrs}) != 1:
            raise ValueError("a and b must have equal length")  # src/provers/inner_product_lsb.rs
        ptrs = [a.ctypes.data_as(N.u64p) for a in arrs]
        count = len(arrs[0]) // input_ext
    h = C.c_void_p()
    N.check(cls_fn(C.byref(cfg), ptrs, count, h))
    return h


class MultilinearProverLSB(NativeProver):
    """MultilinearProverLSB (src/provers/multilinear_lsb.rs): pairs (2k, 2k+1),
    degree 1; any length (an unpaired last element counts in the round value
    and is dropped by the fold, as in the reference)."""

    def __init__(self, evals, *, ext=1, input_ext=None, repr=CANONICAL, w=W_FROM_RAW7, device=0, borrow=False,
                 n=None):
        super().__init__(ext, repr)
        self._finish(_lsb(lambda cfg, p, cnt, h: N.lib().effsc_hip_multilinear_lsb_new(cfg, p[0], cnt, C.byref(h)),
                          [evals], ext, input_ext, repr, w, device, borrow, n))

    def evals(self):
        return self.table(0)

    def num_variables(self):
        n = self._size(N.lib().effsc_hip_table_len, 0)
        return 0 if n == 0 else (n - 1).bit_length()


class InnerProductProverLSB(NativeProver):
    """InnerProductProverLSB (src/provers/inner_product_lsb.rs), degree 2."""

    def __init__(self, a, b, *, ext=1, input_ext=None, repr=CANONICAL, w=W_FROM_RAW7, device=0, borrow=False,
                 n=None):
        super().__init__(ext, repr)
        self._finish(_lsb(lambda cfg, p, cnt, h: N.lib().effsc_hip_inner_product_lsb_new(cfg, p[0], p[1], cnt,
                                                                                          C.byref(h)),
                          [a, b], ext, input_ext, repr, w, device, borrow, n))

    def evaluations(self):
        return self.table(0), self.table(1)

    def final_evaluations(self):
        a, b = self.evaluations()
        if len(a) == 1:
            return a[0], b[0]
        return self.field.zero(), self.field.zero()


class CoefficientProverLSB(NativeProver):
    """CoefficientProverLSB with the product-of-k evaluator
    (src/provers/coefficient_lsb.rs), degree k."""

    def __init__(self, pairwise, *, ext=1, input_ext=None, repr=CANONICAL, w=W_FROM_RAW7, device=0, borrow=False,
                 n=None):
        super().__init__(ext, repr)
        k = len(pairwise)

        def make(cfg, p, cnt, h):
            arr = (N.u64p * k)(*p)
            return N.lib().effsc_hip_product_lsb_new(cfg, arr, k, cnt, C.byref(h))
        self._finish(_lsb(make, pairwise, ext, input_ext, repr, w, device, borrow, n))

    def pairwise(self):
        return [self.table(i) for i in range(self._size(N.lib().effsc_hip_table_count))]


def bit_reverse(values, device=0):
    """reorder_vec_msb (src/streams/memory/core.rs:20-24) on the GPU:
    out[i] = values[bit_reverse(i)] for a power-of-two list of base-field words."""
    w = np.ascontiguousarray(values, dtype=np.uint64)
    n = len(w)
    L = N.lib()
    src, dst = N.u64p(), N.u64p()
    N.check(L.effsc_hip_device_alloc(device, max(n, 1), C.byref(src)))
    N.check(L.effsc_hip_device_alloc(device, max(n, 1), C.byref(dst)))
    try:
        out = np.zeros(n, dtype=np.uint64)
        N.check(L.effsc_hip_copy(device, src, w.ctypes.data_as(N.u64p), n))
        N.check(L.effsc_hip_bit_reverse(device, dst, src, n, 1))
        N.check(L.effsc_hip_copy(device, out.ctypes.data_as(N.u64p), dst, n))
        return out
    finally:
        L.effsc_hip_device_free(device, src)
        L.effsc_hip_device_free(device, dst)


class GkrProver(NativeProver):
    """GkrProver (src/provers/gkr.rs:38-202): the GKR layer polynomial
    add(b,c) (W(b) + W(c)) + mult(b,c) W(b) W(c) over 2k variables, degree 2."""

    def __init__(self, add_evals, mult_evals, w_evals, *, ext=1, input_ext=None, repr=CANONICAL, w=W_FROM_RAW7,
                 device=0):
        super().__init__(ext, repr)
        input_ext = ext if input_ext is None else input_ext
        cfg = _config(ext, input_ext, repr, w, device, 0)
        A, M, Wt = _words(add_evals, input_ext), _words(mult_evals, input_ext), _words(w_evals, input_ext)
        n = len(Wt) // input_ext
        if len(A) != n * n * input_ext:
            raise ValueError("add_evals must have len w^2")  # src/provers/gkr.rs:61
        if len(M) != n * n * input_ext:
            raise ValueError("mult_evals must have len w^2")  # src/provers/gkr.rs:62
        h = C.c_void_p()
        N.check(N.lib().effsc_hip_gkr_new(C.byref(cfg), A.ctypes.data_as(N.u64p), M.ctypes.data_as(N.u64p),
                                           Wt.ctypes.data_as(N.u64p), n, C.byref(h)))
        self._finish(h)

    def claimed_w_values(self):
        out = np.zeros(2 * self.ext, dtype=np.uint64)
        N.check(N.lib().effsc_hip_gkr_claimed_w(self.h, out.ctypes.data_as(N.u64p)))
        return self._elem(out[:self.ext]), self._elem(out[self.ext:])


def fold(values, weight, device=0, *, ext=1, w=W_FROM_RAW7):
    """fold(&mut Vec<F>, weight) (src/fold.rs:14): returns the folded list
    (length next_pow2(n)/2), computed on the device.  ext = 1 takes ints;
    ext = 2 or 3 takes ext-tuples (or AoS words) and an ext-tuple weight."""
    v = _words(values, ext)
    n = len(v) // ext
    if n <= 1:
        return [int(x) for x in v] if ext == 1 else [tuple(int(c) for c in v[i * ext:(i + 1) * ext]) for i in range(n)]
    ptr = N.u64p()
    N.check(N.lib().effsc_hip_device_alloc(device, n * ext, C.byref(ptr)))
    try:
        N.check(N.lib().effsc_hip_copy(device, ptr, v.ctypes.data_as(N.u64p), n * ext))
        if ext == 1:
            m = fold_device(C.cast(ptr, C.c_void_p).value, n, weight, device)
        else:
            m = fold_device_ext(C.cast(ptr, C.c_void_p).value, n, weight, ext, w, device)
        out = np.zeros(m * ext, dtype=np.uint64)
        N.check(N.lib().effsc_hip_copy(device, out.ctypes.data_as(N.u64p), ptr, m * ext))
    finally:
        N.lib().effsc_hip_device_free(device, ptr)
    if ext == 1:
        return [int(x) for x in out]
    return [tuple(int(c) for c in out[i * ext:(i + 1) * ext]) for i in range(m)]


def fold_device_ext(ptr, n, weight, ext, w=W_FROM_RAW7, device=0):
    """In-place fold of n extension elements (AoS canonical words, `ext` per
    element) at device address `ptr` with an ext-tuple weight."""
    out = C.c_size_t(0)
    wt = np.array([int(c) % P for c in weight], dtype=np.uint64)
    N.check(N.lib().effsc_hip_fold_ext(device, ext, w, C.cast(C.c_void_p(ptr), N.u64p), n, wt.ctypes.data_as(N.u64p),
                                       C.byref(out)))
    return out.value


def fold_device(ptr, n, weight, device=0):
    """In-place fold of n canonical words at device address `ptr`."""
    out = C.c_size_t(0)
    N.check(N.lib().effsc_hip_fold(device, C.cast(C.c_void_p(ptr), N.u64p), n, int(weight) % P, C.byref(out)))
    return out.value
