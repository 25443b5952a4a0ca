"""ctypes binding of the C oracle (oracle/effsc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry and the
benchmark's cpu_baseline leg, never by the product library.  Field elements
cross this interface as canonical u64 words (``ext`` words per element).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libeffsc_oracle.so")
P = 0xFFFFFFFF00000001
# F64::from_raw(7): the reference's extension non-residue (src/tests/fields.rs:49,80)
W_FROM_RAW7 = 0xFFFFFFF800000001

ML, IP, PRODUCT, EQ = 1, 2, 3, 4

_lib = None
u64p = C.POINTER(C.c_uint64)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u64 = C.c_uint64
        sz = C.c_size_t
        vp = C.c_void_p
        sig = {
            "or_gl_add": (u64, [u64, u64]),
            "or_gl_sub": (u64, [u64, u64]),
            "or_gl_mul": (u64, [u64, u64]),
            "or_gl_mul_i128": (u64, [u64, u64]),
            "or_gl_inv": (u64, [u64]),
            "or_mont_mul": (u64, [u64, u64]),
            "or_to_mont": (u64, [u64]),
            "or_from_mont": (u64, [u64]),
            "or_ext_mul": (None, [C.c_int, u64, u64p, u64p, u64p]),
            "or_ext_inv": (None, [C.c_int, u64, u64p, u64p]),
            "or_uniform": (u64, [u64, u64]),
            "or_fill_uniform": (None, [u64p, sz, u64]),
            "or_keccak_f1600": (None, [u64p]),
            "or_shake256": (None, [C.c_char_p, sz, C.c_char_p, sz]),
            "or_tr_new_shake": (vp, [C.c_char_p, sz]),
            "or_tr_new_rng": (vp, [u64]),
            "or_tr_send": (None, [vp, C.c_int, u64p]),
            "or_tr_challenge": (None, [vp, C.c_int, u64p]),
            "or_tr_free": (None, [vp]),
            "or_prover_new_ml": (vp, [C.c_int, u64, u64p, sz]),
            "or_prover_new_ip": (vp, [C.c_int, u64, u64p, u64p, sz]),
            "or_prover_new_product": (vp, [C.c_int, u64, C.c_int, C.POINTER(u64p), sz]),
            "or_prover_new_product_ext": (vp, [C.c_int, u64, C.c_int, C.POINTER(u64p), sz]),
            "or_prover_new_product_lsb_ext": (vp, [C.c_int, u64, C.c_int, C.POINTER(u64p), sz]),
            "or_prover_new_eq": (vp, [C.c_int, u64, u64p, sz, u64p, sz]),
            "or_prover_new_gkr": (vp, [C.c_int, u64, u64p, u64p, u64p, sz]),
            "or_prover_new_ml_lsb": (vp, [C.c_int, u64, u64p, sz]),
            "or_prover_new_ip_lsb": (vp, [C.c_int, u64, u64p, u64p, sz]),
            "or_prover_new_product_lsb": (vp, [C.c_int, u64, C.c_int, C.POINTER(u64p), sz]),
            "or_prover_degree": (C.c_int, [vp]),
            "or_prover_round": (None, [vp, u64p, u64p]),
            "or_prover_finalize": (None, [vp, u64p]),
            "or_prover_final_value": (None, [vp, u64p]),
            "or_prover_table_len": (sz, [vp, C.c_int]),
            "or_prover_read_table": (None, [vp, C.c_int, u64p]),
            "or_prover_free": (None, [vp]),
            "or_sumcheck": (None, [vp, sz, vp, u64p, u64p, u64p]),
            "or_sumcheck_verify": (C.c_int, [C.c_int, u64, u64p, sz, sz, vp, u64p, sz, u64p, u64p]),
            "or_bench_ip": (C.c_double, [u64p, u64p, sz, sz, C.c_int, u64p, u64p, u64p]),
            "or_bench_ip_ex": (C.c_double, [u64p, u64p, sz, sz, C.c_int, C.c_int, C.c_char_p, sz, u64p, u64p,
                                            u64p]),
            "or_bench_ml_ex": (C.c_double, [u64p, sz, sz, C.c_int, C.c_char_p, sz, u64p, u64p, u64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _arr(words):
    a = np.ascontiguousarray(np.asarray(words, dtype=np.uint64))
    return a, a.ctypes.data_as(u64p)


# ---------------------------------------------------------------- field
def gl_mul(a, b):
    return lib().or_gl_mul(a, b)


def ext_mul(d, w, a, b):
    A, pa = _arr(a)
    B, pb = _arr(b)
    out = np.zeros(d, dtype=np.uint64)
    lib().or_ext_mul(d, w, pa, pb, out.ctypes.data_as(u64p))
    return [int(x) for x in out]


def ext_inv(d, w, a):
    A, pa = _arr(a)
    out = np.zeros(d, dtype=np.uint64)
    lib().or_ext_inv(d, w, pa, out.ctypes.data_as(u64p))
    return [int(x) for x in out]


def uniform(n, seed):
    out = np.zeros(n, dtype=np.uint64)
    lib().or_fill_uniform(out.ctypes.data_as(u64p), n, seed)
    return out


def shake256(msg: bytes, outlen: int) -> bytes:
    buf = C.create_string_buffer(outlen)
    lib().or_shake256(msg, len(msg), buf, outlen)
    return buf.raw


# ---------------------------------------------------------------- transcripts
class Transcript:
    def __init__(self, handle):
        self.h = handle

    @classmethod
    def shake(cls, domsep: bytes):
        return cls(lib().or_tr_new_shake(domsep, len(domsep)))

    @classmethod
    def rng(cls, seed: int):
        return cls(lib().or_tr_new_rng(seed))

    def send(self, value):
        A, p = _arr(value)
        lib().or_tr_send(self.h, len(A), p)

    def challenge(self, d=1):
        out = np.zeros(d, dtype=np.uint64)
        lib().or_tr_challenge(self.h, d, out.ctypes.data_as(u64p))
        return [int(x) for x in out]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_tr_free(self.h)
            self.h = None


# ---------------------------------------------------------------- provers
class Prover:
    """Oracle prover; tables are numpy uint64 arrays of canonical words."""

    def __init__(self, handle, ext, keep=()):
        if not handle:
            raise ValueError("oracle prover construction failed")
        self.h = handle
        self.ext = ext
        self._keep = keep

    @classmethod
    def multilinear(cls, evals, ext=1, w=W_FROM_RAW7):
        E, p = _arr(evals)
        return cls(lib().or_prover_new_ml(ext, w, p, len(E) // ext), ext)

    @classmethod
    def inner_product(cls, a, b, ext=1, w=W_FROM_RAW7):
        A, pa = _arr(a)
        B, pb = _arr(b)
        return cls(lib().or_prover_new_ip(ext, w, pa, pb, len(A) // ext), ext)

    @classmethod
    def product(cls, tables, ext=1, w=W_FROM_RAW7, tables_in_ext=False):
        arrs = [_arr(t) for t in tables]
        ptrs = (u64p * len(arrs))(*[p for _, p in arrs])
        n = len(arrs[0][0]) // (ext if tables_in_ext else 1)
        fn = lib().or_prover_new_product_ext if tables_in_ext else lib().or_prover_new_product
        return cls(fn(ext, w, len(arrs), ptrs, n), ext)

    @classmethod
    def eq_factored(cls, point, p_evals, ext=1, w=W_FROM_RAW7):
        W, pw = _arr(point)
        Pv, pp = _arr(p_evals)
        return cls(lib().or_prover_new_eq(ext, w, pw, len(W) // ext, pp, len(Pv) // ext), ext)

    @classmethod
    def multilinear_lsb(cls, evals, ext=1, w=W_FROM_RAW7):
        """MultilinearProverLSB (src/provers/multilinear_lsb.rs), pairs (2k, 2k+1)."""
        E, p = _arr(evals)
        return cls(lib().or_prover_new_ml_lsb(ext, w, p, len(E) // ext), ext)

    @classmethod
    def inner_product_lsb(cls, a, b, ext=1, w=W_FROM_RAW7):
        """InnerProductProverLSB (src/provers/inner_product_lsb.rs)."""
        A, pa = _arr(a)
        B, pb = _arr(b)
        return cls(lib().or_prover_new_ip_lsb(ext, w, pa, pb, len(A) // ext), ext)

    @classmethod
    def product_lsb(cls, tables, ext=1, w=W_FROM_RAW7, tables_in_ext=False):
        """CoefficientProverLSB with the product evaluator; base tables lifted
        (or tables already in the extension with tables_in_ext)."""
        arrs = [_arr(t) for t in tables]
        ptrs = (u64p * len(arrs))(*[p for _, p in arrs])
        if tables_in_ext:
            return cls(lib().or_prover_new_product_lsb_ext(ext, w, len(arrs), ptrs, len(arrs[0][0]) // ext), ext)
        return cls(lib().or_prover_new_product_lsb(ext, w, len(arrs), ptrs, len(arrs[0][0])), ext)

    @classmethod
    def gkr(cls, add, mult, wit, ext=1, w=W_FROM_RAW7):
        """GkrProver::new(add_evals, mult_evals, w_evals), src/provers/gkr.rs:58-87."""
        A, pa = _arr(add)
        M, pm = _arr(mult)
        Wt, pw = _arr(wit)
        return cls(lib().or_prover_new_gkr(ext, w, pa, pm, pw, len(Wt) // ext), ext)

    def claimed_w_values(self):
        """(W(b*), W(c*)) after the full proof, src/provers/gkr.rs:89-97."""
        if len(self.table(2)) // self.ext == 1:
            return self.table(2), self.table(3)
        return None

    def degree(self):
        return lib().or_prover_degree(self.h)

    def round(self, challenge=None):
        d = max(self.degree(), 1)
        out = np.zeros(d * self.ext, dtype=np.uint64)
        if challenge is None:
            lib().or_prover_round(self.h, None, out.ctypes.data_as(u64p))
        else:
            C_, pc = _arr(challenge)
            lib().or_prover_round(self.h, pc, out.ctypes.data_as(u64p))
        return [[int(x) for x in out[i * self.ext:(i + 1) * self.ext]] for i in range(d)]

    def finalize(self, challenge):
        C_, pc = _arr(challenge)
        lib().or_prover_finalize(self.h, pc)

    def final_value(self):
        out = np.zeros(self.ext, dtype=np.uint64)
        lib().or_prover_final_value(self.h, out.ctypes.data_as(u64p))
        return [int(x) for x in out]

    def table(self, i):
        n = lib().or_prover_table_len(self.h, i)
        out = np.zeros(n * self.ext, dtype=np.uint64)
        lib().or_prover_read_table(self.h, i, out.ctypes.data_as(u64p))
        return out

    def sumcheck(self, num_rounds, transcript):
        d = max(self.degree(), 1)
        polys = np.zeros(max(num_rounds, 1) * d * self.ext, dtype=np.uint64)
        chals = np.zeros(max(num_rounds, 1) * self.ext, dtype=np.uint64)
        fin = np.zeros(self.ext, dtype=np.uint64)
        lib().or_sumcheck(self.h, num_rounds, transcript.h, polys.ctypes.data_as(u64p),
                          chals.ctypes.data_as(u64p), fin.ctypes.data_as(u64p))
        e = self.ext
        rp = [[[int(x) for x in polys[(r * d + i) * e:(r * d + i + 1) * e]] for i in range(d)]
              for r in range(num_rounds)]
        ch = [[int(x) for x in chals[r * e:(r + 1) * e]] for r in range(num_rounds)]
        return rp, ch, [int(x) for x in fin]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_prover_free(self.h)
            self.h = None


def verify(claimed_sum, degree, num_rounds, transcript, proof_words, ext=1, w=W_FROM_RAW7):
    """Returns (status, challenges, final_claim); status 0 = ok, 1+round = transcript error."""
    cs, pcs = _arr(claimed_sum)
    pw_arr, ppw = _arr(proof_words if len(proof_words) else [0])
    chals = np.zeros(max(num_rounds, 1) * ext, dtype=np.uint64)
    fin = np.zeros(ext, dtype=np.uint64)
    st = lib().or_sumcheck_verify(ext, w, pcs, degree, num_rounds, transcript.h, ppw, len(proof_words),
                                  chals.ctypes.data_as(u64p), fin.ctypes.data_as(u64p))
    ch = [[int(x) for x in chals[r * ext:(r + 1) * ext]] for r in range(num_rounds)]
    return st, ch, [int(x) for x in fin]


def bench_ip(a, b, num_rounds, threads=1, fused=False, domsep=b"effsc-hip/bench/ip"):
    """Times a reference-structured inner-product proof; a, b are consumed.
    fused=False: InnerProductProver's round (two parallel folds, serial
    evaluate); fused=True: the reference's legacy fused fold + evaluate."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    polys = np.zeros(2 * max(num_rounds, 1), dtype=np.uint64)
    chals = np.zeros(max(num_rounds, 1), dtype=np.uint64)
    fin = np.zeros(1, dtype=np.uint64)
    t = lib().or_bench_ip_ex(a.ctypes.data_as(u64p), b.ctypes.data_as(u64p), len(a), num_rounds, threads,
                             1 if fused else 0, domsep, len(domsep), polys.ctypes.data_as(u64p),
                             chals.ctypes.data_as(u64p), fin.ctypes.data_as(u64p))
    return t, polys, chals, int(fin[0])


def bench_ml(v, num_rounds, threads=1, domsep=b"effsc-hip/bench/ml"):
    """Times a MultilinearProver proof with the reference's parallel fused
    kernels (src/multilinear_sumcheck.rs:84-254); v is consumed."""
    v = np.ascontiguousarray(v, dtype=np.uint64)
    polys = np.zeros(max(num_rounds, 1), dtype=np.uint64)
    chals = np.zeros(max(num_rounds, 1), dtype=np.uint64)
    fin = np.zeros(1, dtype=np.uint64)
    t = lib().or_bench_ml_ex(v.ctypes.data_as(u64p), len(v), num_rounds, threads, domsep, len(domsep),
                             polys.ctypes.data_as(u64p), chals.ctypes.data_as(u64p), fin.ctypes.data_as(u64p))
    return t, polys, chals, int(fin[0])
