"""verifier::sumcheck_verify (src/verifier.rs:61-200).

Returns SumcheckResult(challenges, final_claim) or raises SumcheckError.  The
oracle check final_claim == g(r) stays with the caller, as in the reference.
A native transcript built with `.verifier(proof)` runs in the library's C++
verifier; other transcripts (objects with receive/challenge) run here.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .field import P, W_FROM_RAW7, Field
from .proof import SumcheckError, SumcheckResult
from .transcript import NativeTranscript


def noop_hook_verify(round, transcript):
    return None


def _elem_words(x, ext):
    return [int(x) % P] if ext == 1 else [int(c) % P for c in x]


def sumcheck_verify(claimed_sum, expected_degree, num_rounds, transcript, hook=None, *, ext=None,
                    w=W_FROM_RAW7) -> SumcheckResult:
    if hook is noop_hook_verify:
        hook = None
    if ext is None:
        ext = getattr(transcript, "ext", 1 if isinstance(claimed_sum, int) else len(claimed_sum))
    if isinstance(transcript, NativeTranscript) and transcript._tape is not None:
        return _native_verify(claimed_sum, expected_degree, num_rounds, transcript, hook, ext, w)
    F = Field(ext, w)
    claim = claimed_sum
    d = expected_degree
    challenges = []
    for rnd in range(num_rounds):
        recv = []
        for _ in range(max(d, 1)):
            try:
                recv.append(transcript.receive())
            except Exception:
                raise SumcheckError("TranscriptError", rnd)
        h0 = recv[0]
        h1 = F.sub(claim, h0)
        if hook is not None:
            try:
                hook(rnd, transcript)
            except SumcheckError:
                raise
            except Exception:
                raise SumcheckError("HookError", rnd)
        r = transcript.challenge()
        challenges.append(r)
        if d == 0:
            claim = h0
            continue
        h_inf = recv[1] if d >= 2 else F.sub(h1, h0)
        q = [h0, F.sub(h1, h_inf)]
        for i in range(2, d):
            i_d = F.from_u64(pow(i, d, P))
            q.append(F.sub(recv[i], F.mul(h_inf, i_d)))
        q_r = _evaluate_from_evals(F, q, r)
        r_d = F.one()
        for _ in range(d):
            r_d = F.mul(r_d, r)
        claim = F.add(q_r, F.mul(h_inf, r_d))
    return SumcheckResult(challenges, claim)


def _evaluate_from_evals(F, evals, r):
    d = len(evals)
    if d == 0:
        return F.zero()
    if d == 1:
        return evals[0]
    if d == 2:
        return F.add(evals[0], F.mul(r, F.sub(evals[1], evals[0])))
    result = F.zero()
    for i in range(d):
        num, den = F.one(), F.one()
        for j in range(d):
            if j != i:
                num = F.mul(num, F.sub(r, F.from_u64(j)))
                den = F.mul(den, F.from_u64((i - j) % P))
        if not F.is_zero(den):
            result = F.add(result, F.mul(F.mul(evals[i], num), F.inv(den)))
    return result


def _native_verify(claimed_sum, d, num_rounds, t, hook, ext, w):
    words = t.proof_words()
    cs = np.array(_elem_words(claimed_sum, ext), dtype=np.uint64)
    chals = np.zeros(max(num_rounds, 1) * ext, dtype=np.uint64)
    fin = np.zeros(ext, dtype=np.uint64)
    err_round = C.c_size_t(0)
    err_kind = C.c_int(0)
    if hook is None:
        cb = C.cast(None, N.HOOK)
    else:
        def _cb(ctx, rnd, tp):
            try:
                hook(rnd, t)
                return 0
            except Exception:
                return 1
        cb = N.HOOK(_cb)
    wp = words.ctypes.data_as(N.u64p) if len(words) else None
    rc = N.lib().effsc_hip_sumcheck_verify(ext, w, cs.ctypes.data_as(N.u64p), d, num_rounds, t.h, wp, len(words),
                                            cb, None, chals.ctypes.data_as(N.u64p), fin.ctypes.data_as(N.u64p),
                                            C.byref(err_round), C.byref(err_kind))
    if rc == 5:
        raise SumcheckError("TranscriptError" if err_kind.value == 1 else "HookError", err_round.value)
    N.check(rc)
    el = (lambda a: int(a[0])) if ext == 1 else (lambda a: tuple(int(x) for x in a[:ext]))
    return SumcheckResult([el(chals[r * ext:(r + 1) * ext]) for r in range(num_rounds)], el(fin))
