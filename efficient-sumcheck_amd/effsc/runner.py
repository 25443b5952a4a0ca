"""runner::sumcheck (src/runner.rs:29-69).

A native prover driven with a native transcript runs the whole loop in the
library's C++ runner (effsc_hip_sumcheck); any other combination runs the
same loop here.  The hook is called after the round's values are sent and
before the challenge is read; `None` is the reference's noop_hook
(src/lib.rs:28-30).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .field import CANONICAL, MONT64, from_mont, to_mont
from .proof import SumcheckProof, proof_from_words
from .provers import NativeProver
from .transcript import NativeTranscript


def noop_hook(round, transcript):
    return None


def _map(x, fn):
    return fn(x) if isinstance(x, int) else tuple(fn(c) for c in x)


def sumcheck(prover, num_rounds, transcript, hook=None) -> SumcheckProof:
    if hook is noop_hook:
        hook = None
    # A prover built with repr=MONT64 speaks arkworks Montgomery words (the
    # reference's F64 encoding, src/tests/fields.rs:26-30); transcripts take
    # and give canonical values.  Its proof is the canonical proof with every
    # element in Montgomery form.
    mont = getattr(prover, "repr", CANONICAL) == MONT64
    if isinstance(prover, NativeProver) and isinstance(transcript, NativeTranscript):
        if transcript.ext != prover.ext:
            raise ValueError("transcript and prover fields differ")
        # the library's runner works on canonical words whatever the prover's
        # boundary encoding
        proof = _native_sumcheck(prover, num_rounds, transcript, hook)
        if not mont:
            return proof
        return SumcheckProof([[_map(x, to_mont) for x in r] for r in proof.round_polys],
                             [_map(c, to_mont) for c in proof.challenges], _map(proof.final_value, to_mont))
    round_polys, challenges = [], []
    prev = None
    for rnd in range(num_rounds):
        evals = prover.round(prev)
        for v in evals:
            transcript.send(_map(v, from_mont) if mont else v)
        round_polys.append(evals)
        if hook is not None:
            hook(rnd, transcript)
        r = transcript.challenge()
        if mont:
            r = _map(r, to_mont)
        challenges.append(r)
        prev = r
    if prev is not None:
        prover.finalize(prev)
    return SumcheckProof(round_polys, challenges, prover.final_value())


def _native_sumcheck(prover, num_rounds, transcript, hook):
    E = prover.ext
    nw = max(prover.degree(), 1)
    polys = np.zeros(max(num_rounds, 1) * nw * E, dtype=np.uint64)
    chals = np.zeros(max(num_rounds, 1) * E, dtype=np.uint64)
    fin = np.zeros(E, dtype=np.uint64)
    errors = []
    if hook is None:
        cb = C.cast(None, N.HOOK)
    else:
        def _cb(ctx, rnd, t):
            try:
                hook(rnd, transcript)
            except BaseException as e:  # surfaced after the native call returns
                errors.append(e)
            return 0
        cb = N.HOOK(_cb)
    N.check(N.lib().effsc_hip_sumcheck(prover.h, num_rounds, transcript.h, cb, None, polys.ctypes.data_as(N.u64p),
                                       chals.ctypes.data_as(N.u64p), fin.ctypes.data_as(N.u64p)))
    if errors:
        raise errors[0]
    return proof_from_words(polys, chals, fin, num_rounds, nw, E)
