"""Transcripts (src/transcript/transcript.rs:11-43).

`Transcript`      - Fiat-Shamir transcript in the native library (SHAKE256 over
                    a fixed byte encoding; the reference binds spongefish here,
                    src/transcript/spongefish.rs:17-44).
`TestTranscript`  - native TestTranscript/SanityTranscript: sends are no-ops
                    and challenge = from_u64(splitmix64 next), lifted
                    (src/transcript/sanity.rs:12-63).
Verifier side: `.verifier(proof)` returns a transcript of the same kind that
`receive()`s prover messages from a proof's round polynomials, absorbing
each as `send` does (VerifierTranscript::receive).
Any Python object with send/challenge (prover) or receive/challenge
(verifier) also works with the Python runner and verifier.
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from .field import P


class NativeTranscript:
    def __init__(self, handle, ext, factory):
        if not handle:
            raise N.NativeError(N.lib().effsc_hip_last_error().decode() or "transcript allocation failed")
        self.h = handle
        self.ext = ext
        self._factory = factory
        self._tape = None
        self._cursor = 0

    def _words(self, value):
        if self.ext == 1:
            return np.array([int(value) % P], dtype=np.uint64)
        return np.array([int(c) % P for c in value], dtype=np.uint64)

    def send(self, value):
        w = self._words(value)
        N.check(N.lib().effsc_hip_transcript_send(self.h, self.ext, w.ctypes.data_as(N.u64p)))

    def challenge(self):
        out = np.zeros(self.ext, dtype=np.uint64)
        N.check(N.lib().effsc_hip_transcript_challenge(self.h, self.ext, out.ctypes.data_as(N.u64p)))
        return int(out[0]) if self.ext == 1 else tuple(int(x) for x in out)

    # ---- verifier side
    def verifier(self, proof):
        """A fresh transcript of this kind reading prover messages from
        `proof` (a SumcheckProof or a flat list of elements)."""
        t = self._factory()
        flat = proof.round_polys if hasattr(proof, "round_polys") else [proof]
        t._tape = [x for rp in flat for x in rp]
        t._cursor = 0
        return t

    def receive(self):
        if self._tape is None or self._cursor >= len(self._tape):
            raise IndexError("transcript exhausted")
        v = self._tape[self._cursor]
        self._cursor += 1
        self.send(v)
        return v

    def proof_words(self):
        words = []
        for v in (self._tape or []):
            words.extend([int(v)] if self.ext == 1 else [int(c) for c in v])
        return np.array(words, dtype=np.uint64)

    def close(self):
        if getattr(self, "h", None):
            N.lib().effsc_hip_transcript_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Transcript(NativeTranscript):
    def __init__(self, domain_separator: bytes = b"effsc-hip", ext: int = 1):
        ds = bytes(domain_separator)
        super().__init__(N.lib().effsc_hip_transcript_new(ds, len(ds)), ext,
                         lambda: Transcript(ds, ext))


class TestTranscript(NativeTranscript):
    __test__ = False  # not a pytest class

    def __init__(self, seed: int = 0, ext: int = 1):
        super().__init__(N.lib().effsc_hip_transcript_new_test(seed & (2**64 - 1)), ext,
                         lambda: TestTranscript(seed, ext))


SanityTranscript = TestTranscript
