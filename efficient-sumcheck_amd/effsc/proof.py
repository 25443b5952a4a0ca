"""Proof and error types (src/proof.rs:17-53, src/verifier.rs:38-46)."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, List


@dataclass
class SumcheckProof:
    """round_polys[j]: EvalsInfty values of round j; challenges r_1..r_v;
    final_value = g(r_1, ..., r_v) (src/proof.rs:17-34)."""

    round_polys: List[List[Any]] = field(default_factory=list)
    challenges: List[Any] = field(default_factory=list)
    final_value: Any = 0


def proof_from_words(polys, chals, fin, rounds, nw, E):
    """A proof from canonical word arrays (numpy uint64): `rounds` round
    polynomials of `nw` values, `rounds` challenges, one final value; an
    element is an int (E = 1) or an E-tuple.  One tolist() per array:
    per-element slicing and int() cost as much host time as a small proof on
    the device."""
    vals = _elements(polys[:rounds * nw * E].tolist(), E)
    rp = [vals[r * nw:(r + 1) * nw] for r in range(rounds)]
    return SumcheckProof(rp, _elements(chals[:rounds * E].tolist(), E), _elements(fin[:E].tolist(), E)[0])


def _elements(words, E):
    if E == 1:
        return words
    it = iter(words)
    return list(zip(*([it] * E)))


@dataclass
class SumcheckResult:
    """Verifier output (src/verifier.rs:38-46)."""

    challenges: List[Any]
    final_claim: Any


class SumcheckError(Exception):
    """SumcheckError variants (src/proof.rs:37-53)."""

    def __init__(self, kind: str, round: int | None = None, expected=None, got=None):
        self.kind = kind
        self.round = round
        self.expected = expected
        self.got = got
        if kind == "ConsistencyCheck":
            msg = f"round {round}: consistency check failed"
        elif kind == "DegreeMismatch":
            msg = f"round {round}: degree mismatch: expected <= {expected}, got {got}"
        elif kind == "FinalEvaluation":
            msg = "final evaluation mismatch"
        elif kind == "TranscriptError":
            msg = f"round {round}: transcript error"
        else:
            msg = f"round {round}: hook error"
        super().__init__(msg)

    def __eq__(self, other):
        return isinstance(other, SumcheckError) and (self.kind, self.round) == (other.kind, other.round)
