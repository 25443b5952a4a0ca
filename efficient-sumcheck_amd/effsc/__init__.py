"""effsc (MI355X): GPU-backed Goldilocks sumcheck provers behind the
reference crate's API (compsec-epfl/efficient-sumcheck, `effsc` 0.0.2).

    from effsc import sumcheck, sumcheck_verify, Transcript
    from effsc.provers import InnerProductProver
    proof = sumcheck(InnerProductProver(a, b), v, Transcript(b"ds"))
"""
from .field import CANONICAL, MONT64, P, W_FROM_RAW7, Field
from .proof import SumcheckError, SumcheckProof, SumcheckResult
from .provers import (CoefficientProver, CoefficientProverLSB, EqFactoredProver, GkrProver, InnerProductProver,
                      InnerProductProverLSB, MultilinearProver, MultilinearProverLSB, ProductProver, bit_reverse,
                      fold)
from .runner import noop_hook, sumcheck
from .transcript import SanityTranscript, TestTranscript, Transcript
from .verifier import noop_hook_verify, sumcheck_verify

__all__ = [
    "CANONICAL", "MONT64", "P", "W_FROM_RAW7", "Field", "SumcheckError", "SumcheckProof", "SumcheckResult",
    "CoefficientProver", "CoefficientProverLSB", "EqFactoredProver", "GkrProver", "InnerProductProver",
    "InnerProductProverLSB", "MultilinearProverLSB", "bit_reverse", "MultilinearProver", "ProductProver", "fold",
    "noop_hook", "sumcheck", "SanityTranscript", "TestTranscript", "Transcript", "noop_hook_verify",
    "sumcheck_verify",
]
