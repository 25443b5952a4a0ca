"""C-ABI checks that need no GPU: the library loads, exports every entry
point declared in include/effsc_hip.h, and its host-side protocol code
(transcripts, verifier) matches the oracle and the golden vectors."""
import os
import re

import numpy as np
import pytest

import oracle as O
from effsc import _native as N
from effsc import Transcript, TestTranscript, sumcheck_verify, SumcheckError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = O.P


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "effsc_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(effsc_hip_\w+)\s*\(", src)) - {"effsc_hip_hook"})


def test_library_exports_every_declared_symbol():
    L = N.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    assert set(N.EXPORTS) == set(syms)


def test_host_transcript_matches_golden(golden):
    for case in golden["transcript"]:
        if "rng_seed" in case:
            t = TestTranscript(case["rng_seed"])
            assert [[t.challenge()] for _ in case["challenges"]] == case["challenges"]
            continue
        ext = case["ext"]
        t = Transcript(bytes.fromhex(case["domsep"]), ext=ext)
        for op, val in case["ops"]:
            if op == "send":
                t.send(val[0] if ext == 1 else tuple(val))
            else:
                got = t.challenge()
                assert (list(got) if ext > 1 else [got]) == val


def _golden_proof_tape(case):
    ext = case["ext"]
    rps = [[x[0] if ext == 1 else tuple(x) for x in rp] for rp in case["round_polys"]]
    from effsc import SumcheckProof
    return SumcheckProof(rps, [], None)


def test_host_verifier_accepts_golden_proofs(golden):
    import brute as B
    for case in golden["proofs"]:
        if case["transcript"] != "shake" or case["kind"] == "eq":
            continue
        ext, n = case["ext"], case["n"]
        kind = case["kind"]
        k = 1 if kind == "ml" else 2 if kind == "ip" else int(kind[-1])
        tabs = [B.uniform_table(n, s) for s in case["seeds"][:k]]
        F = B.Ext(ext)
        claim = B.claimed_sum_product(F, tabs)
        claim = claim[0] if ext == 1 else tuple(claim)
        proof = _golden_proof_tape(case)
        t = Transcript(case["domsep"].encode(), ext=ext).verifier(proof)
        res = sumcheck_verify(claim, k, case["v"], t)
        fin = case["final"]
        assert res.final_claim == (fin[0] if ext == 1 else tuple(fin))
        assert [[c] if ext == 1 else list(c) for c in res.challenges] == case["challenges"]


def test_host_verifier_transcript_error():
    t = Transcript(b"x").verifier([1, 2, 3])
    with pytest.raises(SumcheckError) as e:
        sumcheck_verify(5, 2, 2, t)
    assert e.value.kind == "TranscriptError" and e.value.round == 1


def test_host_verifier_hook_error():
    t = Transcript(b"x").verifier([1, 2, 3, 4])

    def hook(rnd, tr):
        if rnd == 1:
            raise RuntimeError("pow failed")

    with pytest.raises(SumcheckError) as e:
        sumcheck_verify(5, 2, 2, t, hook)
    assert e.value.kind == "HookError" and e.value.round == 1


def test_host_verifier_matches_oracle_verifier_degree3_ext3():
    rng = np.random.default_rng(3)
    words = [int(x) % P for x in rng.integers(0, 2**63, size=3 * 3 * 5)]
    claim = [7, 8, 9]
    st, och, ofin = O.verify(claim, 3, 5, O.Transcript.shake(b"d3"), words, ext=3)
    assert st == 0
    tape = [tuple(words[i:i + 3]) for i in range(0, len(words), 3)]
    t = Transcript(b"d3", ext=3).verifier(tape)
    res = sumcheck_verify(tuple(claim), 3, 5, t)
    assert list(res.final_claim) == ofin
    assert [list(c) for c in res.challenges] == och
