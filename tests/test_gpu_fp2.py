"""F_p^2 = F_p[X]/(X^2 - w), the reference's F64Ext2 (src/tests/fields.rs:41-70),
on the device: the extension free fold (fold<F>, src/fold.rs:14), proofs
accepted by the verifier, and an inner product and a multilinear proof of
2^20 elements against the oracle.  The per-prover parity cases over F_p^2
(pow2, non-pow2, device loop, product of k, EqFactored, GKR, LSB) run in
test_gpu_parity.py, test_gpu_gkr.py and test_gpu_lsb.py beside the other
fields.
"""
import sys

import numpy as np
import pytest

import oracle as O
import effsc
from effsc import CoefficientProver, InnerProductProver, MultilinearProver, Transcript, sumcheck, sumcheck_verify
from effsc.field import W_FROM_RAW7, Field

pytestmark = pytest.mark.gpu
P = O.P


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def as_oracle(proof, ext):
    rp = [[[x] if ext == 1 else list(x) for x in r] for r in proof.round_polys]
    ch = [[c] if ext == 1 else list(c) for c in proof.challenges]
    fin = [proof.final_value] if ext == 1 else list(proof.final_value)
    return rp, ch, fin


def lift(words, ext):
    out = np.zeros(len(words) * ext, dtype=np.uint64)
    out[::ext] = words
    return out


def fold_definition(F, vals, wt):
    # src/multilinear_sumcheck.rs:132-178: half = next_pow2(n) / 2, new[k] =
    # v[k] + w (v[k + half] - v[k]) for paired k, v[k] (1 - w) for the rest
    n = len(vals)
    half = (1 << (n - 1).bit_length()) // 2
    out = []
    for k in range(half):
        lo = vals[k]
        hi = vals[k + half] if k + half < n else F.zero()
        out.append(F.add(lo, F.mul(wt, F.sub(hi, lo))))
    return out


@pytest.mark.parametrize("ext", [2, 3])
@pytest.mark.parametrize("n", [2, 3, 7, 8, 1000, 4097])
def test_extension_free_fold_matches_definition(ext, n):
    F = Field(ext, W_FROM_RAW7)
    words = O.uniform(n * ext, 40 + n + ext)
    vals = [tuple(int(x) for x in words[i * ext:(i + 1) * ext]) for i in range(n)]
    for wt in [tuple(int(x) for x in O.uniform(ext, 77 + n)), (P - 1,) + (P - 1,) * (ext - 1), (0,) * ext,
               (1,) + (0,) * (ext - 1)]:
        say(f"fold ext={ext} n={n} weight={wt}")
        got = effsc.fold(vals, wt, ext=ext)
        assert got == fold_definition(F, vals, wt)


def test_extension_free_fold_lifted_base_equals_base_fold():
    n = 1 << 12
    base = [int(x) for x in O.uniform(n, 5)]
    wt = 123456789
    assert effsc.fold([(x, 0) for x in base], (wt, 0), ext=2) == [(x, 0) for x in effsc.fold(base, wt)]


@pytest.mark.parametrize("kind,k", [("ml", 1), ("ip", 2), ("p3", 3)])
def test_fp2_proofs_accepted_by_verifier(kind, k):
    v = 10
    F = Field(2, W_FROM_RAW7)
    tabs = [O.uniform((1 << v) * 2, 90 + i) for i in range(k)]
    el = [[(int(t[2 * i]), int(t[2 * i + 1])) for i in range(1 << v)] for t in tabs]
    claim = F.zero()
    for i in range(1 << v):
        x = F.one()
        for t in el:
            x = F.mul(x, t[i])
        claim = F.add(claim, x)
    g = (MultilinearProver(tabs[0], ext=2) if k == 1 else InnerProductProver(*tabs, ext=2) if k == 2
         else CoefficientProver(tabs, ext=2))
    say(f"verify fp2 kind={kind} v={v}")
    proof = sumcheck(g, v, Transcript(b"verify2", ext=2))
    res = sumcheck_verify(claim, k, v, Transcript(b"verify2", ext=2).verifier(proof))
    assert res.final_claim == proof.final_value
    assert res.challenges == proof.challenges


@pytest.mark.parametrize("kind,input_ext", [("ip", 1), ("ip", 2), ("ml", 2)])
def test_fp2_large_parity(kind, input_ext):
    v = 20
    n = 1 << v
    k = 2 if kind == "ip" else 1
    tabs = [O.uniform(n * input_ext, 0x5EED0F00 + i) for i in range(k)]
    say(f"fp2 large kind={kind} input_ext={input_ext} v={v}")
    if kind == "ip":
        g = InnerProductProver(tabs[0], tabs[1], ext=2, input_ext=input_ext)
        o = O.Prover.inner_product(*[t if input_ext == 2 else lift(t, 2) for t in tabs], ext=2)
    else:
        g = MultilinearProver(tabs[0], ext=2, input_ext=input_ext)
        o = O.Prover.multilinear(tabs[0], ext=2)
    pg = sumcheck(g, v, Transcript(b"fp2-large", ext=2))
    po = o.sumcheck(v, O.Transcript.shake(b"fp2-large"))
    assert tuple(as_oracle(pg, 2)) == tuple(po)
    for i in range(k):
        assert np.array_equal(g.table_words(i), o.table(i))


@pytest.mark.parametrize("env", ["EFFSC_HIP_ONE_ROUND", "EFFSC_HIP_EXT3_PASS"])
@pytest.mark.parametrize("kind,ext,input_ext", [("ml", 2, 2), ("ip", 2, 1), ("ml", 3, 1), ("ip", 3, 3)])
@pytest.mark.parametrize("v", [2, 5, 9, 14, 17])
def test_extension_device_loop_forms(kind, ext, input_ext, v, env, monkeypatch):
    # the extension two-round passes (default over F_p^2, EFFSC_HIP_EXT3_PASS
    # over F_p^3) and one round per launch give the same proofs
    monkeypatch.setenv(env, "1")
    monkeypatch.setenv("EFFSC_HIP_TAIL", "64")
    n = 1 << v
    k = 2 if kind == "ip" else 1
    tabs = [O.uniform(n * input_ext, 0xE70 + v + i) for i in range(k)]
    say(f"ext device loop {env} kind={kind} ext={ext}/{input_ext} v={v}")
    lifted = [t if input_ext == ext else lift(t, ext) for t in tabs]
    if kind == "ip":
        g = InnerProductProver(tabs[0], tabs[1], ext=ext, input_ext=input_ext)
        o = O.Prover.inner_product(lifted[0], lifted[1], ext=ext)
    else:
        g = MultilinearProver(tabs[0], ext=ext, input_ext=input_ext)
        o = O.Prover.multilinear(lifted[0], ext=ext)
    pg = sumcheck(g, v, Transcript(b"ext-forms", ext=ext))
    po = o.sumcheck(v, O.Transcript.shake(b"ext-forms"))
    assert tuple(as_oracle(pg, ext)) == tuple(po)
    for i in range(k):
        assert np.array_equal(g.table_words(i), o.table(i))
    # a partial device run, then the rest
    g2 = InnerProductProver(tabs[0], tabs[1], ext=ext, input_ext=input_ext) if kind == "ip" else \
        MultilinearProver(tabs[0], ext=ext, input_ext=input_ext)
    o2 = O.Prover.inner_product(lifted[0], lifted[1], ext=ext) if kind == "ip" else O.Prover.multilinear(lifted[0], ext=ext)
    r1 = max(1, v // 2)
    t1, u1 = Transcript(b"ext-part", ext=ext), O.Transcript.shake(b"ext-part")
    assert tuple(as_oracle(sumcheck(g2, r1, t1), ext)) == tuple(o2.sumcheck(r1, u1))
    assert tuple(as_oracle(sumcheck(g2, v - r1, t1), ext)) == tuple(o2.sumcheck(v - r1, u1))
