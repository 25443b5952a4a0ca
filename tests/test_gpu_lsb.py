"""LSB (pair-split) provers on the GPU against the oracle's restatement of
src/provers/*_lsb.rs, and the device bit-reversal against its definition."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
from effsc import (CoefficientProverLSB, InnerProductProverLSB, MultilinearProverLSB, TestTranscript, Transcript,
                   bit_reverse, sumcheck, sumcheck_verify)
from effsc import _native as N

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
    if ext == 1:
        return words
    out = np.zeros(len(words) * ext, dtype=np.uint64)
    out[::ext] = words
    return out


def rev_perm(v):
    return np.array([int(format(i, f"0{v}b")[::-1], 2) if v else 0 for i in range(1 << v)], dtype=np.int64)


@pytest.mark.parametrize("v", [0, 1, 2, 5, 9, 10, 11, 14, 17])
def test_bit_reverse_matches_definition(v):
    n = 1 << v
    x = O.uniform(n, 40 + v)
    say(f"bit reverse v={v}")
    assert np.array_equal(bit_reverse(x), x[rev_perm(v)])


def test_bit_reverse_planes():
    v, planes = 12, 3
    n = 1 << v
    x = O.uniform(n * planes, 5)
    src, dst = N.u64p(), N.u64p()
    L = N.lib()
    N.check(L.effsc_hip_device_alloc(0, n * planes, C.byref(src)))
    N.check(L.effsc_hip_device_alloc(0, n * planes, C.byref(dst)))
    try:
        N.check(L.effsc_hip_copy(0, src, x.ctypes.data_as(N.u64p), n * planes))
        N.check(L.effsc_hip_bit_reverse(0, dst, src, n, planes))
        out = np.zeros(n * planes, dtype=np.uint64)
        N.check(L.effsc_hip_copy(0, out.ctypes.data_as(N.u64p), dst, n * planes))
        r = rev_perm(v)
        for p in range(planes):
            assert np.array_equal(out[p * n:(p + 1) * n], x[p * n:(p + 1) * n][r])
        assert L.effsc_hip_bit_reverse(0, dst, src, n - 1, 1) == 1  # not a power of two
    finally:
        L.effsc_hip_device_free(0, src)
        L.effsc_hip_device_free(0, dst)


def make(kind, ext, input_ext, tabs):
    if kind == "ml":
        return (MultilinearProverLSB(tabs[0], ext=ext, input_ext=input_ext),
                O.Prover.multilinear_lsb(lift(tabs[0], ext // input_ext), ext=ext))
    if kind == "ip":
        return (InnerProductProverLSB(tabs[0], tabs[1], ext=ext, input_ext=input_ext),
                O.Prover.inner_product_lsb(lift(tabs[0], ext // input_ext), lift(tabs[1], ext // input_ext), ext=ext))
    k = int(kind[1:])
    return (CoefficientProverLSB(tabs[:k], ext=ext, input_ext=input_ext),
            O.Prover.product_lsb(tabs[:k], ext=ext, tables_in_ext=(input_ext == ext and ext > 1)))


@pytest.mark.parametrize("kind", ["ml", "ip", "p3"])
@pytest.mark.parametrize("ext,input_ext", [(1, 1), (3, 1), (3, 3), (2, 1), (2, 2)])
@pytest.mark.parametrize("v", [0, 1, 2, 3, 6, 10, 13])
def test_lsb_parity(kind, ext, input_ext, v):
    n = 1 << v
    k = {"ml": 1, "ip": 2}.get(kind) or 3
    tabs = [O.uniform(n * input_ext, 7000 + 10 * v + i) for i in range(k)]
    say(f"lsb kind={kind} ext={ext} input_ext={input_ext} v={v}")
    g, o = make(kind, ext, input_ext, tabs)
    pg = sumcheck(g, v, Transcript(b"lsb-test", ext=ext))
    po = o.sumcheck(v, O.Transcript.shake(b"lsb-test"))
    assert tuple(as_oracle(pg, ext)) == tuple(po)
    for i in range(k):
        assert np.array_equal(g.table_words(i), o.table(i)), f"table {i}"


@pytest.mark.parametrize("kind", ["ml", "ip"])
def test_lsb_partial_rounds_post_state_and_large(kind):
    # post-state after a partial proof is in the reference's pair-split order
    v = 20
    n = 1 << v
    tabs = [O.uniform(n, 90 + i) for i in range(2)]
    g, o = make(kind, 1, 1, tabs)
    pg = sumcheck(g, 7, TestTranscript(4))
    po = o.sumcheck(7, O.Transcript.rng(4))
    assert tuple(as_oracle(pg, 1)) == tuple(po)
    assert np.array_equal(g.table_words(0), o.table(0))
    g.reset()
    full = sumcheck(g, v, Transcript(b"lsb-large"))
    g2, o2 = make(kind, 1, 1, tabs)
    assert tuple(as_oracle(full, 1)) == tuple(o2.sumcheck(v, O.Transcript.shake(b"lsb-large")))


def test_lsb_verifier_accepts():
    # src/provers/inner_product_lsb.rs tests: the claim chain closes
    v = 9
    a, b = O.uniform(1 << v, 1), O.uniform(1 << v, 2)
    claim = sum(int(x) * int(y) for x, y in zip(a, b)) % P
    proof = sumcheck(InnerProductProverLSB(a, b), v, Transcript(b"lsb-v"))
    res = sumcheck_verify(claim, 2, v, Transcript(b"lsb-v").verifier(proof))
    assert res.final_claim == proof.final_value


@pytest.mark.parametrize("ext,input_ext", [(1, 1), (2, 2), (3, 1), (3, 3)])
@pytest.mark.parametrize("n", [0, 3, 5, 6, 7, 11, 12, 100, 1000, 4097])
def test_lsb_multilinear_any_length(n, ext, input_ext):
    # src/provers/multilinear_lsb.rs:59-126: the round value counts an
    # unpaired last element, the fold drops it; native pair-split rounds
    rounds = max(1, (n - 1).bit_length()) if n > 1 else 1
    t = O.uniform(n * input_ext, 7700 + n)
    say(f"lsb any length n={n} ext={ext}/{input_ext} rounds={rounds}")
    g = MultilinearProverLSB(t, ext=ext, input_ext=input_ext)
    o = O.Prover.multilinear_lsb(lift(t, ext // input_ext), ext=ext)
    pg = sumcheck(g, rounds, Transcript(b"lsb-any", ext=ext))
    po = o.sumcheck(rounds, O.Transcript.shake(b"lsb-any"))
    assert tuple(as_oracle(pg, ext)) == tuple(po)
    assert np.array_equal(g.table_words(0), o.table(0))
    # partial rounds, then the rest, on a fresh pair
    if rounds >= 3:
        g2 = MultilinearProverLSB(t, ext=ext, input_ext=input_ext)
        o2 = O.Prover.multilinear_lsb(lift(t, ext // input_ext), ext=ext)
        tg, to = TestTranscript(3, ext=ext), O.Transcript.rng(3)
        assert tuple(as_oracle(sumcheck(g2, 2, tg), ext)) == tuple(o2.sumcheck(2, to))
        assert np.array_equal(g2.table_words(0), o2.table(0))


def test_lsb_products_reject_other_lengths():
    # the reference's InnerProductProverLSB / CoefficientProverLSB panic on
    # these lengths
    with pytest.raises(N.NativeError):
        InnerProductProverLSB([1, 2, 3], [4, 5, 6])
    with pytest.raises(N.NativeError):
        CoefficientProverLSB([O.uniform(12, 1), O.uniform(12, 2)])
