"""Parity of the HIP provers with the oracle and the golden vectors.

Every test runs through the C ABI (libeffsc_hip.so) on the GPU and compares
round polynomials, challenges, final values and post-state bit for bit.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
import effsc
from effsc import (CoefficientProver, EqFactoredProver, InnerProductProver, MultilinearProver, TestTranscript,
                   Transcript, sumcheck, sumcheck_verify)
from effsc import _native as N
from effsc.field import MONT64, to_mont

pytestmark = pytest.mark.gpu
P = O.P


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def el(words, ext):
    return int(words[0]) if ext == 1 else tuple(int(x) for x in words)


def as_oracle(proof, ext):
    rp = [[[x] if ext == 1 else list(x) for x in r] for r in proof.round_polys]
    ch = [[c] if ext == 1 else list(c) for c in proof.challenges]
    fin = [proof.final_value] if ext == 1 else list(proof.final_value)
    return rp, ch, fin


def lift(words, d):
    """Base words -> extension words (c0 = x, other components 0); d = 1 keeps them."""
    if d == 1:
        return words
    out = np.zeros(len(words) * d, dtype=np.uint64)
    out[::d] = words
    return out


def make_pair(kind, ext, input_ext, tabs, wpt=None):
    """(gpu prover, oracle prover) over the same canonical input words."""
    if kind == "ml":
        # base inputs of an extension prover: the oracle prover over the
        # lifted words (same ML padding rules for any length)
        g = MultilinearProver(tabs[0], ext=ext, input_ext=input_ext)
        o = O.Prover.multilinear(lift(tabs[0], ext // input_ext), ext=ext)
    elif kind == "ip":
        g = InnerProductProver(tabs[0], tabs[1], ext=ext, input_ext=input_ext)
        o = O.Prover.inner_product(lift(tabs[0], ext // input_ext), lift(tabs[1], ext // input_ext), ext=ext)
    elif kind.startswith("p"):
        k = int(kind[1:])
        g = CoefficientProver(tabs[:k], ext=ext, input_ext=input_ext)
        o = O.Prover.product(tabs[:k], ext=ext, tables_in_ext=(input_ext == ext and ext > 1))
    else:
        g = EqFactoredProver(wpt, tabs[0], ext=ext)
        o = O.Prover.eq_factored(wpt, tabs[0], ext=ext)
    return g, o


def check_proof(kind, ext, input_ext, n, v, seed, transcript="shake", rounds=None):
    rounds = v if rounds is None else rounds
    k = {"ml": 1, "ip": 2, "eq": 1}.get(kind) or int(kind[1:])
    tabs = [O.uniform(n * input_ext, seed + i) for i in range(k)]
    wpt = O.uniform(v * ext, seed + 100) if kind == "eq" else None
    say(f"case kind={kind} ext={ext} input_ext={input_ext} n={n} v={v} rounds={rounds} tr={transcript}")
    g, o = make_pair(kind, ext, input_ext, tabs, wpt)
    if transcript == "shake":
        tg, to = Transcript(b"effsc-hip/test", ext=ext), O.Transcript.shake(b"effsc-hip/test")
    else:
        tg, to = TestTranscript(seed, ext=ext), O.Transcript.rng(seed)
    pg = sumcheck(g, rounds, tg)
    po = o.sumcheck(rounds, to)
    assert tuple(as_oracle(pg, ext)) == tuple(po)
    # post-state
    for i in range(len(tabs) if kind != "eq" else 3):
        assert np.array_equal(g.table_words(i), o.table(i)), f"table {i}"
    return g, pg


# ------------------------------------------------------------------ golden
def test_golden_proofs(golden):
    for case in golden["proofs"]:
        ext, n, v, kind = case["ext"], case["n"], case["v"], case["kind"]
        seeds = case["seeds"]
        say("golden", kind, ext, n, v)
        if kind == "ml":
            tabs = [O.uniform(n, seeds[0])]
            g = MultilinearProver(tabs[0], ext=ext, input_ext=1)
        elif kind == "ip":
            tabs = [O.uniform(n, s) for s in seeds[:2]]
            g = InnerProductProver(tabs[0], tabs[1], ext=ext, input_ext=1)
        elif kind.startswith("product"):
            k = int(kind[-1])
            g = CoefficientProver([O.uniform(n, s) for s in seeds[:k]], ext=ext, input_ext=1)
        else:
            g = EqFactoredProver(O.uniform(v * ext, seeds[1]), O.uniform(n, seeds[0]), ext=ext)
        t = (Transcript(case["domsep"].encode(), ext=ext) if case["transcript"] == "shake"
             else TestTranscript(7, ext=ext))
        pr = sumcheck(g, v, t)
        rp, ch, fin = as_oracle(pr, ext)
        assert rp == case["round_polys"]
        assert ch == case["challenges"]
        assert fin == case["final"]


# ------------------------------------------------------------- oracle parity
@pytest.mark.parametrize("kind", ["ml", "ip"])
@pytest.mark.parametrize("ext,input_ext", [(1, 1), (3, 3), (3, 1), (2, 2), (2, 1)])
@pytest.mark.parametrize("v", [0, 1, 2, 3, 4, 5, 8, 11, 14])
def test_pow2_parity(kind, ext, input_ext, v):
    check_proof(kind, ext, input_ext, 1 << v, v, 1000 + v)


@pytest.mark.parametrize("kind", ["ml", "ip"])
@pytest.mark.parametrize("n", [1, 3, 5, 13, 33, 100, 1000, 4097])
def test_non_pow2_parity(kind, n):
    v = (n - 1).bit_length()
    check_proof(kind, 1, 1, n, v, 77 + n)
    check_proof(kind, 3, 3, n, v, 78 + n, transcript="rng")
    check_proof(kind, 2, 2, n, v, 79 + n)
    check_proof(kind, 2, 1, n, v, 80 + n, transcript="rng")


@pytest.mark.parametrize("k", [1, 2, 3, 4])
@pytest.mark.parametrize("ext,input_ext", [(1, 1), (3, 1), (3, 3), (2, 1), (2, 2)])
@pytest.mark.parametrize("v", [1, 2, 3, 6, 12])
def test_product_parity(k, ext, input_ext, v):
    check_proof(f"p{k}", ext, input_ext, 1 << v, v, 500 + 10 * k + v)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
@pytest.mark.parametrize("ext,input_ext", [(1, 1), (3, 1), (3, 3), (2, 1), (2, 2)])
@pytest.mark.parametrize("n", [3, 5, 6, 7, 12, 13, 100, 1000, 4097])
def test_product_non_pow2_parity(k, ext, input_ext, n):
    # msb_fold_vec (src/provers/coefficient.rs:131-140): half = n/2, an odd
    # last element is dropped; the rounds run on the per-round path
    rounds = max(1, (n - 1).bit_length() - 1)
    check_proof(f"p{k}", ext, input_ext, n, rounds, 3000 + 10 * k + n, rounds=rounds)


@pytest.mark.parametrize("ext", [1, 2, 3])
@pytest.mark.parametrize("v", [0, 1, 2, 3, 5, 8, 13])
def test_eq_factored_parity(ext, v):
    check_proof("eq", ext, ext, 1 << v, v, 900 + v)


@pytest.mark.parametrize("ext,v", [(1, 13), (1, 14), (1, 15), (1, 17), (1, 18), (1, 20), (1, 23), (3, 11), (3, 12),
                                   (3, 15), (3, 17), (3, 18), (3, 20), (3, 21), (2, 11), (2, 12), (2, 14), (2, 17),
                                   (2, 19), (2, 20)])
def test_eq_factored_device_loop(ext, v, monkeypatch):
    # large p: multi-workgroup rounds with in-reduction eq folds, then the
    # one-workgroup tail; against the oracle and the per-round host loop
    # base field: two rounds per pass against EQ = eq_l x eq_r, in the
    # split-value form while both variables of a pass are left ones, else
    # computed on the fly (v = 23 has a pass across the left/right split,
    # v_L = 11), then the register tail over p and EQ from 2^11..2^14
    # elements; over F_p^2 and F_p^3 two rounds per pass in the split-value
    # form while both variables of a pass are left ones (lazy Karatsuba sums
    # against the prepared eq_r; from v = 17 on), then one round per launch,
    # then the F_p^3 register tail over p and EQ from 2^10 elements (F_p^2:
    # the one-workgroup eq tail);
    # EFFSC_HIP_EQ_DIRECT computes EQ per element in every pass (over the
    # extensions: one round per launch in the pair form),
    # EFFSC_HIP_EQ_TAIL keeps the one-workgroup eq tail,
    # EFFSC_HIP_ONE_ROUND keeps one round per launch,
    # EFFSC_HIP_EQ_FOLD_LIMIT=0 folds the eq tables in their own launches
    # instead of inside the transcript kernel
    check_proof("eq", ext, ext, 1 << v, v, 950 + v)
    monkeypatch.setenv("EFFSC_HIP_EQ_DIRECT", "1")
    check_proof("eq", ext, ext, 1 << v, v, 950 + v)
    monkeypatch.delenv("EFFSC_HIP_EQ_DIRECT")
    monkeypatch.setenv("EFFSC_HIP_EQ_FOLD_LIMIT", "0")
    check_proof("eq", ext, ext, 1 << v, v, 950 + v)
    monkeypatch.delenv("EFFSC_HIP_EQ_FOLD_LIMIT")
    if ext >= 2:
        monkeypatch.setenv("EFFSC_HIP_EQ_TAIL", "1")
        check_proof("eq", ext, ext, 1 << v, v, 950 + v, transcript="rng")
        monkeypatch.delenv("EFFSC_HIP_EQ_TAIL")
        check_proof("eq", ext, ext, 1 << v, v, 950 + v, transcript="rng")
    if ext == 1:
        # the one-workgroup eq tail instead of the register tail over p and EQ
        monkeypatch.setenv("EFFSC_HIP_EQ_TAIL", "1")
        check_proof("eq", ext, ext, 1 << v, v, 950 + v)
        check_proof("eq", ext, ext, 1 << v, v, 950 + v, transcript="rng")
        monkeypatch.delenv("EFFSC_HIP_EQ_TAIL")
        check_proof("eq", ext, ext, 1 << v, v, 950 + v, transcript="rng")
        monkeypatch.setenv("EFFSC_HIP_ONE_ROUND", "1")
        check_proof("eq", ext, ext, 1 << v, v, 950 + v)
        monkeypatch.delenv("EFFSC_HIP_ONE_ROUND")
    monkeypatch.setenv("EFFSC_HIP_HOST_LOOP", "1")
    check_proof("eq", ext, ext, 1 << v, v, 950 + v)
    monkeypatch.delenv("EFFSC_HIP_HOST_LOOP")
    # a hook forces the host loop too; a partial run is per round
    wpt = O.uniform(v * ext, 960 + v)
    tab = O.uniform((1 << v) * ext, 961 + v)
    g = EqFactoredProver(wpt, tab, ext=ext)
    part = sumcheck(g, 3, Transcript(b"eq-part", ext=ext))
    rest = sumcheck(g, v - 3, Transcript(b"eq-part-2", ext=ext))
    o = O.Prover.eq_factored(wpt, tab, ext=ext)
    po = o.sumcheck(3, O.Transcript.shake(b"eq-part"))
    assert tuple(as_oracle(part, ext)) == tuple(po)
    po2 = o.sumcheck(v - 3, O.Transcript.shake(b"eq-part-2"))
    assert tuple(as_oracle(rest, ext)) == tuple(po2)


@pytest.mark.parametrize("kind", ["ml", "ip", "eq"])
@pytest.mark.parametrize("v", [12, 16, 20])
def test_proof_block_from_the_tail(kind, v, monkeypatch):
    # the pipelined tail writes the proof block into the pinned host staging
    # itself; EFFSC_HIP_PROOF_COPY=1 keeps the separate copy: both give the
    # oracle's proof, twice in a row on one prover
    for copy in (None, "1"):
        if copy:
            monkeypatch.setenv("EFFSC_HIP_PROOF_COPY", copy)
        for tr in ("shake", "rng"):
            g, _ = check_proof(kind, 1, 1, 1 << v, v, 5600 + v, transcript=tr)
        monkeypatch.delenv("EFFSC_HIP_PROOF_COPY", raising=False)


@pytest.mark.parametrize("v", [14, 17, 20, 23])
def test_eq_factored_fused_small_passes(v, monkeypatch):
    # base-field split-value passes of at most 128 workgroups take the previous
    # pass's transcript step and eq_l folds into their launch
    # (k_eq_pass_sv_fs): against the oracle and the unfused loop, with both
    # transcripts and a second proof after reset (the eq_l buffers swap)
    wpt = O.uniform(v, 970 + v)
    tab = O.uniform(1 << v, 971 + v)
    for tr in ("shake", "rng"):
        g = EqFactoredProver(wpt, tab)
        o = O.Prover.eq_factored(wpt, tab)
        mk = (lambda: Transcript(b"eq-fused")) if tr == "shake" else (lambda: TestTranscript(5))
        po = o.sumcheck(v, O.Transcript.shake(b"eq-fused") if tr == "shake" else O.Transcript.rng(5))
        first = sumcheck(g, v, mk())
        assert tuple(as_oracle(first, 1)) == tuple(po)
        g.reset()
        again = sumcheck(g, v, mk())
        assert again.round_polys == first.round_polys and again.final_value == first.final_value
        monkeypatch.setenv("EFFSC_HIP_NO_FUSE", "1")
        plain = sumcheck(EqFactoredProver(wpt, tab), v, mk())
        monkeypatch.delenv("EFFSC_HIP_NO_FUSE")
        assert plain.round_polys == first.round_polys and plain.final_value == first.final_value


def test_eq_factored_short_input_is_zero_padded():
    check_proof("eq", 1, 1, 37, 7, 4242)


@pytest.mark.parametrize("kind", ["ml", "ip", "p3"])
def test_large_parity(kind):
    v = 20
    check_proof(kind, 1, 1, 1 << v, v, 31337)


def test_large_ext3_product_parity():
    check_proof("p3", 3, 1, 1 << 16, 16, 4711)


# --------------------------------------------------------- runner semantics
def test_partial_then_continue_and_per_round_abi():
    # tests/canonical_api.rs:146-178 and 258-290
    a, b = O.uniform(256, 1), O.uniform(256, 2)
    full = sumcheck(InnerProductProver(a, b), 8, Transcript(b"p"))
    g = InnerProductProver(a, b)
    t = Transcript(b"p")
    first = sumcheck(g, 3, t)
    second = sumcheck(g, 5, t)
    assert first.round_polys + second.round_polys == full.round_polys
    assert first.challenges + second.challenges == full.challenges
    assert second.final_value == full.final_value
    # per-round C-ABI calls driven from Python (Python runner) == native runner
    class PyT:
        def __init__(self):
            self.t = Transcript(b"p")

        def send(self, v):
            self.t.send(v)

        def challenge(self):
            return self.t.challenge()

    py = sumcheck(InnerProductProver(a, b), 8, PyT())
    assert py.round_polys == full.round_polys and py.final_value == full.final_value


def test_hook_order_and_zero_rounds():
    # tests/canonical_api.rs:180-197; tests/inner_product_sumcheck.rs:142-157
    calls = []
    g = MultilinearProver(O.uniform(32, 5))
    sumcheck(g, 5, Transcript(b"h"), lambda r, t: calls.append(r))
    assert calls == list(range(5))
    g = InnerProductProver(O.uniform(16, 1), O.uniform(16, 2))
    pr = sumcheck(g, 0, Transcript(b"z"))
    assert pr.round_polys == [] and pr.challenges == []
    assert len(g.evaluations()[0]) == 16


def test_reset_reproduces_proof():
    a, b = O.uniform(1 << 12, 3), O.uniform(1 << 12, 4)
    g = InnerProductProver(a, b)
    p1 = sumcheck(g, 12, Transcript(b"r"))
    g.reset()
    p2 = sumcheck(g, 12, Transcript(b"r"))
    assert p1.round_polys == p2.round_polys and p1.final_value == p2.final_value


def test_montgomery_boundary():
    # arkworks words in, arkworks words out (src/field.rs:99-145)
    a, b = O.uniform(1 << 10, 5), O.uniform(1 << 10, 6)
    am = np.array([to_mont(int(x)) for x in a], dtype=np.uint64)
    bm = np.array([to_mont(int(x)) for x in b], dtype=np.uint64)
    gm = InnerProductProver(am, bm, repr=MONT64)
    gc = InnerProductProver(a, b)
    r0m, r0c = gm.round(None), gc.round(None)
    assert r0m == [to_mont(x) for x in r0c]
    cm = to_mont(123456789)
    assert gm.round(cm) == [to_mont(x) for x in gc.round(123456789)]
    assert [int(x) for x in gm.table_words(0)] == [to_mont(int(x)) for x in gc.table_words(0)]


class _ForwardingTranscript:
    """A transcript the runner does not know as native: forces the Python loop."""

    def __init__(self, inner):
        self.inner = inner
        self.ext = inner.ext

    def send(self, value):
        self.inner.send(value)

    def challenge(self):
        return self.inner.challenge()


@pytest.mark.parametrize("ext", [1, 2, 3])
@pytest.mark.parametrize("python_loop", [False, True])
def test_montgomery_prover_runner(ext, python_loop):
    # sumcheck() over a repr=MONT64 prover: the proof of the canonical prover
    # with every element in Montgomery form, accepted by the verifier
    from effsc.field import from_mont
    v = 9
    a, b = O.uniform((1 << v) * ext, 15), O.uniform((1 << v) * ext, 16)
    am = np.array([to_mont(int(x)) for x in a], dtype=np.uint64)
    bm = np.array([to_mont(int(x)) for x in b], dtype=np.uint64)
    mp = (lambda x: to_mont(x)) if ext == 1 else (lambda x: tuple(to_mont(c) for c in x))
    say(f"mont runner ext={ext} python_loop={python_loop}")
    gc = InnerProductProver(a, b, ext=ext)
    pc = sumcheck(gc, v, Transcript(b"mont", ext=ext))
    gm = InnerProductProver(am, bm, ext=ext, repr=MONT64)
    t = Transcript(b"mont", ext=ext)
    pm = sumcheck(gm, v, _ForwardingTranscript(t) if python_loop else t)
    assert pm.round_polys == [[mp(x) for x in r] for r in pc.round_polys]
    assert pm.challenges == [mp(c) for c in pc.challenges]
    assert pm.final_value == mp(pc.final_value)
    assert [int(x) for x in gm.table_words(0)] == [to_mont(int(x)) for x in gc.table_words(0)]
    # the canonical form of the Montgomery proof verifies against the sum
    # computed from the tables
    F = effsc.Field(ext)
    el = (lambda w, i: int(w[i])) if ext == 1 else (lambda w, i: tuple(int(x) for x in w[i * ext:(i + 1) * ext]))
    total = F.zero()
    for i in range(1 << v):
        total = F.add(total, F.mul(el(a, i), el(b, i)))
    back = effsc.SumcheckProof([[(from_mont(x) if ext == 1 else tuple(from_mont(c) for c in x)) for x in r]
                                for r in pm.round_polys], [], None)
    res = sumcheck_verify(total, 2, v, Transcript(b"mont", ext=ext).verifier(back))
    assert res.challenges == pc.challenges
    assert res.final_claim == pc.final_value


def test_borrowed_device_input_and_fill_uniform():
    n, v = 1 << 14, 14
    ptrs = []
    for seed in (8, 9):
        p = N.u64p()
        N.check(N.lib().effsc_hip_device_alloc(0, n, C.byref(p)))
        N.check(N.lib().effsc_hip_fill_uniform(0, p, n, seed, 0))
        ptrs.append(p)
    try:
        host = np.zeros(n, dtype=np.uint64)
        N.check(N.lib().effsc_hip_copy(0, host.ctypes.data_as(N.u64p), ptrs[0], n))
        assert np.array_equal(host, O.uniform(n, 8))
        addr = [C.cast(p, C.c_void_p).value for p in ptrs]
        g = InnerProductProver(addr[0], addr[1], borrow=True, n=n)
        pg = sumcheck(g, v, Transcript(b"b"))
        po = O.Prover.inner_product(O.uniform(n, 8), O.uniform(n, 9)).sumcheck(v, O.Transcript.shake(b"b"))
        assert tuple(as_oracle(pg, 1)) == tuple(po)
        g.close()
        # input untouched by the proof
        N.check(N.lib().effsc_hip_copy(0, host.ctypes.data_as(N.u64p), ptrs[0], n))
        assert np.array_equal(host, O.uniform(n, 8))
    finally:
        for p in ptrs:
            N.lib().effsc_hip_device_free(0, p)


@pytest.mark.parametrize("ext", [1, 2, 3])
def test_eq_factored_borrowed_input_and_reset(ext):
    v = 11
    n = 1 << v
    aos = O.uniform(n * ext, 610 + ext)
    wpt = O.uniform(v * ext, 620 + ext)
    planes = np.ascontiguousarray(aos.reshape(n, ext).T).reshape(-1)  # one plane per component
    p = N.u64p()
    N.check(N.lib().effsc_hip_device_alloc(0, n * ext, C.byref(p)))
    try:
        N.check(N.lib().effsc_hip_copy(0, p, planes.ctypes.data_as(N.u64p), n * ext))
        say(f"case eq borrowed ext={ext} v={v}")
        g = EqFactoredProver(wpt, C.cast(p, C.c_void_p).value, ext=ext, borrow=True)
        po = O.Prover.eq_factored(wpt, aos, ext=ext).sumcheck(v, O.Transcript.shake(b"eqb"))
        for rep in range(2):
            pg = sumcheck(g, v, Transcript(b"eqb", ext=ext))
            assert tuple(as_oracle(pg, ext)) == tuple(po), f"proof {rep}"
            g.reset()
        # a partial proof, reset, then the full proof again
        sumcheck(g, 3, Transcript(b"eqb", ext=ext))
        g.reset()
        assert tuple(as_oracle(sumcheck(g, v, Transcript(b"eqb", ext=ext)), ext)) == tuple(po)
        g.close()
        back = np.zeros(n * ext, dtype=np.uint64)
        N.check(N.lib().effsc_hip_copy(0, back.ctypes.data_as(N.u64p), p, n * ext))
        assert np.array_equal(back, planes)
    finally:
        N.lib().effsc_hip_device_free(0, p)


def test_eq_factored_borrow_requires_full_length():
    p = N.u64p()
    N.check(N.lib().effsc_hip_device_alloc(0, 64, C.byref(p)))
    try:
        cfg = effsc.provers._config(1, 1, 0, effsc.provers.W_FROM_RAW7, 0, 1)
        h = C.c_void_p()
        wpt = O.uniform(7, 1)
        rc = N.lib().effsc_hip_eq_factored_new(C.byref(cfg), wpt.ctypes.data_as(N.u64p), 7, p, 64, C.byref(h))
        assert rc == 1 and not h.value  # EINVAL: 64 < 2^7
    finally:
        N.lib().effsc_hip_device_free(0, p)


def test_free_fold_matches_reference_semantics():
    for n in (2, 3, 7, 8, 1000, 1 << 15):
        vals = [int(x) for x in O.uniform(n, n)]
        w = 987654321987
        got = effsc.fold(vals, w)
        pr = O.Prover.multilinear(vals)
        pr.finalize([w])
        assert got == [int(x) for x in pr.table(0)]


def test_gpu_proofs_accepted_by_verifier():
    for kind, k in (("ml", 1), ("ip", 2), ("p3", 3)):
        v = 10
        tabs = [O.uniform(1 << v, 60 + i) for i in range(k)]
        claim = 0
        for i in range(1 << v):
            x = 1
            for t in tabs:
                x = x * int(t[i]) % P
            claim = (claim + x) % P
        g = (MultilinearProver(tabs[0]) if k == 1 else InnerProductProver(*tabs) if k == 2
             else CoefficientProver(tabs))
        proof = sumcheck(g, v, Transcript(b"verify"))
        res = sumcheck_verify(claim, k, v, Transcript(b"verify").verifier(proof))
        assert res.final_claim == proof.final_value
        assert res.challenges == proof.challenges


def test_errors_are_reported_not_raised_natively():
    with pytest.raises(ValueError):
        InnerProductProver([1, 2, 3], [1, 2])
    with pytest.raises(N.NativeError):
        CoefficientProver([[1, 2]] * 5)  # more tables than the built-in evaluators take
    with pytest.raises(N.NativeError):
        EqFactoredProver([1, 2], [1, 2, 3, 4, 5])  # more than 2^v evaluations


# ------------------------------------------- device-resident proof loop
def _proof_both_loops(make, rounds, transcript_factory, monkeypatch, tail):
    monkeypatch.setenv("EFFSC_HIP_TAIL", str(tail))
    monkeypatch.delenv("EFFSC_HIP_HOST_LOOP", raising=False)
    dev = sumcheck(make(), rounds, transcript_factory())
    monkeypatch.setenv("EFFSC_HIP_HOST_LOOP", "1")
    host = sumcheck(make(), rounds, transcript_factory())
    monkeypatch.delenv("EFFSC_HIP_HOST_LOOP")
    return dev, host


@pytest.mark.parametrize("tail", [2, 64, 1 << 14, 1 << 22])
@pytest.mark.parametrize("kind,ext,input_ext", [("ml", 1, 1), ("ip", 1, 1), ("ip", 3, 3), ("p3", 1, 1),
                                                ("p3", 3, 1), ("p4", 3, 3), ("p2", 3, 1), ("ml", 2, 1),
                                                ("ip", 2, 2), ("p3", 2, 1), ("p4", 2, 2)])
@pytest.mark.parametrize("v", [1, 2, 5, 11, 16])
def test_device_loop_matches_oracle(kind, ext, input_ext, v, tail, monkeypatch):
    k = {"ml": 1, "ip": 2}.get(kind) or int(kind[1:])
    n = 1 << v
    tabs = [O.uniform(n * input_ext, 7000 + 13 * v + i) for i in range(k)]
    for rounds in sorted({v, max(v - 1, 1), 1}):
        for tr in ("shake", "rng"):
            say(f"device loop kind={kind} ext={ext}/{input_ext} v={v} rounds={rounds} tail={tail} tr={tr}")

            def make():
                return make_pair(kind, ext, input_ext, tabs)[0]

            tf = (lambda: Transcript(b"dev", ext=ext)) if tr == "shake" else (lambda: TestTranscript(99, ext=ext))
            dev, host = _proof_both_loops(make, rounds, tf, monkeypatch, tail)
            assert dev.round_polys == host.round_polys
            assert dev.challenges == host.challenges
            assert dev.final_value == host.final_value
            o = make_pair(kind, ext, input_ext, tabs)[1]
            po = o.sumcheck(rounds, O.Transcript.shake(b"dev") if tr == "shake" else O.Transcript.rng(99))
            assert tuple(as_oracle(dev, ext)) == tuple(po)


@pytest.mark.parametrize("tail", [None, 1 << 12, 1 << 13])
@pytest.mark.parametrize("kind", ["ml", "ip"])
@pytest.mark.parametrize("v", [11, 12, 13, 14, 15, 16, 17])
def test_register_tail_parity(kind, v, tail, monkeypatch):
    # base-field tails of 2^11..2^14 elements run in registers: stop in every
    # phase (all waves, one wave with in-lane pairs, one wave with shuffles)
    if tail is not None:
        monkeypatch.setenv("EFFSC_HIP_TAIL", str(tail))
    n = 1 << v
    for rounds in sorted({1, 2, 3, max(v - 11, 1), v - 10, v - 9, v - 6, v - 5, v - 3, v - 1, v}):
        for tr in ("shake", "rng"):
            check_proof(kind, 1, 1, n, v, 4000 + v, transcript=tr, rounds=rounds)


@pytest.mark.parametrize("tail", [None, 2, 1 << 9])
@pytest.mark.parametrize("kind", ["ml", "ip"])
@pytest.mark.parametrize("v", [9, 12, 17, 20])
def test_fused_small_passes(kind, v, tail, monkeypatch):
    # fold passes of at most 128 workgroups take the previous pass's transcript
    # step into their launch (k_pass_fs2): every workgroup runs the step, the
    # first writes the proof words; the sponge and the partials alternate
    # between two buffers.  Against the oracle, partial runs included, and the
    # same proof with EFFSC_HIP_NO_FUSE (one transcript launch per pass)
    if tail is not None:
        monkeypatch.setenv("EFFSC_HIP_TAIL", str(tail))
    n = 1 << v
    for rounds in sorted({2, 5, v - 1, v}):
        for tr in ("shake", "rng"):
            _, fused = check_proof(kind, 1, 1, n, v, 4800 + v, transcript=tr, rounds=rounds)
            monkeypatch.setenv("EFFSC_HIP_NO_FUSE", "1")
            _, plain = check_proof(kind, 1, 1, n, v, 4800 + v, transcript=tr, rounds=rounds)
            monkeypatch.delenv("EFFSC_HIP_NO_FUSE")
            assert plain.round_polys == fused.round_polys and plain.final_value == fused.final_value


@pytest.mark.parametrize("tail", [None, 1 << 9, 1 << 10])
@pytest.mark.parametrize("kind,input_ext", [("p3", 1), ("p3", 3), ("p4", 3)])
@pytest.mark.parametrize("v", [10, 11, 12])
def test_ext3_register_tail_parity(kind, input_ext, v, tail, monkeypatch):
    # F_p^3 products with tails of 2^8..2^9 work elements run in registers:
    # stop in every phase (all waves, in-lane pairs, shuffles); the plain
    # one-workgroup tail gives the same proof
    if tail is not None:
        monkeypatch.setenv("EFFSC_HIP_TAIL", str(tail))
    n = 1 << v
    for rounds in sorted({1, 2, 3, 4, 5, 6, v - 4, v - 1, v}):
        for tr in ("shake", "rng"):
            g, pg = check_proof(kind, 3, input_ext, n, v, 4600 + v, transcript=tr, rounds=rounds)
    monkeypatch.setenv("EFFSC_HIP_TAIL_PLAIN", "1")
    _, plain = check_proof(kind, 3, input_ext, n, v, 4600 + v, transcript="rng", rounds=v)
    assert plain.round_polys == pg.round_polys and plain.final_value == pg.final_value


@pytest.mark.parametrize("v", [3, 9, 12, 19, 20])
def test_base_input_ext3_product_first_fold(v, monkeypatch):
    # round 1 of a product of three base tables under F_p^3 challenges sums
    # base-field coefficients of c^k (one or two quadruples per thread, from
    # 2^18 quadruples at v = 20); EFFSC_HIP_FSEQ_GENERIC keeps the F_p^3 form
    _, pg = check_proof("p3", 3, 1, 1 << v, v, 5200 + v, transcript="rng")
    _, ps = check_proof("p3", 3, 1, 1 << v, v, 5200 + v, transcript="shake")
    monkeypatch.setenv("EFFSC_HIP_FSEQ_GENERIC", "1")
    _, pgen = check_proof("p3", 3, 1, 1 << v, v, 5200 + v, transcript="rng")
    assert pgen.round_polys == pg.round_polys and pgen.final_value == pg.final_value


@pytest.mark.parametrize("tail", [2, 1 << 14])
def test_device_loop_partial_then_continue(tail, monkeypatch):
    monkeypatch.setenv("EFFSC_HIP_TAIL", str(tail))
    a, b = O.uniform(1 << 16, 1), O.uniform(1 << 16, 2)
    full = sumcheck(InnerProductProver(a, b), 16, Transcript(b"p"))
    g = InnerProductProver(a, b)
    t = Transcript(b"p")
    parts = [sumcheck(g, r, t) for r in (3, 9, 4)]
    assert sum((p.round_polys for p in parts), []) == full.round_polys
    assert sum((p.challenges for p in parts), []) == full.challenges
    assert parts[-1].final_value == full.final_value
    # post-state after partial runs equals the oracle's
    o = O.Prover.inner_product(a, b)
    ot = O.Transcript.shake(b"p")
    for r in (3, 9):
        o.sumcheck(r, ot)
    g2 = InnerProductProver(a, b)
    t2 = Transcript(b"p")
    for r in (3, 9):
        sumcheck(g2, r, t2)
    assert np.array_equal(g2.table_words(0), o.table(0))
    assert np.array_equal(g2.table_words(1), o.table(1))


@pytest.mark.parametrize("kind", ["ml", "ip"])
@pytest.mark.parametrize("v", [2, 3, 6, 9])
def test_per_round_post_state_matches_oracle(kind, v, monkeypatch):
    # per-round C-ABI calls with post-state reads between rounds: the
    # two-round pass keeps folds pending, reads materialise them
    monkeypatch.setenv("EFFSC_HIP_HOST_LOOP", "1")
    n = 1 << v
    tabs = [O.uniform(n, 300 + v + i) for i in range(2)]
    g = MultilinearProver(tabs[0]) if kind == "ml" else InnerProductProver(tabs[0], tabs[1])
    o = O.Prover.multilinear(tabs[0]) if kind == "ml" else O.Prover.inner_product(tabs[0], tabs[1])
    prev = None
    for r in range(v):
        chal = None if prev is None else prev
        got = g.round(chal)
        exp = o.round(None if chal is None else [chal])
        assert [[x] for x in got] == exp, f"round {r}"
        if r % 2 == 1:
            for i in range(1 if kind == "ml" else 2):
                assert np.array_equal(g.table_words(i), o.table(i)), f"round {r} table {i}"
        prev = (r * 0x9E3779B97F4A7C15 + 12345) % P
    g.finalize(prev)
    o.finalize([prev])
    assert [g.final_value()] == o.final_value()


@pytest.mark.parametrize("v", [4, 13, 18])
def test_two_round_pass_equals_one_round_kernels(v, monkeypatch):
    a, b = O.uniform(1 << v, 5), O.uniform(1 << v, 6)
    monkeypatch.setenv("EFFSC_HIP_ONE_ROUND", "1")
    one = sumcheck(InnerProductProver(a, b), v, Transcript(b"2r"))
    one_ml = sumcheck(MultilinearProver(a), v, Transcript(b"2r"))
    monkeypatch.delenv("EFFSC_HIP_ONE_ROUND")
    two = sumcheck(InnerProductProver(a, b), v, Transcript(b"2r"))
    two_ml = sumcheck(MultilinearProver(a), v, Transcript(b"2r"))
    assert one.round_polys == two.round_polys and one.final_value == two.final_value
    assert one_ml.round_polys == two_ml.round_polys and one_ml.final_value == two_ml.final_value


def test_measurement_probes_report_rates_and_reject_bad_arguments():
    L = N.lib()
    r = C.c_double()
    assert L.effsc_hip_mul_rate(0, C.byref(r)) == 0 and r.value > 1e9
    for pat in (0, 1, 2):
        r.value = 0
        assert L.effsc_hip_hbm_rate(0, pat, 1 << 20, C.byref(r)) == 0, L.effsc_hip_last_error()
        assert r.value > 1e9
    assert L.effsc_hip_hbm_rate(0, 3, 1 << 20, C.byref(r)) == 1  # EINVAL: unknown pattern
    assert L.effsc_hip_hbm_rate(0, 0, 3 << 10, C.byref(r)) == 1  # EINVAL: not a power of two
    assert L.effsc_hip_hbm_rate(0, 0, 512, C.byref(r)) == 1  # EINVAL: below 1024


def test_full_size_inner_product_matches_oracle():
    # the benchmark configuration itself: two 2^28-element tables from the
    # counter-based generator, proved on the GPU (two-round passes, fused
    # small passes, pipelined tail, device transcript) and by the oracle's
    # threaded restatement of the
    # reference prover (two folds per round, serial evaluate): bit-identical
    # round polynomials, challenges and final value
    v = 28
    n = 1 << v
    say(f"full-size inner product v={v}")
    a, b = O.uniform(n, 0x5EED0002), O.uniform(n, 0x5EED0003)
    proof = sumcheck(InnerProductProver(a, b), v, Transcript(b"effsc-hip/bench/ip"))
    _, polys, chals, fin = O.bench_ip(a, b, v, 16)  # consumes a and b
    assert [w for r in proof.round_polys for w in r] == [int(x) for x in polys]
    assert proof.challenges == [int(x) for x in chals]
    assert proof.final_value == fin


@pytest.mark.parametrize("kind,ext,input_ext", [("p3", 3, 1), ("eq", 1, 1), ("eq", 3, 3)])
def test_full_size_degree3_and_eq_configurations(kind, ext, input_ext):
    # the v = 24 configurations at full size against the oracle: product of
    # three base tables under F_p^3 challenges, EqFactored over the base
    # field and over F_p^3 (proof words and post-state)
    check_proof(kind, ext, input_ext, 1 << 24, 24, 0x5EED0004)
