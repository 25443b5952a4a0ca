"""Pins the C oracle: golden vectors from Python big integers and hashlib,
the reference's constants and known-answer tests, and the brute force."""
import hashlib
import random

import numpy as np
import pytest

import brute as B
import oracle as O

P = O.P


def test_field_constants_match_reference():
    # avx512.rs:19-35 and tests/fields.rs:61-64
    assert O.lib().or_to_mont(1) == 0xFFFFFFFF
    assert O.lib().or_to_mont(P - 1) == 0xFFFFFFFE00000002
    assert O.lib().or_from_mont(0xFFFFFFFF) == 1
    # from_raw(7) has canonical value 7 * 2^-64 = -7 * 2^32 mod p
    assert O.lib().or_from_mont(7) == O.W_FROM_RAW7 == (P - 7 * 2**32) % P
    # INV = -p^-1 mod 2^64
    assert (P * 0xFFFFFFFEFFFFFFFF) % 2**64 == 2**64 - 1


def test_mont_mul_edge_cases():
    # avx512.rs:689-701: 0*(-1) = 0, 1*(-1) = -1, (-1)(-1) = 1 in Montgomery form
    one, neg = O.lib().or_to_mont(1), O.lib().or_to_mont(P - 1)
    assert O.lib().or_mont_mul(0, neg) == 0
    assert O.lib().or_mont_mul(one, neg) == neg
    assert O.lib().or_mont_mul(neg, neg) == one


def test_field_golden(golden):
    L = O.lib()
    for a, b, c in golden["field"]["mul"]:
        assert L.or_gl_mul(a, b) == c
        assert L.or_gl_mul_i128(a, b) == c
    for a, b, c in golden["field"]["add"]:
        assert L.or_gl_add(a, b) == c
    for a, b, c in golden["field"]["sub"]:
        assert L.or_gl_sub(a, b) == c
    for a, c in golden["field"]["inv"]:
        assert L.or_gl_inv(a) == c


def test_field_random_vs_bigint():
    rng = random.Random(1)
    L = O.lib()
    for _ in range(20000):
        a, b = rng.randrange(P), rng.randrange(P)
        assert L.or_gl_mul(a, b) == a * b % P
        m = L.or_mont_mul(a, b)
        assert m == a * b * pow(2**64, P - 2, P) % P


def test_ext_golden(golden):
    for key, cases in golden["field"]["ext"].items():
        d = int(key[1])
        w = int(key.split("_w")[1], 16)
        for a, b, c in cases:
            assert O.ext_mul(d, w, a, b) == c


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("w", [7, O.W_FROM_RAW7])
def test_ext_inverse_and_nonresidue(d, w):
    rng = random.Random(d * 31 + w % 97)
    one = [1] + [0] * (d - 1)
    for _ in range(50):
        a = [rng.randrange(P) for _ in range(d)]
        assert O.ext_mul(d, w, a, O.ext_inv(d, w, a)) == one
    # X^d = w generates a field: w is not a d-th power (Euler criterion)
    assert pow(w, (P - 1) // d, P) != 1


def test_uniform_generator_matches_python():
    assert [int(x) for x in O.uniform(64, 5)] == B.uniform_table(64, 5)
    assert all(int(x) < P for x in O.uniform(1000, 9))


def test_shake256_golden(golden):
    for msg, out in golden["shake256"]:
        assert O.shake256(bytes.fromhex(msg), 200).hex() == out
    assert O.shake256(b"x" * 1000, 400) == hashlib.shake_256(b"x" * 1000).digest(400)


def test_keccak_permutation_of_zero_state():
    # FIPS 202 / KeccakCodePackage: Keccak-f[1600] applied to the all-zero state
    st = (O.C.c_uint64 * 25)() if hasattr(O, "C") else None
    import ctypes
    st = (ctypes.c_uint64 * 25)()
    O.lib().or_keccak_f1600(st)
    assert st[0] == 0xF1258F7940E1DDE7 and st[1] == 0x84D5CCF933C0478A


def test_transcript_golden(golden):
    for case in golden["transcript"]:
        if "rng_seed" in case:
            t = O.Transcript.rng(case["rng_seed"])
            assert [t.challenge(1) for _ in case["challenges"]] == case["challenges"]
            continue
        t = O.Transcript.shake(bytes.fromhex(case["domsep"]))
        for op, val in case["ops"]:
            if op == "send":
                t.send(val)
            else:
                assert t.challenge(case["ext"]) == val


def _oracle_proof(case):
    ext, n, v = case["ext"], case["n"], case["v"]
    seeds = case["seeds"]
    kind = case["kind"]
    tabs = [O.uniform(n, s) for s in seeds]
    if kind == "ml":
        pr = O.Prover.multilinear(tabs[0] if ext == 1 else _lift(tabs[0], ext), ext=ext)
    elif kind == "ip":
        pr = O.Prover.inner_product(*[t if ext == 1 else _lift(t, ext) for t in tabs[:2]], ext=ext)
    elif kind.startswith("product"):
        k = int(kind[-1])
        pr = O.Prover.product(tabs[:k], ext=ext)
    else:
        wpt = O.uniform(v * ext, seeds[1])
        pr = O.Prover.eq_factored(wpt, tabs[0], ext=ext)
    t = (O.Transcript.shake(case["domsep"].encode()) if case["transcript"] == "shake" else O.Transcript.rng(7))
    return pr.sumcheck(v, t)


def test_oracle_proofs_match_brute_force_golden(golden):
    for case in golden["proofs"]:
        rp, ch, fin = _oracle_proof(case)
        assert rp == case["round_polys"], case["kind"]
        assert ch == case["challenges"]
        assert fin == case["final"]


@pytest.mark.parametrize("v", [1, 3, 6])
@pytest.mark.parametrize("ext", [1, 2, 3])
def test_eq_factored_equals_inner_product_with_eq_table(v, ext):
    # src/provers/eq_factored.rs:411-446: EqFactored == IP(p, build_eq_table(w))
    n = 1 << v
    p = O.uniform(n * ext, 3)
    wpt = O.uniform(v * ext, 4)
    F = B.Ext(ext)
    W = [tuple(int(x) for x in wpt[i * ext:(i + 1) * ext]) for i in range(v)]
    eq = []
    for x in range(n):
        eq.extend(B.eq_point(F, W, B._bits(x, v, F)))
    a = O.Prover.eq_factored(wpt, p, ext=ext).sumcheck(v, O.Transcript.shake(b"eq"))
    b = O.Prover.inner_product(p, np.array(eq, dtype=np.uint64), ext=ext).sumcheck(v, O.Transcript.shake(b"eq"))
    assert a == b


def test_eq_table_known_answer():
    # src/hypercube/eq_evals.rs:70-79: LSB eq([2,3]) = [2, -4, -3, 6]; the MSB
    # table of build_eq_table swaps the middle entries.
    pr = O.Prover.eq_factored([2, 3], [0, 0, 0, 0])
    # v = 2 -> v_l = 1: eq_l = eq([2]) = [-1, 2], eq_r = eq([3]) = [-2, 3]
    assert [int(x) for x in pr.table(1)] == [P - 1, 2]
    assert [int(x) for x in pr.table(2)] == [P - 2, 3]
    full = [(l * r) % P for l in (P - 1, 2) for r in (P - 2, 3)]
    assert full == [2, P - 3, P - 4, 6]


def test_coefficient_degree1_equals_multilinear():
    # src/provers/coefficient.rs:378-414
    t = O.uniform(64, 8)
    a = O.Prover.product([t]).sumcheck(6, O.Transcript.shake(b"c"))
    b = O.Prover.multilinear(t).sumcheck(6, O.Transcript.shake(b"c"))
    assert a == b


def test_round0_closed_forms():
    # tests/multilinear_sumcheck.rs:140-157 and tests/inner_product_sumcheck.rs:159-186
    n = 32
    a, b = O.uniform(n, 1), O.uniform(n, 2)
    A, Bv = [int(x) for x in a], [int(x) for x in b]
    (s0,), = O.Prover.multilinear(a).round(None)
    assert s0 == sum(A[:n // 2]) % P
    q = O.Prover.inner_product(a, b).round(None)
    assert q[0] == [sum(x * y for x, y in zip(A[:16], Bv[:16])) % P]
    assert q[1] == [sum((A[16 + i] - A[i]) * (Bv[16 + i] - Bv[i]) for i in range(16)) % P]


def test_verifier_accepts_honest_and_catches_corruption():
    # tests/adversarial_verifier.rs:110-230, 492-530
    for kind, k in (("ml", 1), ("ip", 2), ("p3", 3)):
        v, n = 6, 64
        tabs = [O.uniform(n, 40 + i) for i in range(k)]
        claimed = [0]
        for i in range(n):
            x = 1
            for t in tabs:
                x = x * int(t[i]) % P
            claimed[0] = (claimed[0] + x) % P
        pr = O.Prover.multilinear(tabs[0]) if k == 1 else (
            O.Prover.inner_product(*tabs) if k == 2 else O.Prover.product(tabs))
        rp, ch, fin = pr.sumcheck(v, O.Transcript.shake(b"adv"))
        words = [w for r in rp for e in r for w in e]
        st, vch, vfin = O.verify(claimed, k, v, O.Transcript.shake(b"adv"), words)
        assert st == 0 and vch == ch and vfin == fin
        bad = list(words)
        bad[len(bad) // 2] = (bad[len(bad) // 2] + 1) % P
        st, _, vfin = O.verify(claimed, k, v, O.Transcript.shake(b"adv"), bad)
        assert st == 0 and vfin != fin  # caller's oracle check catches it
        st, _, vfin = O.verify([(claimed[0] + 1) % P], k, v, O.Transcript.shake(b"adv"), words)
        assert vfin != fin
        st, _, _ = O.verify(claimed, k, v, O.Transcript.shake(b"adv"), words[:-1])
        assert st == v  # TranscriptError in the last round


def test_final_value_is_mle_at_challenges():
    # tests/canonical_api.rs:22-34,139
    v = 5
    a, b = O.uniform(32, 5), O.uniform(32, 6)
    F = B.Ext(1)
    pr = O.Prover.inner_product(a, b)
    rp, ch, fin = pr.sumcheck(v, O.Transcript.rng(3))
    pt = [tuple(c) for c in ch]
    fa = B.mle(F, [(int(x),) for x in a], pt)
    fb = B.mle(F, [(int(x),) for x in b], pt)
    assert fin == [fa[0] * fb[0] % P]
    assert [int(pr.table(0)[0])] == list(fa)


def test_partial_then_continue_equals_full():
    # tests/canonical_api.rs:146-178
    a, b = O.uniform(256, 1), O.uniform(256, 2)
    full = O.Prover.inner_product(a, b).sumcheck(8, O.Transcript.shake(b"p"))
    pr = O.Prover.inner_product(a, b)
    t = O.Transcript.shake(b"p")
    first = pr.sumcheck(3, t)
    second = pr.sumcheck(5, t)
    assert first[0] + second[0] == full[0]
    assert first[1] + second[1] == full[1]
    assert second[2] == full[2]


def test_plonky3_style_counter_closed_form():
    # tests/plonky3_roundtrip.rs:158-181, MSB analogue: evals x+1, challenges
    # 2,3,4,5 give f(r) = 1 + 8*2 + 4*3 + 2*4 + 5 = 42
    evals = list(range(1, 17))
    pr = O.Prover.multilinear(evals)
    out = [pr.round(None)]
    for c in (2, 3, 4):
        out.append(pr.round([c]))
    pr.finalize([5])
    assert out[0] == [[36]]
    assert pr.final_value() == [42]


# ---------------------------------------------------------------- GKR
@pytest.mark.parametrize("ext", [1, 2, 3])
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_gkr_oracle_matches_brute_force(ext, n):
    # src/provers/gkr.rs restated in C against the definition (tails included)
    add = O.uniform(n * n * ext, 70 + n)
    mult = O.uniform(n * n * ext, 80 + n)
    wit = O.uniform(n * ext, 90 + n)
    F = B.Ext(ext)

    def els(words):
        return [tuple(int(x) for x in words[i * ext:(i + 1) * ext]) for i in range(len(words) // ext)]
    v = (n * n - 1).bit_length() if n > 1 else 0
    rp, ch, fin = O.Prover.gkr(add, mult, wit, ext=ext).sumcheck(v, O.Transcript.shake(b"gkr"))
    brp, bch, bfin = B.prove_gkr(F, els(add), els(mult), els(wit), B.ShakeTranscript(b"gkr"), lift=False)
    assert [[list(x) for x in r] for r in brp] == rp
    assert [list(c) for c in bch] == ch
    assert list(bfin) == fin


def test_gkr_reference_identities():
    # src/provers/gkr.rs:208-266 (claim chain, claimed W values = MLE of W at
    # the b and c halves of the challenges) and :291-316 (zero gates)
    for k in (1, 2, 3, 4):
        n = 1 << k
        add, mult, wit = O.uniform(n * n, 0x100 * k), O.uniform(n * n, 0x100 * k + 1), O.uniform(n, 0x100 * k + 2)
        claim = 0
        for b in range(n):
            for c in range(n):
                i = b * n + c
                wb, wc = int(wit[b]), int(wit[c])
                claim = (claim + int(add[i]) * (wb + wc) + int(mult[i]) * wb * wc) % P
        p = O.Prover.gkr(add, mult, wit)
        rp, ch, fin = p.sumcheck(2 * k, O.Transcript.rng(99))
        for (q0, qi), (r,) in zip(rp, ch):
            q0, qi = q0[0], qi[0]
            q1 = (claim - q0) % P
            claim = (q0 + r * (q1 - q0 - qi) + qi * r * r) % P
        assert claim == fin[0]
        wb, wc = p.claimed_w_values()
        F = B.Ext(1)
        T = [(int(x),) for x in wit]
        assert int(wb[0]) == B.mle(F, T, [(c[0],) for c in ch[:k]])[0]
        assert int(wc[0]) == B.mle(F, T, [(c[0],) for c in ch[k:]])[0]
    z = np.zeros(16, dtype=np.uint64)
    rp, _, _ = O.Prover.gkr(z, z, O.uniform(4, 0x500)).sumcheck(4, O.Transcript.rng(99))
    assert all(x == [0] for r in rp for x in r)


# ---------------------------------------------------------------- LSB provers
@pytest.mark.parametrize("kind,k", [("ml", 1), ("ip", 2), ("p3", 3)])
@pytest.mark.parametrize("ext", [1, 2, 3])
@pytest.mark.parametrize("v", [0, 1, 2, 3, 5])
def test_lsb_oracle_matches_brute_force(kind, k, ext, v):
    n = 1 << v
    tabs = [O.uniform(n, 300 + 10 * k + v + i) for i in range(k)]
    if kind == "ml":
        o = O.Prover.multilinear_lsb(tabs[0] if ext == 1 else _lift(tabs[0], ext), ext=ext)
    elif kind == "ip":
        o = O.Prover.inner_product_lsb(*[t if ext == 1 else _lift(t, ext) for t in tabs], ext=ext)
    else:
        o = O.Prover.product_lsb(tabs, ext=ext)
    rp, ch, fin = o.sumcheck(v, O.Transcript.shake(b"lsb"))
    F = B.Ext(ext)
    brp, bch, bfin = B.prove_lsb(F, [[int(x) for x in t] for t in tabs], v, B.ShakeTranscript(b"lsb"), k)
    assert [[list(x) for x in r] for r in brp] == rp
    assert [list(c) for c in bch] == ch
    assert list(bfin) == fin


def _lift(words, ext):
    out = np.zeros(len(words) * ext, dtype=np.uint64)
    out[::ext] = words
    return out


def test_lsb_equals_msb_on_bit_reversed_table():
    # src/streams/memory/core.rs:20-47: reorder_vec_msb(out[i] = src[rev(i)])
    # turns the LSB statement into the MSB one, proof for proof
    for v in (1, 4, 7):
        n = 1 << v
        a, b = O.uniform(n, 7 + v), O.uniform(n, 8 + v)
        rev = np.array([int(format(i, f"0{v}b")[::-1], 2) for i in range(n)])
        lsb = O.Prover.inner_product_lsb(a, b).sumcheck(v, O.Transcript.shake(b"r"))
        msb = O.Prover.inner_product(a[rev], b[rev]).sumcheck(v, O.Transcript.shake(b"r"))
        assert lsb == msb


def test_base_coefficient_form_of_first_fold():
    # Round 1 of a product of three base tables under an F_p^3 challenge c:
    # the GPU sums base-field coefficients e_k of c^k (k_fused_base3).  The
    # same formulas, restated here, equal the F_p^3 evaluation from the
    # definition: fold each pair with c, then h(0), h(inf), h(2) of the
    # product of the three folded tables, summed over the quadruples.
    F = B.Ext(3)
    rng = random.Random(77)
    for trial in range(20):
        c = tuple(rng.randrange(P) for _ in range(3))
        q = rng.randrange(1, 9)
        quads = [[[rng.randrange(P) for _ in range(4)] for _ in range(3)] for _ in range(q)]
        if trial == 0:
            quads[0] = [[0, P - 1, 1, P - 1] for _ in range(3)]  # edge values
        want = [F.zero()] * 3
        e = [[0] * 4 for _ in range(3)]
        for quad in quads:
            prods = [F.one()] * 3
            ab = []
            for x0, x1, x2, x3 in quad:
                lo = F.add(F.lift(x0), F.scale(c, (x2 - x0) % P))
                hi = F.add(F.lift(x1), F.scale(c, (x3 - x1) % P))
                d = F.sub(hi, lo)
                for t, h in enumerate((lo, d, F.add(hi, d))):
                    prods[t] = F.mul(prods[t], h)
                dl, dh = (x2 - x0) % P, (x3 - x1) % P
                ai, bi = (x1 - x0) % P, (dh - dl) % P
                ab.append(((x0, dl), (ai, bi), ((x1 + ai) % P, (dh + bi) % P)))
            want = [F.add(want[t], prods[t]) for t in range(3)]
            for t in range(3):
                (a0, b0), (a1, b1), (a2, b2) = (ab[i][t] for i in range(3))
                a12, b12, m12 = a1 * a2 % P, b1 * b2 % P, (a1 * b2 + b1 * a2) % P
                for k, term in enumerate((a0 * a12, b0 * a12 + a0 * m12, b0 * m12 + a0 * b12, b0 * b12)):
                    e[t][k] = (e[t][k] + term) % P
        c2 = F.mul(c, c)
        c3 = F.mul(c2, c)
        got = [F.add(F.add(F.lift(e[t][0]), F.scale(c, e[t][1])), F.add(F.scale(c2, e[t][2]), F.scale(c3, e[t][3])))
               for t in range(3)]
        assert got == want


# ------------------------------------------------------- threaded CPU baselines
@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("threads", [1, 8])
@pytest.mark.parametrize("v", [1, 2, 5, 11, 16])
def test_bench_ip_equals_prover_and_reference_order(v, threads, fused):
    # the timed CPU baselines (InnerProductProver round structure, and the
    # legacy fused path of src/inner_product_sumcheck.rs:199-360) must produce
    # the oracle prover's proof: they check the full-size GPU proofs
    n = 1 << v
    a, b = O.uniform(n, 0x77 + v), O.uniform(n, 0x78 + v)
    rp, ch, fin = O.Prover.inner_product(a, b).sumcheck(v, O.Transcript.shake(b"effsc-hip/bench/ip"))
    _, polys, chals, f = O.bench_ip(a.copy(), b.copy(), v, threads, fused=fused)
    assert [int(x) for x in polys[:2 * v]] == [x[0] for r in rp for x in r]
    assert [int(x) for x in chals[:v]] == [c[0] for c in ch]
    assert [f] == fin


@pytest.mark.parametrize("fused", [False, True])
def test_bench_ip_matches_brute_force_golden(golden, fused):
    for case in golden["proofs"]:
        if case["kind"] != "ip" or case["ext"] != 1 or case["n"] != 1 << case["v"] or case["transcript"] != "shake":
            continue
        v = case["v"]
        a, b = O.uniform(case["n"], case["seeds"][0]), O.uniform(case["n"], case["seeds"][1])
        _, polys, chals, f = O.bench_ip(a, b, v, 4, fused=fused, domsep=case["domsep"].encode())
        assert [[[int(polys[2 * r])], [int(polys[2 * r + 1])]] for r in range(v)] == case["round_polys"]
        assert [[int(c)] for c in chals[:v]] == case["challenges"]
        assert [f] == case["final"]


@pytest.mark.parametrize("threads", [1, 8])
@pytest.mark.parametrize("v", [1, 2, 5, 12, 16])
def test_bench_ml_equals_prover(v, threads):
    # the timed CPU baseline of the multilinear workloads (the reference's
    # parallel fused kernels, src/multilinear_sumcheck.rs:84-254) must produce
    # the oracle prover's proof
    n = 1 << v
    x = O.uniform(n, 0x91 + v)
    rp, ch, fin = O.Prover.multilinear(x).sumcheck(v, O.Transcript.shake(b"effsc-hip/bench/ml"))
    _, polys, chals, f = O.bench_ml(x.copy(), v, threads)
    assert [int(p) for p in polys[:v]] == [r[0][0] for r in rp]
    assert [int(c) for c in chals[:v]] == [c[0] for c in ch]
    assert [f] == fin


def test_bench_ml_matches_brute_force_golden(golden):
    seen = 0
    for case in golden["proofs"]:
        if case["kind"] != "ml" or case["ext"] != 1 or case["n"] != 1 << case["v"] or case["transcript"] != "shake":
            continue
        v = case["v"]
        _, polys, chals, f = O.bench_ml(O.uniform(case["n"], case["seeds"][0]), v, 4, domsep=case["domsep"].encode())
        assert [[[int(polys[r])]] for r in range(v)] == case["round_polys"]
        assert [[int(c)] for c in chals[:v]] == case["challenges"]
        assert [f] == case["final"]
        seen += 1
    assert seen == 4


# ------------------------------------------------ general coefficient shapes
def _pylist(words, ext):
    return [int(x) for x in words] if ext == 1 else [tuple(int(x) for x in words[i * ext:(i + 1) * ext])
                                                    for i in range(len(words) // ext)]


@pytest.mark.parametrize("ext", [1, 2, 3])
@pytest.mark.parametrize("k", [1, 2, 3, 4])
@pytest.mark.parametrize("n", [3, 5, 6, 7, 12, 13])
def test_product_non_pow2_matches_definition(n, k, ext):
    # msb_fold_vec halves with len / 2 and drops an odd last element
    # (src/provers/coefficient.rs:131-140)
    rounds = max(1, (n - 1).bit_length() - 1) if n > 3 else 1
    tabs = [O.uniform(n, 400 + 10 * k + n + i) for i in range(k)]
    F = B.Ext(ext)
    rp, ch, fin = O.Prover.product(tabs, ext=ext).sumcheck(rounds, O.Transcript.shake(b"halve"))
    brp, bch, bfin = B.prove_halving(F, [[int(x) for x in t] for t in tabs], rounds, B.ShakeTranscript(b"halve"))
    assert [[list(x) for x in r] for r in brp] == rp
    assert [list(c) for c in bch] == ch
    assert list(bfin) == fin


@pytest.mark.parametrize("ext", [1, 3])
@pytest.mark.parametrize("n", [1, 3, 5, 6, 7, 11, 12, 100])
def test_lsb_multilinear_any_length_matches_definition(n, ext):
    # src/provers/multilinear_lsb.rs: s0 counts an unpaired last element, the
    # fold drops it
    rounds = max(1, (n - 1).bit_length())
    t = O.uniform(n, 500 + n)
    F = B.Ext(ext)
    o = O.Prover.multilinear_lsb(t if ext == 1 else _lift(t, ext), ext=ext)
    rp, ch, fin = o.sumcheck(rounds, O.Transcript.shake(b"lsb-odd"))
    brp, bch, bfin, bT = B.prove_ml_pairs(F, [int(x) for x in t], rounds, B.ShakeTranscript(b"lsb-odd"))
    assert [[list(x) for x in r] for r in brp] == rp
    assert [list(c) for c in bch] == ch
    assert list(bfin) == fin
    assert _pylist(o.table(0), ext) == [x[0] if ext == 1 else x for x in bT]


@pytest.mark.parametrize("ext", [2, 3])
@pytest.mark.parametrize("v", [1, 3, 4])
def test_product_lsb_extension_inputs_match_brute_force(v, ext):
    n = 1 << v
    tabs = [O.uniform(n * ext, 800 + v + i) for i in range(3)]
    o = O.Prover.product_lsb(tabs, ext=ext, tables_in_ext=True)
    rp, ch, fin = o.sumcheck(v, O.Transcript.shake(b"lsb-ext"))
    F = B.Ext(ext)
    el = [[tuple(int(x) for x in t[i * ext:(i + 1) * ext]) for i in range(n)] for t in tabs]
    brp, bch, bfin = B.prove_lsb(F, el, v, B.ShakeTranscript(b"lsb-ext"), 3, lift=False)
    assert [[list(x) for x in r] for r in brp] == rp
    assert [list(c) for c in bch] == ch
    assert list(bfin) == fin
