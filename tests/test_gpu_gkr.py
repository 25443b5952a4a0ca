"""GkrProver on the GPU against the oracle (src/provers/gkr.rs semantics,
non-power-of-two witness lengths included)."""
import sys

import numpy as np
import pytest

import oracle as O
from effsc import GkrProver, TestTranscript, Transcript, sumcheck, sumcheck_verify

pytestmark = pytest.mark.gpu
P = O.P


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def lift(words, ext):
    if ext == 1:
        return words
    out = np.zeros(len(words) * ext, dtype=np.uint64)
    out[::ext] = words
    return out


def as_oracle(proof, ext):
    rp = [[[x] if ext == 1 else list(x) for x in r] for r in proof.round_polys]
    ch = [[c] if ext == 1 else list(c) for c in proof.challenges]
    fin = [proof.final_value] if ext == 1 else list(proof.final_value)
    return rp, ch, fin


def rounds_for(n):
    return (n * n - 1).bit_length() if n > 1 else 0


@pytest.mark.parametrize("ext,input_ext", [(1, 1), (3, 3), (3, 1), (2, 2), (2, 1)])
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 16, 37, 64])
def test_gkr_parity(n, ext, input_ext):
    seed = 1000 * n + 10 * ext + input_ext
    add = O.uniform(n * n * input_ext, seed)
    mult = O.uniform(n * n * input_ext, seed + 1)
    wit = O.uniform(n * input_ext, seed + 2)
    v = rounds_for(n)
    say(f"gkr n={n} ext={ext} input_ext={input_ext} rounds={v}")
    g = GkrProver(add, mult, wit, ext=ext, input_ext=input_ext)
    o = O.Prover.gkr(lift(add, ext // input_ext), lift(mult, ext // input_ext), lift(wit, ext // input_ext), ext=ext)
    pg = sumcheck(g, v, Transcript(b"gkr-test", ext=ext))
    po = o.sumcheck(v, O.Transcript.shake(b"gkr-test"))
    assert tuple(as_oracle(pg, ext)) == tuple(po)
    for i in range(4):
        assert np.array_equal(g.table_words(i), o.table(i)), f"table {i}"
    wb, wc = g.claimed_w_values()
    cw = o.claimed_w_values()
    if cw is not None:
        el = (lambda a: int(a[0])) if ext == 1 else (lambda a: tuple(int(x) for x in a))
        assert (wb, wc) == (el(cw[0]), el(cw[1]))


def test_gkr_large_and_test_transcript():
    n = 512  # 2^18 gate entries
    add, mult, wit = O.uniform(n * n, 5), O.uniform(n * n, 6), O.uniform(n, 7)
    v = 18
    pg = sumcheck(GkrProver(add, mult, wit), v, TestTranscript(3))
    po = O.Prover.gkr(add, mult, wit).sumcheck(v, O.Transcript.rng(3))
    assert tuple(as_oracle(pg, 1)) == tuple(po)


def test_gkr_verifier_accepts_and_claims_match_witness_mle():
    # src/provers/gkr.rs:208-266 on the GPU: claim chain and W(b*), W(c*)
    k = 5
    n = 1 << k
    add, mult, wit = O.uniform(n * n, 11), O.uniform(n * n, 12), O.uniform(n, 13)
    claim = 0
    for b in range(n):
        for c in range(n):
            i = b * n + c
            wb, wc = int(wit[b]), int(wit[c])
            claim = (claim + int(add[i]) * (wb + wc) + int(mult[i]) * wb * wc) % P
    g = GkrProver(add, mult, wit)
    proof = sumcheck(g, 2 * k, Transcript(b"gkr-verify"))
    res = sumcheck_verify(claim, 2, 2 * k, Transcript(b"gkr-verify").verifier(proof))
    assert res.final_claim == proof.final_value

    def mle(table, pt):
        t = [int(x) for x in table]
        for r in pt:
            h = len(t) // 2
            t = [(t[j] + r * (t[j + h] - t[j])) % P for j in range(h)]
        return t[0]
    wb, wc = g.claimed_w_values()
    assert wb == mle(wit, proof.challenges[:k]) and wc == mle(wit, proof.challenges[k:])


def test_gkr_length_errors():
    with pytest.raises(ValueError):
        GkrProver(O.uniform(15, 1), O.uniform(16, 2), O.uniform(4, 3))
    with pytest.raises(ValueError):
        GkrProver(O.uniform(16, 1), O.uniform(9, 2), O.uniform(4, 3))


@pytest.mark.parametrize("ext,input_ext", [(1, 1), (3, 3), (3, 1), (2, 2), (2, 1)])
@pytest.mark.parametrize("k", [6, 7, 8, 9])
def test_gkr_device_loop(k, ext, input_ext, monkeypatch):
    # hookless proofs of power-of-two layers run on the device (factored
    # witness rounds while the b variables are bound, four-table round
    # launches, then one workgroup); partial runs end on large tables, inside
    # the factored rounds (2, 3), at their end (k) and after it (k + 1)
    n = 1 << k
    seed = 50 * k + ext + input_ext
    add, mult, wit = (O.uniform(n * n * input_ext, seed), O.uniform(n * n * input_ext, seed + 1),
                      O.uniform(n * input_ext, seed + 2))
    lifted = [lift(x, ext // input_ext) for x in (add, mult, wit)]
    for rounds in sorted({1, 2, 3, k, k + 1, 2 * k - 1, 2 * k}):
        for tr in ("shake", "rng"):
            say(f"gkr device loop k={k} ext={ext}/{input_ext} rounds={rounds} tr={tr}")
            mk = (lambda: Transcript(b"gkr-dev", ext=ext)) if tr == "shake" else (lambda: TestTranscript(21, ext=ext))
            g = GkrProver(add, mult, wit, ext=ext, input_ext=input_ext)
            pg = sumcheck(g, rounds, mk())
            o = O.Prover.gkr(*lifted, ext=ext)
            po = o.sumcheck(rounds, O.Transcript.shake(b"gkr-dev") if tr == "shake" else O.Transcript.rng(21))
            assert tuple(as_oracle(pg, ext)) == tuple(po)
            for i in range(4):
                assert np.array_equal(g.table_words(i), o.table(i)), f"table {i}"
            # the per-round path gives the same proof
            monkeypatch.setenv("EFFSC_HIP_HOST_LOOP", "1")
            h = GkrProver(add, mult, wit, ext=ext, input_ext=input_ext)
            ph = sumcheck(h, rounds, mk())
            monkeypatch.delenv("EFFSC_HIP_HOST_LOOP")
            assert ph.round_polys == pg.round_polys and ph.final_value == pg.final_value
            if rounds == 2 * k:
                assert g.claimed_w_values() == h.claimed_w_values()
            if ext == 1:
                # one transcript launch per round instead of the small
                # factored launches taking the previous round's step
                monkeypatch.setenv("EFFSC_HIP_NO_FUSE", "1")
                pl = GkrProver(add, mult, wit, ext=ext, input_ext=input_ext)
                pp = sumcheck(pl, rounds, mk())
                monkeypatch.delenv("EFFSC_HIP_NO_FUSE")
                assert pp.round_polys == pg.round_polys and pp.final_value == pg.final_value
                for i in range(4):
                    assert np.array_equal(pl.table_words(i), g.table_words(i)), f"table {i}"
            # and so does the device loop over the four broadcast tables
            monkeypatch.setenv("EFFSC_HIP_GKR_BROADCAST", "1")
            f = GkrProver(add, mult, wit, ext=ext, input_ext=input_ext)
            pf = sumcheck(f, rounds, mk())
            monkeypatch.delenv("EFFSC_HIP_GKR_BROADCAST")
            assert pf.round_polys == pg.round_polys and pf.final_value == pg.final_value
            for i in range(4):
                assert np.array_equal(f.table_words(i), g.table_words(i)), f"table {i}"


def test_gkr_full_size_layer_matches_oracle():
    # the benchmark layer: k = 12, 2^24 gate entries, all 24 rounds, against
    # the oracle (proof, post-state tables, claimed W values)
    k = 12
    n = 1 << k
    add, mult, wit = O.uniform(n * n, 0x5EED0041), O.uniform(n * n, 0x5EED0042), O.uniform(n, 0x5EED0043)
    say(f"gkr full-size layer k={k}")
    g = GkrProver(add, mult, wit)
    pg = sumcheck(g, 2 * k, Transcript(b"gkr-full"))
    o = O.Prover.gkr(add, mult, wit)
    po = o.sumcheck(2 * k, O.Transcript.shake(b"gkr-full"))
    assert tuple(as_oracle(pg, 1)) == tuple(po)
    for i in range(4):
        assert np.array_equal(g.table_words(i), o.table(i)), f"table {i}"
    cw = o.claimed_w_values()
    assert tuple(g.claimed_w_values()) == (int(cw[0][0]), int(cw[1][0]))
