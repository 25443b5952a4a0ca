"""Edge operands through the device arithmetic.

The golden field vectors (0, 1, p - 1, p - 2, 2^32 - 1, 2^32, 2^32 + 1, 2^63,
w, ...) and the F_p^2 / F_p^3 product vectors of tests/golden/golden.json
are driven through the gfx950 kernels: a two-element table [0, b] folded
with challenge a is a + a (b - 0) - 0 ... = a b, computed by the fold kernel's
field product, and read back through the post-state accessor.  Tables and
challenges drawn only from those edge values then run through every prover
(base field, F_p^2, F_p^3; sizes crossing the two-round passes, the
one-quadruple and two-quadruple product kernels and the register tails),
against the oracle, both in the device loop and round by round with edge
challenges.
"""
import sys

import numpy as np
import pytest

import oracle as O
from effsc import (GkrProver, InnerProductProverLSB, MultilinearProver, MultilinearProverLSB, Transcript,
                   TestTranscript, sumcheck)
from effsc.field import W_FROM_RAW7

from test_gpu_parity import as_oracle, make_pair

pytestmark = pytest.mark.gpu
P = O.P

SPECIALS = [0, 1, 2, P - 1, P - 2, (1 << 32) - 1, 1 << 32, (1 << 32) + 1, 1 << 63, P >> 1, W_FROM_RAW7,
            0xFFFFFFFF00000000, P - (1 << 32), 0x8000000080000000]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def edge_words(n, seed):
    rng = np.random.default_rng(seed)
    return np.array([SPECIALS[i] for i in rng.integers(0, len(SPECIALS), n)], dtype=np.uint64)


def edge_elem(rng, ext):
    x = [SPECIALS[i] for i in rng.integers(0, len(SPECIALS), ext)]
    return x[0] if ext == 1 else tuple(x)


def lift(words, ext):
    if ext == 1:
        return words
    out = np.zeros(len(words) * ext, dtype=np.uint64)
    out[::ext] = words
    return out


def test_device_base_products_match_golden(golden):
    cases = [(a, b, c) for a, b, c in golden["field"]["mul"] if a < P and b < P]
    say(f"golden base products: {len(cases)}")
    for a, b, c in cases:
        g = MultilinearProver([0, b])
        g.round(None)
        g.finalize(a)
        assert g.table(0) == [c], (a, b)


@pytest.mark.parametrize("ext", [2, 3])
def test_device_extension_products_match_golden(golden, ext):
    for key, cases in golden["field"]["ext"].items():
        if not key.startswith(f"d{ext}_"):
            continue
        w = int(key.split("_w")[1], 16)
        say(f"golden {key}: {len(cases)}")
        for a, b, c in cases:
            g = MultilinearProver([tuple([0] * ext), tuple(b)], ext=ext, w=w)
            g.round(None)
            g.finalize(tuple(a))
            assert g.table(0) == [tuple(c)], (key, a, b)


def _round_by_round(g, o, v, ext, seed):
    rng = np.random.default_rng(seed)
    prev = None
    for r in range(v):
        rg = g.round(prev)
        ro = o.round(None if prev is None else ([prev] if ext == 1 else list(prev)))
        assert [[x] if ext == 1 else list(x) for x in rg] == ro, f"round {r}"
        prev = edge_elem(rng, ext)
    if prev is not None:
        g.finalize(prev)
        o.finalize([prev] if ext == 1 else list(prev))
    fg = g.final_value()
    assert ([fg] if ext == 1 else list(fg)) == o.final_value()


KINDS = [("ml", 1, 1), ("ip", 1, 1), ("ml", 2, 1), ("ip", 2, 2), ("ip", 3, 3), ("p3", 1, 1), ("p3", 2, 1),
         ("p3", 3, 1), ("p3", 3, 3), ("p4", 2, 2), ("p2", 3, 1)]


@pytest.mark.parametrize("kind,ext,input_ext", KINDS)
@pytest.mark.parametrize("v", [1, 2, 5, 12, 15])
def test_edge_tables_device_loop(kind, ext, input_ext, v):
    k = {"ml": 1, "ip": 2}.get(kind) or int(kind[1:])
    n = 1 << v
    tabs = [edge_words(n * input_ext, 31 * v + i + 7 * ext) for i in range(k)]
    say(f"edge device loop kind={kind} ext={ext}/{input_ext} v={v}")
    g, o = make_pair(kind, ext, input_ext, tabs)
    pg = sumcheck(g, v, Transcript(b"edge", ext=ext))
    po = o.sumcheck(v, O.Transcript.shake(b"edge"))
    assert tuple(as_oracle(pg, ext)) == tuple(po)
    for i in range(k):
        assert np.array_equal(g.table_words(i), o.table(i)), f"table {i}"


@pytest.mark.parametrize("kind,ext,input_ext", KINDS)
@pytest.mark.parametrize("v", [3, 13])
def test_edge_tables_and_edge_challenges_round_by_round(kind, ext, input_ext, v):
    k = {"ml": 1, "ip": 2}.get(kind) or int(kind[1:])
    n = 1 << v
    tabs = [edge_words(n * input_ext, 57 * v + i + 3 * ext) for i in range(k)]
    say(f"edge rounds kind={kind} ext={ext}/{input_ext} v={v}")
    g, o = make_pair(kind, ext, input_ext, tabs)
    _round_by_round(g, o, v, ext, 1000 + v)
    for i in range(k):
        assert np.array_equal(g.table_words(i), o.table(i)), f"table {i}"


@pytest.mark.parametrize("kind", ["p3", "ip"])
def test_edge_tables_large(kind):
    # v = 21: the product kernels' two-quadruple form (2^18 quadruples and up)
    # and the base-coefficient first fold; IP: nontemporal passes from v = 24
    v = 21 if kind == "p3" else 24
    ext = 3 if kind == "p3" else 1
    k = 3 if kind == "p3" else 2
    n = 1 << v
    tabs = [edge_words(n, 4242 + i) for i in range(k)]
    say(f"edge large kind={kind} v={v}")
    g, o = make_pair(kind, ext, 1, tabs)
    pg = sumcheck(g, v, TestTranscript(5, ext=ext))
    po = o.sumcheck(v, O.Transcript.rng(5))
    assert tuple(as_oracle(pg, ext)) == tuple(po)


@pytest.mark.parametrize("ext", [1, 2, 3])
@pytest.mark.parametrize("v", [4, 13, 17])
def test_edge_eq_factored(ext, v):
    n = 1 << v
    p = edge_words(n * ext, 900 + v + ext)
    wpt = edge_words(v * ext, 950 + v + ext)
    say(f"edge eq ext={ext} v={v}")
    g, o = make_pair("eq", ext, ext, [p], wpt)
    pg = sumcheck(g, v, Transcript(b"edge-eq", ext=ext))
    po = o.sumcheck(v, O.Transcript.shake(b"edge-eq"))
    assert tuple(as_oracle(pg, ext)) == tuple(po)
    for i in range(3):
        assert np.array_equal(g.table_words(i), o.table(i)), f"table {i}"
    g2, o2 = make_pair("eq", ext, ext, [p], wpt)
    _round_by_round(g2, o2, v, ext, 77 + v)


@pytest.mark.parametrize("ext,input_ext", [(1, 1), (2, 1), (3, 3)])
@pytest.mark.parametrize("n", [8, 64, 256])
def test_edge_gkr(n, ext, input_ext):
    add = edge_words(n * n * input_ext, n + 1)
    mult = edge_words(n * n * input_ext, n + 2)
    wit = edge_words(n * input_ext, n + 3)
    v = (n * n - 1).bit_length()
    say(f"edge gkr n={n} ext={ext}/{input_ext}")
    g = GkrProver(add, mult, wit, ext=ext, input_ext=input_ext)
    o = O.Prover.gkr(lift(add, ext // input_ext), lift(mult, ext // input_ext), lift(wit, ext // input_ext), ext=ext)
    pg = sumcheck(g, v, Transcript(b"edge-gkr", ext=ext))
    po = o.sumcheck(v, O.Transcript.shake(b"edge-gkr"))
    assert tuple(as_oracle(pg, ext)) == tuple(po)


@pytest.mark.parametrize("kind", ["ml", "ip"])
def test_edge_lsb(kind):
    v = 12
    n = 1 << v
    k = 1 if kind == "ml" else 2
    tabs = [edge_words(n, 600 + i) for i in range(k)]
    say(f"edge lsb kind={kind}")
    if kind == "ml":
        g, o = MultilinearProverLSB(tabs[0]), O.Prover.multilinear_lsb(tabs[0])
    else:
        g, o = InnerProductProverLSB(*tabs), O.Prover.inner_product_lsb(*tabs)
    pg = sumcheck(g, v, Transcript(b"edge-lsb"))
    po = o.sumcheck(v, O.Transcript.shake(b"edge-lsb"))
    assert tuple(as_oracle(pg, 1)) == tuple(po)
