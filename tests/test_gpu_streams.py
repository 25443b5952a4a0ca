"""Streamed input on the GPU: claims and provers built from FileStreams equal
the oracle on the same tables, whatever the chunking of the upload."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from effsc import TestTranscript, Transcript, sumcheck
from effsc import _native as N
from effsc.streams import (AUTO, OUT_OF_CORE, FileStream, inner_product_from_stream, multilinear_from_stream,
                           multivariate_claim, multivariate_product_claim)

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


def files(tmp_path, count, n, seed, ext=1):
    out = []
    for i in range(count):
        d = O.uniform(n * ext, seed + i)
        p = tmp_path / f"t{i}.bin"
        FileStream.write_to_file(p, d, ext=ext)
        out.append((p, d))
    return out


@pytest.mark.parametrize("chunk", ["16", "4096", None])
def test_claims(tmp_path, monkeypatch, chunk):
    if chunk:
        monkeypatch.setenv("EFFSC_HIP_STREAM_CHUNK", chunk)
    else:
        monkeypatch.delenv("EFFSC_HIP_STREAM_CHUNK", raising=False)
    fs = files(tmp_path, 3, 1 << 14, 50)
    say(f"claims chunk={chunk}")
    s = [FileStream(p) for p, _ in fs]
    assert multivariate_claim(s[0]) == sum(int(x) for x in fs[0][1]) % P
    prod = sum(int(a) * int(b) * int(c) for a, b, c in zip(fs[0][1], fs[1][1], fs[2][1])) % P
    assert multivariate_product_claim(s) == prod
    # ext3 product claim against the oracle's extension arithmetic
    e = files(tmp_path, 2, 256, 60, ext=3)
    s3 = [FileStream(p, ext=3) for p, _ in e]
    acc = [0, 0, 0]
    for i in range(256):
        m = O.ext_mul(3, O.W_FROM_RAW7, e[0][1][3 * i:3 * i + 3], e[1][1][3 * i:3 * i + 3])
        acc = [(x + int(y)) % P for x, y in zip(acc, m)]
    assert multivariate_product_claim(s3) == tuple(acc)


@pytest.mark.parametrize("kind", ["ml", "ip"])
@pytest.mark.parametrize("v,chunk", [(4, "1"), (10, "8"), (16, "1024"), (20, None), (14, "3")])
def test_stream_prover_matches_oracle(tmp_path, monkeypatch, kind, v, chunk):
    if chunk:
        monkeypatch.setenv("EFFSC_HIP_STREAM_CHUNK", chunk)
    else:
        monkeypatch.delenv("EFFSC_HIP_STREAM_CHUNK", raising=False)
    k = 1 if kind == "ml" else 2
    fs = files(tmp_path, k, 1 << v, 300 + v)
    say(f"stream prover kind={kind} v={v} chunk={chunk}")
    s = [FileStream(p) for p, _ in fs]
    g = multilinear_from_stream(s[0]) if k == 1 else inner_product_from_stream(s[0], s[1])
    o = O.Prover.multilinear(fs[0][1]) if k == 1 else O.Prover.inner_product(fs[0][1], fs[1][1])
    po = o.sumcheck(v, O.Transcript.shake(b"stream"))
    assert tuple(as_oracle(sumcheck(g, v, Transcript(b"stream")), 1)) == tuple(po)
    g.reset()  # the accumulated first pass stays valid: the input is never written
    assert tuple(as_oracle(sumcheck(g, v, Transcript(b"stream")), 1)) == tuple(po)
    # the per-round path (host loop) over the same uploaded tables
    g.reset()
    o2 = O.Prover.multilinear(fs[0][1]) if k == 1 else O.Prover.inner_product(fs[0][1], fs[1][1])
    hooked = sumcheck(g, v, TestTranscript(9), lambda r, t: None)
    assert tuple(as_oracle(hooked, 1)) == tuple(o2.sumcheck(v, O.Transcript.rng(9)))


def test_stream_prover_ext3_challenges(tmp_path):
    fs = files(tmp_path, 2, 1 << 12, 77)
    s = [FileStream(p) for p, _ in fs]
    g = inner_product_from_stream(s[0], s[1], ext=3)
    o = O.Prover.product([fs[0][1], fs[1][1]], ext=3)
    pg = sumcheck(g, 12, Transcript(b"s3", ext=3))
    assert tuple(as_oracle(pg, 3)) == tuple(o.sumcheck(12, O.Transcript.shake(b"s3")))


def test_stream_prover_rejects_non_canonical(tmp_path, monkeypatch):
    monkeypatch.setenv("EFFSC_HIP_STREAM_CHUNK", "64")
    d = O.uniform(1 << 12, 5)
    d[1000] = P + 5
    p = tmp_path / "bad.bin"
    FileStream.write_to_file(p, d)
    with pytest.raises(N.NativeError):
        multilinear_from_stream(FileStream(p))
    with pytest.raises(N.NativeError):
        multivariate_claim(FileStream(p))


def _make(kind, s, mode):
    return multilinear_from_stream(s[0], mode=mode) if kind == "ml" else inner_product_from_stream(s[0], s[1], mode=mode)


@pytest.mark.parametrize("kind", ["ml", "ip"])
@pytest.mark.parametrize("v,chunk,tail", [(17, None, None), (18, "4096", None), (19, "64", None), (16, "1024", "4096")])
def test_out_of_core_prover_matches_oracle(tmp_path, monkeypatch, kind, v, chunk, tail):
    # the input stays in the files: rounds 0-1 while it streams through the
    # staging at construction, rounds 2-3 while it streams again in the proof
    if os.environ.get("EFFSC_HIP_ONE_ROUND") or os.environ.get("EFFSC_HIP_HOST_LOOP"):
        pytest.skip("out-of-core provers run on the device loop with two-round passes only")
    for name, val in (("EFFSC_HIP_STREAM_CHUNK", chunk), ("EFFSC_HIP_TAIL", tail)):
        if val:
            monkeypatch.setenv(name, val)
        else:
            monkeypatch.delenv(name, raising=False)
    k = 1 if kind == "ml" else 2
    fs = files(tmp_path, k, 1 << v, 900 + v)
    say(f"out-of-core prover kind={kind} v={v} chunk={chunk} tail={tail}")
    s = [FileStream(p) for p, _ in fs]
    g = _make(kind, s, OUT_OF_CORE)
    # before its whole proof the prover has no resident input
    with pytest.raises(N.NativeError):
        g.table_words(0)
    with pytest.raises(N.NativeError):
        sumcheck(g, v - 1, Transcript(b"ooc"))
    with pytest.raises(N.NativeError):
        sumcheck(g, v, Transcript(b"ooc"), lambda r, t: None)
    o = O.Prover.multilinear(fs[0][1]) if k == 1 else O.Prover.inner_product(fs[0][1], fs[1][1])
    po = o.sumcheck(v, O.Transcript.shake(b"ooc"))
    assert tuple(as_oracle(sumcheck(g, v, Transcript(b"ooc")), 1)) == tuple(po)
    for i in range(k):
        assert np.array_equal(g.table_words(i), o.table(i)), f"table {i}"
    # reset streams the input again; the test transcript gives other folds
    g.reset()
    o2 = O.Prover.multilinear(fs[0][1]) if k == 1 else O.Prover.inner_product(fs[0][1], fs[1][1])
    assert tuple(as_oracle(sumcheck(g, v, TestTranscript(5)), 1)) == tuple(o2.sumcheck(v, O.Transcript.rng(5)))
    g.close()


def test_out_of_core_rejects_small_and_auto_picks_resident(tmp_path, monkeypatch):
    monkeypatch.delenv("EFFSC_HIP_TAIL", raising=False)
    # 2^14 elements: n/4 = 4096, the inner product's tail size
    fs = files(tmp_path, 2, 1 << 14, 31)
    s = [FileStream(p) for p, _ in fs]
    with pytest.raises(N.NativeError):
        inner_product_from_stream(s[0], s[1], mode=OUT_OF_CORE)  # n/4 is not above the tail
    g = inner_product_from_stream(s[0], s[1], mode=AUTO)  # fits: resident, per-round calls allowed
    o = O.Prover.inner_product(fs[0][1], fs[1][1])
    hooked = sumcheck(g, 14, TestTranscript(3), lambda r, t: None)
    assert tuple(as_oracle(hooked, 1)) == tuple(o.sumcheck(14, O.Transcript.rng(3)))
