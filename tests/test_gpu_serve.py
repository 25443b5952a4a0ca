"""Served proofs (effsc_hip_server_*): every request's proof equals the
oracle's proof of that request alone, whether it ran alone or in a mix of
requests that arrived while others were on the device."""
import ctypes as C
import random
import sys
import time

import pytest

import oracle as O
from effsc import TestTranscript, Transcript
from effsc import _native as N
from effsc.batch import DeviceTable
from effsc.serve import INNER_PRODUCT, MULTILINEAR, PRODUCT3_EXT3, Server

pytestmark = pytest.mark.gpu

EXT = {MULTILINEAR: 1, INNER_PRODUCT: 1, PRODUCT3_EXT3: 3}
K = {MULTILINEAR: 1, INNER_PRODUCT: 2, PRODUCT3_EXT3: 3}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_proof(kind, v, seeds, domsep=None, rng_seed=None):
    n = 1 << v
    tabs = [O.uniform(n, s) for s in seeds[:K[kind]]]
    if kind == MULTILINEAR:
        p = O.Prover.multilinear(tabs[0])
    elif kind == INNER_PRODUCT:
        p = O.Prover.inner_product(tabs[0], tabs[1])
    else:
        p = O.Prover.product(tabs, ext=3)
    t = O.Transcript.shake(domsep) if domsep is not None else O.Transcript.rng(rng_seed)
    return p.sumcheck(v, t)


def words(served, ext):
    pr = served.proof
    rp = [[[x] if ext == 1 else list(x) for x in r] for r in pr.round_polys]
    ch = [[c] if ext == 1 else list(c) for c in pr.challenges]
    fin = [pr.final_value] if ext == 1 else list(pr.final_value)
    return rp, ch, fin


def make_requests(count, vmax, seed):
    rng = random.Random(seed)
    reqs = []
    for i in range(count):
        kind = (MULTILINEAR, INNER_PRODUCT, PRODUCT3_EXT3)[i % 3]
        v = rng.randint(0, vmax)
        seeds = [rng.getrandbits(40) for _ in range(3)]
        reqs.append((kind, v, seeds, b"serve-%d" % i))
    return reqs


def test_mixed_requests_match_oracle_alone_and_in_the_mix():
    reqs = make_requests(30, 12, 99)
    # a small arena and few slots force queueing; the second half arrives
    # while the first half is on the device
    with Server(hbm_budget=4 << 20, max_active=5) as srv:
        ids = []
        for i, (kind, v, seeds, ds) in enumerate(reqs):
            say(f"submit {i} kind={kind} v={v}")
            ids.append(srv.submit(kind, v, Transcript(ds, ext=EXT[kind]), seeds=seeds))
            if i == 14:
                first = srv.wait(ids[0])
                time.sleep(0.002)
        results = [first] + [srv.wait(r) for r in ids[1:]]
        st = srv.stats()
    say("stats", st)
    assert st["completed"] == len(reqs) and st["peak_active"] <= 5 and st["peak_bytes"] <= 4 << 20
    for i, ((kind, v, seeds, ds), res) in enumerate(zip(reqs, results)):
        assert tuple(words(res, EXT[kind])) == tuple(oracle_proof(kind, v, seeds, domsep=ds)), f"request {i}"
        assert res.submitted <= res.admitted <= res.completed
    # the same requests alone, one server each, give the same proofs
    for i in (0, 7, 14, 22, 29):
        kind, v, seeds, ds = reqs[i]
        with Server(hbm_budget=64 << 20, max_active=1) as srv:
            alone = srv.wait(srv.submit(kind, v, Transcript(ds, ext=EXT[kind]), seeds=seeds))
        assert alone.proof == results[i].proof, f"request {i} alone"


def test_large_requests_and_test_transcript():
    with Server(hbm_budget=1 << 30, max_active=8) as srv:
        reqs = [(INNER_PRODUCT, 20, [11, 12, 0]), (MULTILINEAR, 18, [13, 0, 0]), (PRODUCT3_EXT3, 16, [14, 15, 16]),
                (INNER_PRODUCT, 1, [17, 18, 0]), (PRODUCT3_EXT3, 0, [19, 20, 21])]
        ids = [srv.submit(k, v, TestTranscript(5 + i, ext=EXT[k]), seeds=s) for i, (k, v, s) in enumerate(reqs)]
        for i, ((k, v, s), rid) in enumerate(zip(reqs, ids)):
            say(f"large request {i} kind={k} v={v}")
            res = srv.wait(rid)
            assert tuple(words(res, EXT[k])) == tuple(oracle_proof(k, v, s, rng_seed=5 + i))


def test_served_large_requests_under_poisson_arrivals():
    # the served configuration's upper range: v = 22..26 of every kind,
    # arriving as a Poisson process while the earlier ones run, against the
    # oracle (the inner products through its threaded restatement of the
    # reference's fused path)
    rng = random.Random(0x5EED0007)
    reqs = [(MULTILINEAR, 26), (INNER_PRODUCT, 25), (PRODUCT3_EXT3, 22), (MULTILINEAR, 23), (INNER_PRODUCT, 22),
            (PRODUCT3_EXT3, 24), (INNER_PRODUCT, 26), (MULTILINEAR, 22)]
    reqs = [(k, v, [rng.getrandbits(40) for _ in range(3)], b"served-large-%d" % i) for i, (k, v) in enumerate(reqs)]
    with Server(hbm_budget=24 << 30, max_active=4) as srv:
        ids = []
        for i, (k, v, seeds, ds) in enumerate(reqs):
            say(f"submit large {i} kind={k} v={v}")
            ids.append(srv.submit(k, v, Transcript(ds, ext=EXT[k]), seeds=seeds))
            time.sleep(rng.expovariate(1000.0))
        results = [srv.wait(r) for r in ids]
    for i, ((k, v, seeds, ds), res) in enumerate(zip(reqs, results)):
        say(f"check large {i} kind={k} v={v}")
        if k == INNER_PRODUCT:
            n = 1 << v
            _, polys, chals, fin = O.bench_ip(O.uniform(n, seeds[0]), O.uniform(n, seeds[1]), v, 16, fused=True,
                                              domsep=ds)
            assert [w for r in res.proof.round_polys for w in r] == [int(x) for x in polys], f"request {i}"
            assert res.proof.challenges == [int(x) for x in chals]
            assert res.proof.final_value == fin
        else:
            assert tuple(words(res, EXT[k])) == tuple(oracle_proof(k, v, seeds, domsep=ds)), f"request {i}"


def test_borrowed_device_tables():
    v = 13
    a, b = DeviceTable.uniform(1 << v, 31), DeviceTable.uniform(1 << v, 32)
    pa, pb = C.cast(a.ptr, C.c_void_p).value, C.cast(b.ptr, C.c_void_p).value
    with Server(hbm_budget=16 << 20, max_active=4) as srv:
        assert Server.footprint(INNER_PRODUCT, v) - Server.footprint(INNER_PRODUCT, v, tables=[pa, pb]) \
            == 2 * (1 << v) * 8
        both = srv.submit(INNER_PRODUCT, v, Transcript(b"borrowed"), tables=[pa, pb])
        # table a borrowed, table b made on the device from seed 32
        mixed = srv.submit(INNER_PRODUCT, v, Transcript(b"borrowed"), seeds=[0, 32, 0], tables=[pa, None])
        r_both, r_mixed = srv.wait(both), srv.wait(mixed)
    expected = tuple(oracle_proof(INNER_PRODUCT, v, [31, 32], domsep=b"borrowed"))
    assert tuple(words(r_both, 1)) == expected
    assert tuple(words(r_mixed, 1)) == expected
    a.free()
    b.free()


def test_footprint_over_budget_is_rejected():
    with Server(hbm_budget=1 << 20, max_active=2) as srv:
        with pytest.raises(N.NativeError) as e:
            srv.submit(INNER_PRODUCT, 20, Transcript(b"big"), seeds=[1, 2, 0])
        assert e.value.code == 1
        # the server still serves what fits
        res = srv.wait(srv.submit(MULTILINEAR, 10, Transcript(b"ok"), seeds=[3, 0, 0]))
        assert tuple(words(res, 1)) == tuple(oracle_proof(MULTILINEAR, 10, [3], domsep=b"ok"))


def test_admission_follows_priority():
    with Server(hbm_budget=1 << 30, max_active=1) as srv:
        blocker = srv.submit(INNER_PRODUCT, 22, Transcript(b"blocker"), seeds=[1, 2, 0], priority=0.0)
        ids = {p: srv.submit(MULTILINEAR, 8, Transcript(b"p%d" % p), seeds=[p, 0, 0], priority=float(p))
               for p in (3, 1, 2)}
        b = srv.wait(blocker)
        res = {p: srv.wait(r) for p, r in ids.items()}
    if max(r.submitted for r in res.values()) > b.completed:
        pytest.skip("the blocker finished before the others were submitted")
    assert res[1].admitted <= res[2].admitted <= res[3].admitted
    for p, r in res.items():
        assert tuple(words(r, 1)) == tuple(oracle_proof(MULTILINEAR, 8, [p], domsep=b"p%d" % p))
