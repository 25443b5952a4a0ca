"""Batched proofs (effsc_hip_batch_*): every instance's proof equals the
oracle's proof of that instance alone, on graph replays and plain launches."""
import random

import pytest

import oracle as O
from effsc import InnerProductProver, MultilinearProver, Transcript, sumcheck
from effsc.batch import BatchProver, DeviceTable

pytestmark = pytest.mark.gpu


def _instances(count, vmin, vmax, seed):
    rng = random.Random(seed)
    return [rng.randint(vmin, vmax) for _ in range(count)]


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_oracle_per_instance(kind, graph):
    vs = _instances(24, 0, 12, 1234 + kind)
    seeds = [(0xB0000 + 2 * i, 0xB0000 + 2 * i + 1) for i in range(len(vs))]
    a = [DeviceTable.uniform(1 << v, s[0]) for v, s in zip(vs, seeds)]
    b = [DeviceTable.uniform(1 << v, s[1]) for v, s in zip(vs, seeds)] if kind == 2 else None
    bp = BatchProver(a, b, kind=kind, graph=graph)
    expected = []
    for i, v in enumerate(vs):
        ta = O.uniform(1 << v, seeds[i][0])
        o = O.Prover.multilinear(ta) if kind == 1 else O.Prover.inner_product(ta, O.uniform(1 << v, seeds[i][1]))
        expected.append(o.sumcheck(v, O.Transcript.shake(b"batch-%d" % i)))
    for replay in range(3):
        print(f"batch kind={kind} graph={graph} replay={replay}", flush=True)
        proofs = bp.prove([Transcript(b"batch-%d" % i) for i in range(len(vs))])
        for i, (p, (rp, ch, fin)) in enumerate(zip(proofs, expected)):
            assert p.round_polys == [[x[0] for x in r] for r in rp], f"instance {i} v={vs[i]}"
            assert p.challenges == [c[0] for c in ch]
            assert p.final_value == fin[0]
        # eager work on the same device between graph replays
        sumcheck(InnerProductProver(O.uniform(64, 1), O.uniform(64, 2)), 6, Transcript(b"eager"))


def test_batch_equals_single_prover_at_larger_sizes():
    vs = [16, 12, 18, 3, 14, 17]
    a = [DeviceTable.uniform(1 << v, 0xC1000 + i) for i, v in enumerate(vs)]
    b = [DeviceTable.uniform(1 << v, 0xC2000 + i) for i, v in enumerate(vs)]
    bp = BatchProver(a, b, kind=2)
    proofs = bp.prove([Transcript(b"big-%d" % i) for i in range(len(vs))])
    for i, v in enumerate(vs):
        single = sumcheck(InnerProductProver(O.uniform(1 << v, 0xC1000 + i), O.uniform(1 << v, 0xC2000 + i)), v,
                          Transcript(b"big-%d" % i))
        assert proofs[i].round_polys == [list(r) for r in single.round_polys]
        assert proofs[i].final_value == single.final_value


@pytest.mark.parametrize("graph", [True, False])
def test_autotuned_batch_matches_oracle(graph):
    vs = _instances(16, 0, 11, 4321)
    a = [DeviceTable.uniform(1 << v, 0xD0000 + i) for i, v in enumerate(vs)]
    b = [DeviceTable.uniform(1 << v, 0xD1000 + i) for i, v in enumerate(vs)]
    bp = BatchProver(a, b, kind=2, graph=graph)
    print(f"autotune graph={graph}", flush=True)
    chosen, times = bp.autotune(candidates=(1, 7, 64, 2048), reps=1)
    assert chosen in (1, 7, 64, 2048) and all(t > 0 for t in times.values())
    assert times[chosen] == min(times.values())
    proofs = bp.prove([Transcript(b"tuned-%d" % i) for i in range(len(vs))])
    for i, v in enumerate(vs):
        o = O.Prover.inner_product(O.uniform(1 << v, 0xD0000 + i), O.uniform(1 << v, 0xD1000 + i))
        rp, ch, fin = o.sumcheck(v, O.Transcript.shake(b"tuned-%d" % i))
        assert proofs[i].round_polys == [[x[0] for x in r] for r in rp], f"instance {i} v={v}"
        assert proofs[i].final_value == fin[0]


def test_full_size_batch_matches_oracle():
    # the benchmark batch: 512 inner-product instances, v ~ U[12, 24] (about
    # 21 GB of tables), one batched proof; every instance against the
    # oracle's threaded restatement of the reference's fused path (pinned to
    # the oracle prover and the golden proofs on the CPU)
    import random as _r

    rng = _r.Random(0x5EED0004)
    vs = [rng.randint(12, 24) for _ in range(512)]
    print(f"full-size batch: {len(vs)} instances, {2 * sum(1 << v for v in vs) * 8 / 2**30:.1f} GiB", flush=True)
    a = [DeviceTable.uniform(1 << v, 0x5EED0004 + 2 * i) for i, v in enumerate(vs)]
    b = [DeviceTable.uniform(1 << v, 0x5EED0004 + 2 * i + 1) for i, v in enumerate(vs)]
    bp = BatchProver(a, b, kind=2, graph=True)
    proofs = bp.prove([Transcript(b"effsc-hip/bench/ip") for _ in vs])
    for i, v in enumerate(vs):
        if i % 64 == 0:
            print(f"checking instances {i}..{i + 63}", flush=True)
        _, polys, chals, fin = O.bench_ip(O.uniform(1 << v, 0x5EED0004 + 2 * i), O.uniform(1 << v, 0x5EED0004 + 2 * i + 1),
                                          v, 16, fused=True)
        assert [w for r in proofs[i].round_polys for w in r] == [int(x) for x in polys], f"instance {i} v={v}"
        assert proofs[i].challenges == [int(x) for x in chals]
        assert proofs[i].final_value == fin
