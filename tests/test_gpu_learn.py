"""The learned service-time model: its GPU forward and backward kernels
against the numpy oracle, training, and what it gains the served scheduler
over first-come-first-served admission (with identical proofs)."""
import random
import sys

import numpy as np
import pytest

import mlp as M
from effsc import Transcript
from effsc.learn import INPUTS, Network, ServiceTimeModel, features, measure_service_times, work_priority
from effsc.serve import INNER_PRODUCT, MULTILINEAR, PRODUCT3_EXT3, Server

pytestmark = pytest.mark.gpu
EXT = {MULTILINEAR: 1, INNER_PRODUCT: 1, PRODUCT3_EXT3: 3}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("hidden,n", [(32, 300), (7, 5), (64, 1000)])
def test_forward_and_gradient_match_oracle(hidden, n):
    rng = np.random.default_rng(hidden + n)
    net = Network(INPUTS, hidden, seed=hidden)
    p = rng.normal(0, 0.4, net.count).astype(np.float32)
    net.set_params(p)
    X = rng.normal(0, 1, (n, INPUTS)).astype(np.float32)
    T = rng.normal(0, 1, n).astype(np.float32)
    say(f"mlp hidden={hidden} n={n}")
    y = net.forward(X)
    y_ref = M.forward(p.astype(np.float64), X.astype(np.float64), INPUTS, hidden)
    # fp32 kernels against the float64 oracle: relative 1e-5 of the output scale
    assert np.allclose(y, y_ref, rtol=1e-5, atol=1e-5 * max(1.0, np.abs(y_ref).max()))
    loss, g = net.loss_grad(X, T)
    loss_ref, g_ref = M.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64), INPUTS, hidden)
    assert abs(loss - loss_ref) <= 1e-5 * max(1.0, loss_ref)
    scale = np.abs(g_ref).max()
    assert np.allclose(g, g_ref, rtol=1e-4, atol=1e-5 * scale)
    net.close()


def test_training_fits_a_known_function():
    # targets from a smooth function of the features: the network must fit it
    rng = random.Random(5)
    reqs = [(rng.choice([MULTILINEAR, INNER_PRODUCT, PRODUCT3_EXT3]), rng.randint(4, 26)) for _ in range(400)]
    X = np.array([features(k, v) for k, v in reqs], dtype=np.float32)
    T = np.array([np.log(0.02 * v + 3e-7 * (1 << v) * {1: 1, 2: 2, 3: 9}[k]) for k, v in reqs], dtype=np.float32)
    net = Network(INPUTS, 32, seed=7)
    l0, _ = net.loss_grad(X, T)
    l1 = net.train(X, T, steps=3000, lr=1e-2)
    say(f"loss {l0:.4f} -> {l1:.5f}")
    assert l1 < 0.02 * l0 and l1 < 0.01
    # the trained parameters reproduce the loss on the oracle
    lo, _ = M.loss_grad(net.params().astype(np.float64), X, T, INPUTS, 32)
    assert abs(lo - l1) <= 1e-4 * max(1.0, lo)


def _burst(seed, count=24):
    rng = random.Random(seed)
    return [(rng.choice([MULTILINEAR, INNER_PRODUCT, PRODUCT3_EXT3]), rng.randint(8, 18)) for _ in range(count)]


def _run_burst(srv, reqs, priorities):
    ids = [srv.submit(k, v, Transcript(b"burst-%d" % i, ext=EXT[k]), seeds=[i + 1, i + 2, i + 3], priority=pr)
           for i, ((k, v), pr) in enumerate(zip(reqs, priorities))]
    return [srv.wait(r) for r in ids]


def test_learned_admission_beats_fifo_with_identical_proofs():
    # calibration: every (kind, v) alone, on the device
    grid = [(k, v) for k in (MULTILINEAR, INNER_PRODUCT, PRODUCT3_EXT3) for v in range(6, 21)]
    with Server(hbm_budget=1 << 30, max_active=1) as srv:
        measure_service_times(srv, grid[:3], lambda i, k: Transcript(b"warm", ext=EXT[k]))
        samples = measure_service_times(srv, grid, lambda i, k: Transcript(b"cal-%d" % i, ext=EXT[k]))
    model = ServiceTimeModel(hidden=32, seed=3)
    loss = model.fit(samples, steps=3000, lr=1e-2)
    say(f"service-time model loss {loss:.4f} on {len(samples)} samples")
    assert loss < 0.05
    reqs = _burst(11)
    with Server(hbm_budget=1 << 30, max_active=2) as srv:
        _run_burst(srv, reqs[:4], [0.0] * 4)  # warm-up
        fifo = _run_burst(srv, reqs, [0.0] * len(reqs))
        learned = _run_burst(srv, reqs, [model.priority(k, v) for k, v in reqs])
        # the non-learned ordering on the same features: shortest work first
        swf = _run_burst(srv, reqs, [work_priority(k, v) for k, v in reqs])
    lat_f = np.mean([r.latency for r in fifo])
    lat_l = np.mean([r.latency for r in learned])
    lat_h = np.mean([r.latency for r in swf])
    say(f"mean latency FIFO {lat_f * 1e3:.3f} ms, shortest work first {lat_h * 1e3:.3f} ms, "
        f"learned {lat_l * 1e3:.3f} ms")
    assert all(a.proof == b.proof == c.proof for a, b, c in zip(fifo, learned, swf))
    assert lat_l < 0.85 * lat_f
