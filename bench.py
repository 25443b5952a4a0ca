"""Benchmark: Goldilocks inner-product sumcheck (InnerProductProver, MSB),
v = 28, two 2^28-element tables uniform in [0, p) from fixed seeds, all 28
rounds with the SHAKE256 Fiat-Shamir transcript, on one MI355X.

A step is one full proof (28 rounds + finalize + final value) over tables
already resident in HBM.  Metric: evals/s = 2^v / T_proof (the reference's
criterion convention, Throughput::Elements(2^v), benches/sumcheck.rs:28,80).

Prints one JSON line.  Besides the contract fields it carries
  roofline     - the dominant kernel's achieved bandwidth: the algorithmic
                 bytes of all its launches in the timed region (each input
                 element read once, each output element written once) divided
                 by their summed HIP-event durations (events recorded on the
                 prover's stream around each launch);
  cpu_baseline - the oracle's restatement of the reference's
                 InnerProductProver (parallel folds, serial evaluate), timed on
                 this host's cores on a bounded sample.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "efficient-sumcheck_amd"))

HBM_PEAK_GBS = 8000.0
SEED_A, SEED_B = 0x5EED0002, 0x5EED0003
DOMSEP = b"effsc-hip/bench/ip"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(v, tables=2):
    # per table: round 0 reads N; round j >= 1 reads L_j and writes L_j/2;
    # finalize folds 2 -> 1 (reads 2, writes 1)
    n = 1 << v
    words = n  # round 0
    L = n
    while L >= 2:
        words += L + L // 2
        L //= 2
    return 8 * tables * words


def recorded_traffic(path=os.path.join(ROOT, "profiles", "k_pass_ip_v28_traffic.md")):
    # HBM bytes per k_pass launch from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs (scaled by a 1 GiB copy), or None when absent
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(v_sample, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n = 1 << v_sample
    a = O.uniform(n, SEED_A)
    b = O.uniform(n, SEED_B)
    t, _, _, _ = O.bench_ip(a, b, v_sample, threads)
    return n / t, t


def bench_batch(args):
    # Batched inner-product proofs: `count` instances, v_i uniform in
    # [vmin, vmax] from a fixed seed, uniform tables; the batch (one launch per
    # round for all instances, hipGraph replay) against per-instance proofs
    # (one device-resident proof per instance, the same inputs and outputs).
    import random

    import effsc
    from effsc import _native as N
    from effsc.batch import BatchProver, DeviceTable

    rng = random.Random(0x5EED0004)
    vs = [rng.randint(args.vmin, args.vmax) for _ in range(args.count)]
    a = [DeviceTable.uniform(1 << v, 0x5EED0004 + 2 * i) for i, v in enumerate(vs)]
    b = [DeviceTable.uniform(1 << v, 0x5EED0004 + 2 * i + 1) for i, v in enumerate(vs)]
    total = sum(1 << v for v in vs)
    log(f"bench: batch of {len(vs)} inner-product instances, v in [{args.vmin},{args.vmax}], "
        f"{2 * total * 8 / 2**30:.1f} GiB of tables")
    bp = BatchProver(a, b, kind=2, graph=True)

    def ts():
        return [effsc.Transcript(b"effsc-hip/batch/%d" % i) for i in range(len(vs))]

    for _ in range(args.warmup):
        ref = bp.prove(ts())
    N.check(N.lib().effsc_hip_synchronize(0))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = bp.prove(ts())
    T_batch = (time.perf_counter() - t0) / args.steps
    assert all(p.round_polys == q.round_polys and p.final_value == q.final_value for p, q in zip(out, ref))
    # per-instance baseline
    provers = [effsc.InnerProductProver(C.cast(x.ptr, C.c_void_p).value, C.cast(y.ptr, C.c_void_p).value,
                                        borrow=True, n=x.n)
               for x, y in zip(a, b)]
    def single_all():
        res = []
        for i, (p, v) in enumerate(zip(provers, vs)):
            p.reset()
            res.append(effsc.sumcheck(p, v, effsc.Transcript(b"effsc-hip/batch/%d" % i)))
        return res
    single_all()
    t1 = time.perf_counter()
    reps = max(1, args.steps // 2)
    for _ in range(reps):
        single = single_all()
    T_single = (time.perf_counter() - t1) / reps
    same = all(p.round_polys == [list(r) for r in q.round_polys] and p.final_value == q.final_value
               for p, q in zip(out, single))
    print(json.dumps({"metric": "batched proofs/s (512 inner-product instances, v ~ U[12,24])",
                      "value": len(vs) / T_batch, "unit": "proofs/s", "ms_per_batch": T_batch * 1e3,
                      "per_instance_launches_proofs_per_s": len(vs) / T_single,
                      "per_instance_ms": T_single * 1e3, "speedup": T_single / T_batch,
                      "bit_identical_to_per_instance": same, "instances": len(vs),
                      "table_bytes": 2 * total * 8}), flush=True)
    for p in provers:
        p.close()
    bp.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--v", type=int, default=28)
    ap.add_argument("--cpu-v", type=int, default=None, help="CPU baseline sample size (default: v)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["ip", "batch"], default="ip",
                    help="ip: the v=28 inner-product proof (default, the headline line); batch: batched proofs")
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--vmin", type=int, default=12)
    ap.add_argument("--vmax", type=int, default=24)
    ap.add_argument("--calibrate", action="store_true",
                    help="one 1 GiB device-to-device copy before the timed region (a known-size kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every run here uses one GPU).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        return bench_batch(args)

    import effsc
    from effsc import _native as N

    L = N.lib()
    v = args.v
    n = 1 << v
    dev = 0
    ptrs = []
    for seed in (SEED_A, SEED_B):
        p = N.u64p()
        N.check(L.effsc_hip_device_alloc(dev, n, C.byref(p)))
        N.check(L.effsc_hip_fill_uniform(dev, p, n, seed, 0))
        ptrs.append(p)
    addr = [C.cast(p, C.c_void_p).value for p in ptrs]
    if args.calibrate:
        cw = min(n, 1 << 27)
        tmp = N.u64p()
        N.check(L.effsc_hip_device_alloc(dev, cw, C.byref(tmp)))
        N.check(L.effsc_hip_copy(dev, tmp, ptrs[0], cw))
        N.check(L.effsc_hip_device_free(dev, tmp))
        log(f"bench: calibration copy of {cw * 8} bytes done")
    prover = effsc.InnerProductProver(addr[0], addr[1], borrow=True, n=n, device=dev)
    log(f"bench: inner product v={v}, tables resident ({2 * n * 8 / 2**30:.1f} GiB)")

    def step():
        prover.reset()
        return effsc.sumcheck(prover, v, effsc.Transcript(DOMSEP))

    for _ in range(args.warmup):
        step()
    N.check(L.effsc_hip_synchronize(dev))
    N.check(L.effsc_hip_profile(prover.h, 1))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        proof = step()
    N.check(L.effsc_hip_synchronize(dev))
    elapsed = time.perf_counter() - t0
    launches = (C.c_uint64 * 4)()
    ms = (C.c_double * 4)()
    nbytes = (C.c_double * 4)()
    N.check(L.effsc_hip_profile_read(prover.h, launches, ms, nbytes))
    N.check(L.effsc_hip_profile(prover.h, 0))
    # the profiled region adds event records; time an unprofiled region too
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    N.check(L.effsc_hip_synchronize(dev))
    elapsed_plain = time.perf_counter() - t1
    T = min(elapsed, elapsed_plain) / args.steps

    # dominant kernel class: 0 fused fold+evaluate, 1 evaluate, 2 fold, 3 two-round pass
    names = {0: "k_fused<Fp,Fp,2,2> (fold + evaluate)", 1: "k_eval<Fp,2,2> (evaluate)", 2: "k_fold (fold)",
             3: "k_pass<2,P,2> (two-round pass: P pending folds + evaluate + next-round coefficients)"}
    dom = max(range(4), key=lambda i: ms[i])
    gbs = [nbytes[i] / (ms[i] * 1e-3) / 1e9 if ms[i] > 0 else 0.0 for i in range(4)]
    alg = algorithmic_bytes(v)
    log(f"bench: {T * 1e3:.3f} ms/proof; " + "; ".join(
        f"{names[i].split(' ')[0]} {launches[i]} launches {ms[i]:.2f} ms {gbs[i]:.0f} GB/s" for i in range(4) if launches[i])
        + f"; final {proof.final_value}")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cv = args.cpu_v if args.cpu_v is not None else v
        log(f"bench: CPU baseline (oracle port of the reference IP prover), v={cv}, {threads} threads")
        rate, secs = cpu_baseline(cv, threads)
        cpu = {"value": rate, "unit": "evals/s", "cores": threads, "kind": "port",
               "sample": f"one full InnerProductProver proof at v={cv} ({secs:.2f} s): two parallel folds "
                         f"per round, serial evaluate, scalar Montgomery arithmetic"}

    out = {
        "metric": "Goldilocks sumcheck evals/s (InnerProductProver MSB, v=28)",
        "value": n / T,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": T * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"inner_product_msb_v{v}", "v": v, "field": "goldilocks", "tables": 2,
                   "transcript": "shake256-fs", "input": "uniform [0,p) counter-based, seeds 0x5EED0002/3"},
        # the reference algorithm's bytes (one fold per round, 32 N per table) over the proof time
        "gbps_reference_bytes": alg / T / 1e9,
        "roofline": {"bound": "hbm", "achieved": gbs[dom], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs[dom] / HBM_PEAK_GBS,
                     "traffic": recorded_traffic() if dom == 3 and v == 28 else None, "kernel": names[dom],
                     "launches_timed": int(launches[dom]), "kernel_ms_total": ms[dom],
                     "per_class_gbs": {names[i].split(" ")[0]: gbs[i] for i in range(4) if launches[i]}},
        "cpu_baseline": cpu,
    }
    prover.close()
    for p in ptrs:
        L.effsc_hip_device_free(dev, p)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
