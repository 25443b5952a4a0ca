"""Benchmark of the MI355X sumcheck provers on one GPU.

Default workload (the headline line): Goldilocks inner-product sumcheck
(InnerProductProver, MSB), v = 28, two 2^28-element tables uniform in [0, p)
from fixed seeds, all 28 rounds with the SHAKE256 Fiat-Shamir transcript.

A step is one full proof (v rounds + finalize + final value) over tables
already resident in HBM.  Metric: evals/s = 2^v / T_proof (the reference's
criterion convention, Throughput::Elements(2^v), benches/sumcheck.rs:28,80).

Other workloads (--workload):
  ml        MultilinearProver, v = 20
  ml2, ip2  MultilinearProver / InnerProductProver over F_p^2 (F64Ext2
            inputs), v = 26
  ml3, ip3  the same over F_p^3 (F64Ext3 inputs), v = 24
  matrix    the reference's criterion matrix (benches/sumcheck.rs):
            {multilinear, inner_product} x {F64, F64Ext2, F64Ext3} x
            {2^16, 2^20, 2^24} and fold/{F64, F64Ext2, F64Ext3}, each cell
            beside the oracle on the host
  product3  CoefficientProver, product of 3 multilinears (degree 3), v = 24,
            base-field inputs, challenges and folded tables in F_p^3
  eq, eq3   EqFactoredProver, v = 24, random point, base field / F_p^3
  lsb       InnerProductProverLSB, v = 28 (tables bit-reversed once into the
            MSB path at construction)
  gkr       GkrProver, k = 12: a 2^24-point layer, 24 rounds over four tables
  batch     512 inner-product instances, v ~ U[12, 24], batched vs one by one
  serve     served request mix, saturated and Poisson, learned admission
  stream    the v = 28 inner product from host files, resident and out of core

Prints one JSON line.  The default line also carries the batched (512
inner-product instances) and served (Poisson mix) results under "batched"
and "served" (--no-extras skips them).  Besides the contract fields it
carries
  roofline     - the dominant kernel class's achieved rate: its algorithmic
                 bytes (each input element read once, each output element
                 written once) or Goldilocks products over the summed
                 HIP-event durations of its launches in the timed region
                 (events recorded on the prover's stream around each launch);
  cpu_baseline - the oracle's restatement of the reference prover, timed on
                 this host's cores on a bounded sample: one warm-up and ten
                 timed proofs, median with p10 / p90, the host CPU model and
                 flags; for the inner product both the InnerProductProver
                 round structure and the reference's faster legacy fused
                 path, the faster one as the value.
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
W_FROM_RAW7 = 0xFFFFFFF800000001
SEED_A, SEED_B = 0x5EED0002, 0x5EED0003
SEED_ML = 0x5EED0001
SEEDS_P3 = (0x5EED0031, 0x5EED0032, 0x5EED0033)
SEED_EQ_P, SEED_EQ_W = 0x5EED0034, 0x5EED0035
SEEDS_GKR = (0x5EED0041, 0x5EED0042, 0x5EED0043)  # add, mult, witness
DOMSEP = b"effsc-hip/bench/ip"

CLASS_NAMES = ["k_fused (fold + evaluate)", "k_eval (evaluate)", "k_fold (fold)",
               "k_pass (two-round pass: pending folds + evaluate + next-round coefficients)"]

WORKLOADS = {
    "ip": dict(kind="ip", v=28, ext=1, input_ext=1, tables=2,
               metric="Goldilocks sumcheck evals/s (InnerProductProver MSB, v=28)"),
    "ml": dict(kind="ml", v=20, ext=1, input_ext=1, tables=1,
               metric="Goldilocks sumcheck evals/s (MultilinearProver MSB, v=20)"),
    "ml2": dict(kind="ml", v=26, ext=2, input_ext=2, tables=1,
                metric="Goldilocks sumcheck evals/s (MultilinearProver MSB over F_p^2, v=26)"),
    "ip2": dict(kind="ip", v=26, ext=2, input_ext=2, tables=2,
                metric="Goldilocks sumcheck evals/s (InnerProductProver MSB over F_p^2, v=26)"),
    "ml3": dict(kind="ml", v=24, ext=3, input_ext=3, tables=1,
                metric="Goldilocks sumcheck evals/s (MultilinearProver MSB over F_p^3, v=24)"),
    "ip3": dict(kind="ip", v=24, ext=3, input_ext=3, tables=2,
                metric="Goldilocks sumcheck evals/s (InnerProductProver MSB over F_p^3, v=24)"),
    "product3": dict(kind="product", v=24, ext=3, input_ext=1, tables=3,
                     metric="Goldilocks sumcheck evals/s (CoefficientProver product of 3, F_p^3 challenges, v=24)"),
    "eq": dict(kind="eq", v=24, ext=1, input_ext=1, tables=1,
               metric="Goldilocks sumcheck evals/s (EqFactoredProver, base field, v=24)"),
    "eq3": dict(kind="eq", v=24, ext=3, input_ext=3, tables=1,
                metric="Goldilocks sumcheck evals/s (EqFactoredProver, F_p^3, v=24)"),
    "lsb": dict(kind="ip_lsb", v=28, ext=1, input_ext=1, tables=2,
                metric="Goldilocks sumcheck evals/s (InnerProductProverLSB, v=28)"),
    "gkr": dict(kind="gkr", v=24, ext=1, input_ext=1, tables=4,
                metric="Goldilocks sumcheck evals/s (GkrProver, k=12: 2^24-point layer, 24 rounds)"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(v, tables=2, in_words=1, work_words=1):
    # per table: round 0 reads N input elements; round 1 reads N input
    # elements and writes N/2 work elements; round j >= 2 reads L_j and
    # writes L_j/2 work elements; finalize folds 2 -> 1
    n = 1 << v
    if v == 0:
        return 8 * tables * in_words
    b = n * in_words  # round 0
    b += n * in_words + (n // 2) * work_words  # round 1 (or finalize when v = 1)
    L = n // 2
    while L >= 2:
        b += L * work_words + (L // 2) * work_words
        L //= 2
    return 8 * tables * b


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


# ---------------------------------------------------------------- CPU side
CPU_SAMPLES = 10


def cpu_info():
    """Host CPU model, the SIMD flags the reference's backends key on, and
    the logical CPU count."""
    model, flags = None, []
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if model is None and line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                if not flags and line.startswith("flags"):
                    flags = line.split(":", 1)[1].split()
    except OSError:
        pass
    keep = [x for x in ("avx2", "bmi2", "adx", "avx512f", "avx512dq", "avx512vl", "avx512ifma") if x in flags]
    return {"model": model, "flags": keep, "logical_cpus": os.cpu_count()}


def _sampled(run, samples):
    """One warm-up, then `samples` timed runs; run() returns the seconds of
    its timed part (input set-up is outside it)."""
    import numpy as np
    run()
    t = np.array([run() for _ in range(samples)])
    return {"median_s": float(np.median(t)), "p10_s": float(np.percentile(t, 10)),
            "p90_s": float(np.percentile(t, 90)), "samples": samples}


def cpu_baseline(workload, v_sample, threads, samples=CPU_SAMPLES):
    """The oracle (a C restatement of the reference prover) timed on this
    host over the same counter-based uniform tables."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n = 1 << v_sample
    protocol = f"1 warm-up + {samples} timed proofs at v={v_sample}, median (p10, p90)"
    if workload == "ip":
        a, b = O.uniform(n, SEED_A), O.uniform(n, SEED_B)
        paths = {}
        for name, fused, desc in (
                ("legacy_fused_ip", True, "inner_product_sumcheck_partial with fused_fold_and_compute_polynomial "
                                          "(src/inner_product_sumcheck.rs:199-360): fold and evaluate in one parallel "
                                          "sweep per round"),
                ("inner_product_prover", False, "InnerProductProver::round (src/provers/inner_product.rs:61-113): two "
                                                "parallel folds, serial evaluate")):
            st = _sampled(lambda: O.bench_ip(a.copy(), b.copy(), v_sample, threads, fused=fused)[0], samples)
            paths[name] = dict(st, evals_per_s=n / st["median_s"], threads=threads, path=desc)
            log(f"bench: CPU {name}: {st['median_s'] * 1e3:.1f} ms median at v={v_sample}, {threads} threads")
        best = max(paths, key=lambda k: paths[k]["evals_per_s"])
        return {"value": paths[best]["evals_per_s"], "unit": "evals/s", "cores": threads, "kind": "port",
                "sample": f"{protocol}; best reference CPU path: {best}, scalar Montgomery arithmetic",
                "best_path": best, "paths": paths, "cpu": cpu_info()}

    if workload == "ml":
        x = O.uniform(n, SEED_ML)
        paths = {}
        for name, th, desc in (
                ("multilinear_prover_parallel", threads, "MultilinearProver with the parallel feature "
                 "(src/provers/multilinear.rs:38-58 over src/multilinear_sumcheck.rs:84-254): rayon-split "
                 "evaluate, fused fold + evaluate, fold"),
                ("multilinear_prover_serial", 1, "the same prover on one thread")):
            st = _sampled(lambda: O.bench_ml(x.copy(), v_sample, th)[0], samples)
            paths[name] = dict(st, evals_per_s=n / st["median_s"], threads=th, path=desc)
            log(f"bench: CPU {name}: {st['median_s'] * 1e3:.1f} ms median at v={v_sample}, {th} threads")
        best = max(paths, key=lambda k: paths[k]["evals_per_s"])
        return {"value": paths[best]["evals_per_s"], "unit": "evals/s", "cores": paths[best]["threads"],
                "kind": "port", "sample": f"{protocol}; best reference CPU path: {best}, scalar Montgomery arithmetic",
                "best_path": best, "paths": paths, "cpu": cpu_info()}

    def make():
        if workload == "lsb":
            return O.Prover.inner_product_lsb(O.uniform(n, SEED_A), O.uniform(n, SEED_B))
        if workload == "gkr":
            m = 1 << (v_sample // 2)
            return O.Prover.gkr(O.uniform(m * m, SEEDS_GKR[0]), O.uniform(m * m, SEEDS_GKR[1]),
                                O.uniform(m, SEEDS_GKR[2]))
        if workload in ("ml", "ml2", "ml3"):
            e = WORKLOADS[workload]["ext"]
            return O.Prover.multilinear(O.uniform(n * e, SEED_ML), ext=e)
        if workload in ("ip2", "ip3"):
            e = WORKLOADS[workload]["ext"]
            return O.Prover.inner_product(O.uniform(n * e, SEED_A), O.uniform(n * e, SEED_B), ext=e)
        if workload == "product3":
            return O.Prover.product([O.uniform(n, s) for s in SEEDS_P3], ext=3)
        e = 3 if workload == "eq3" else 1
        return O.Prover.eq_factored(O.uniform(v_sample * e, SEED_EQ_W), O.uniform(n * e, SEED_EQ_P), ext=e)

    def run():
        p = make()
        t0 = time.perf_counter()
        p.sumcheck(v_sample, O.Transcript.shake(b"effsc-hip/bench/cpu"))
        return time.perf_counter() - t0
    st = _sampled(run, samples)
    log(f"bench: CPU oracle prover: {st['median_s'] * 1e3:.1f} ms median at v={v_sample}, 1 thread")
    return {"value": n / st["median_s"], "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": f"{protocol}; serial scalar Montgomery restatement of the reference prover",
            **st, "cpu": cpu_info()}


# ---------------------------------------------------------------- GPU side
def device_tables(L, N, dev, spec):
    """Allocates and fills the workload's input tables on the device."""
    v, ext, kind = spec["v"], spec["ext"], spec["kind"]
    n = 1 << v
    made = []

    def table(words, seed):
        p = N.u64p()
        N.check(L.effsc_hip_device_alloc(dev, words, C.byref(p)))
        N.check(L.effsc_hip_fill_uniform(dev, p, words, seed, 0))
        made.append(p)
        return C.cast(p, C.c_void_p).value

    if kind in ("ip", "ip_lsb"):
        # ext planes of n words each (the borrowed-input layout)
        addrs = [table(n * ext, SEED_A), table(n * ext, SEED_B)]
    elif kind == "gkr":
        m = 1 << (v // 2)
        addrs = [table(n, SEEDS_GKR[0]), table(n, SEEDS_GKR[1]), table(m, SEEDS_GKR[2])]
    elif kind == "ml":
        addrs = [table(n * ext, SEED_ML)]
    elif kind == "product":
        addrs = [table(n, s) for s in SEEDS_P3]
    else:
        addrs = [table(n * ext, SEED_EQ_P)]  # ext planes of n words
    return addrs, made


def eq_point(L, N, dev, v, ext):
    import numpy as np
    words = v * ext
    p = N.u64p()
    N.check(L.effsc_hip_device_alloc(dev, words, C.byref(p)))
    N.check(L.effsc_hip_fill_uniform(dev, p, words, SEED_EQ_W, 0))
    host = np.zeros(words, dtype=np.uint64)
    N.check(L.effsc_hip_copy(dev, host.ctypes.data_as(N.u64p), p, words))
    N.check(L.effsc_hip_device_free(dev, p))
    return host


def make_prover(effsc, L, N, dev, spec, addrs):
    v, n, ext = spec["v"], 1 << spec["v"], spec["ext"]
    kind = spec["kind"]
    if kind == "ip":
        return effsc.InnerProductProver(addrs[0], addrs[1], ext=ext, input_ext=ext, borrow=True, n=n, device=dev)
    if kind == "ip_lsb":
        # the prover bit-reverses the tables into its own planes once
        return effsc.InnerProductProverLSB(addrs[0], addrs[1], borrow=True, n=n, device=dev)
    if kind == "gkr":
        # GkrProver takes host tables (it builds the broadcast witness tables)
        import numpy as np
        m = 1 << (v // 2)
        host = []
        for a, cnt in zip(addrs, (n, n, m)):
            h = np.zeros(cnt, dtype=np.uint64)
            N.check(L.effsc_hip_copy(dev, h.ctypes.data_as(N.u64p), C.cast(a, N.u64p), cnt))
            host.append(h)
        return effsc.GkrProver(host[0], host[1], host[2], device=dev)
    if kind == "ml":
        return effsc.MultilinearProver(addrs[0], ext=ext, input_ext=ext, borrow=True, n=n, device=dev)
    if kind == "product":
        return effsc.CoefficientProver(addrs, ext=ext, input_ext=1, borrow=True, n=n, device=dev)
    w = eq_point(L, N, dev, v, ext)  # v * ext canonical words
    return effsc.EqFactoredProver(w, addrs[0], ext=ext, borrow=True, device=dev)


def proof_bytes(spec):
    v, kind, ext = spec["v"], spec["kind"], spec["ext"]
    if kind == "gkr":
        # four tables of 2^{2k} elements, folded every round (the witness
        # broadcasts included, as the reference materialises them)
        return algorithmic_bytes(v, tables=4, in_words=1, work_words=1)
    if kind == "product":
        return algorithmic_bytes(v, tables=3, in_words=1, work_words=3)
    return algorithmic_bytes(v, tables=spec["tables"], in_words=ext, work_words=ext)


def bench_single(args, workload):
    import effsc
    from effsc import _native as N

    spec = dict(WORKLOADS[workload])
    if args.v is not None:
        spec["v"] = args.v
    L = N.lib()
    v, ext = spec["v"], spec["ext"]
    n = 1 << v
    dev = 0
    addrs, made = device_tables(L, N, dev, spec)
    if args.calibrate:
        cw = min(n, 1 << 27)
        tmp = N.u64p()
        N.check(L.effsc_hip_device_alloc(dev, cw, C.byref(tmp)))
        N.check(L.effsc_hip_copy(dev, tmp, made[0], cw))
        N.check(L.effsc_hip_device_free(dev, tmp))
        log(f"bench: calibration copy of {cw * 8} bytes done")
    prover = make_prover(effsc, L, N, dev, spec, addrs)
    log(f"bench: {workload} v={v} ext={ext}, tables resident "
        f"({sum(1 for _ in made)} x {n * (ext if spec['kind'] in ('eq', 'ml', 'ip') else 1) * 8 / 2**20:.0f} MiB)")

    def step():
        prover.reset()
        return effsc.sumcheck(prover, v, effsc.Transcript(DOMSEP, ext=ext))

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
    muls = (C.c_double * 4)()
    N.check(L.effsc_hip_profile_read(prover.h, launches, ms, nbytes))
    N.check(L.effsc_hip_profile_read_muls(prover.h, muls))
    N.check(L.effsc_hip_profile(prover.h, 0))
    # the profiled region adds event records; time an unprofiled region too
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    N.check(L.effsc_hip_synchronize(dev))
    elapsed_plain = time.perf_counter() - t1
    T = min(elapsed, elapsed_plain) / args.steps

    gbs = [nbytes[i] / (ms[i] * 1e-3) / 1e9 if ms[i] > 0 else 0.0 for i in range(4)]
    gmuls = [muls[i] / (ms[i] * 1e-3) / 1e9 if ms[i] > 0 else 0.0 for i in range(4)]
    alg = proof_bytes(spec)
    log(f"bench: {T * 1e3:.3f} ms/proof; " + "; ".join(
        f"{CLASS_NAMES[i].split(' ')[0]} {launches[i]} launches {ms[i]:.2f} ms {gbs[i]:.0f} GB/s "
        f"{gmuls[i]:.0f} Gmul/s" for i in range(4) if launches[i]) + f"; final {proof.final_value}")

    valu = spec["kind"] == "product"
    profiled = any(launches[i] for i in range(4))
    names = list(CLASS_NAMES)
    if spec["kind"] == "eq":
        names[0] = "k_eq_round (fold p with the challenge + evaluate with eq weights)"
        names[1] = "k_eq_round (round 0: evaluate with eq weights)"
    elif spec["kind"] == "product" and ext == 3:
        names[0] = "k_fused_base3 (round 1) and k_fused_seq (fold + evaluate, one F_p^3 table at a time)"
    elif spec["kind"] == "gkr":
        names[0] = "k_gkr_fused_fw (fold add and mult, factored witness, evaluate the wire)"
        names[1] = "k_gkr_eval_fw (round 0: evaluate the wire, factored witness)"
    if profiled:
        dom = max(range(4), key=lambda i: ms[i])
        kernel = names[dom]
    if valu:
        peak = C.c_double()
        N.check(L.effsc_hip_mul_rate(dev, C.byref(peak)))
        peak_g = peak.value / 1e9
        roof = {"bound": "valu", "achieved": gmuls[dom], "peak": peak_g, "unit": "Gmul/s",
                "frac": gmuls[dom] / peak_g, "traffic": None, "kernel": kernel,
                "peak_source": "effsc_hip_mul_rate: independent Goldilocks product chains on every CU",
                "hbm_achieved_gbs": gbs[dom], "hbm_frac": gbs[dom] / HBM_PEAK_GBS,
                "launches_timed": int(launches[dom]), "kernel_ms_total": ms[dom]}
    elif profiled:
        roof = {"bound": "hbm", "achieved": gbs[dom], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs[dom] / HBM_PEAK_GBS,
                "traffic": recorded_traffic() if (workload == "ip" and dom == 3 and v == 28) else None,
                "kernel": kernel, "launches_timed": int(launches[dom]), "kernel_ms_total": ms[dom]}
    else:
        # no multi-workgroup round kernel is profiled on this path: whole-proof rate
        roof = {"bound": "hbm", "achieved": alg / T / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": alg / T / 1e9 / HBM_PEAK_GBS, "traffic": None,
                "kernel": "whole proof (algorithmic bytes / proof time; latency-bound at this size)"}
    roof["per_class"] = {names[i]: {"launches": int(launches[i]), "ms": ms[i], "gbs": gbs[i], "gmul_s": gmuls[i]}
                         for i in range(4) if launches[i]}
    if roof["bound"] == "hbm":
        # the measured side of the HBM roofline: kernels with no arithmetic
        # moving the same bytes (a stream copy; for the two-round passes also
        # their read-pass and fold-pass access patterns) at this size
        pats = [("copy_gbs", 0)]
        if spec["kind"] in ("ip", "ip_lsb", "ml"):
            pats += [("read_pass_gbs", 1), ("fold_pass_gbs", 2)]
        att = {}
        for name, pat in pats:
            r = C.c_double()
            N.check(L.effsc_hip_hbm_rate(dev, pat, 1 << max(v, 10), C.byref(r)))
            att[name] = r.value / 1e9
        att["frac_of_copy"] = roof["achieved"] / att["copy_gbs"]
        att["source"] = (f"effsc_hip_hbm_rate: no-arithmetic kernels, 2^{max(v, 10)} words per table, "
                         "best of 5 launches")
        roof["attainable"] = att
        log("bench: attainable " + ", ".join(f"{k} {att[k]:.0f}" for k, _ in pats))

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        # bounded samples: 11 proofs of about a second or less each
        default_cpu_v = {"ip": min(v, 26), "ml": min(v, 20), "ml2": min(v, 18), "ip2": min(v, 18),
                         "ml3": min(v, 18), "ip3": min(v, 18), "product3": min(v, 18), "eq": min(v, 20),
                         "eq3": min(v, 18), "lsb": min(v, 20), "gkr": min(v, 18)}[workload]
        cv = args.cpu_v if args.cpu_v is not None else default_cpu_v
        log(f"bench: CPU baseline (oracle port of the reference prover), v={cv}")
        cpu = cpu_baseline(workload, cv, threads)

    inputs = {"ip": "seeds 0x5EED0002/3", "ml": "seed 0x5EED0001", "product3": "seeds 0x5EED0031/32/33",
              "ml2": "seed 0x5EED0001 (F_p^2 words)", "ip2": "seeds 0x5EED0002/3 (F_p^2 words)",
              "ml3": "seed 0x5EED0001 (F_p^3 words)", "ip3": "seeds 0x5EED0002/3 (F_p^3 words)",
              "eq": "p seed 0x5EED0034, point seed 0x5EED0035", "eq3": "p seed 0x5EED0034, point seed 0x5EED0035",
              "lsb": "seeds 0x5EED0002/3 (LSB order)", "gkr": "add, mult, witness seeds 0x5EED0041/42/43"}
    out = {
        "metric": spec["metric"] if v == WORKLOADS[workload]["v"] else spec["metric"].rsplit("v=", 1)[0] + f"v={v})",
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
        "config": {"workload": f"{workload}_v{v}" if workload in ("lsb", "gkr") else f"{workload}_msb_v{v}", "v": v, "field": {1: "goldilocks", 2: "goldilocks^2", 3: "goldilocks^3"}[ext],
                   "tables": spec["tables"], "transcript": "shake256-fs",
                   "input": f"uniform [0,p) counter-based, {inputs[workload]}"},
        # the reference algorithm's bytes (one fold per round) over the proof time
        "gbps_reference_bytes": alg / T / 1e9,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    prover.close()
    for p in made:
        L.effsc_hip_device_free(dev, p)
    return out


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
    tile, tile_ms = bp.autotune()
    log(f"bench: autotuned tile {tile} units per workgroup; " +
        ", ".join(f"{c}: {t:.2f} ms" for c, t in tile_ms.items()))

    def ts():
        return [effsc.Transcript(b"effsc-hip/batch/%d" % i) for i in range(len(vs))]

    # timed at the C ABI: transcripts made and output buffers allocated
    # beforehand, so the figure is the library's (H2D of the sponges, the
    # graph, D2H of every proof word, synchronisation)
    import numpy as np
    polys, chals, fins = bp.buffers()
    for _ in range(args.warmup):
        ref = bp.prove(ts())
    sets = [ts() for _ in range(args.steps)]
    handles = [(C.c_void_p * len(vs))(*[t.h for t in tset]) for tset in sets]
    N.check(N.lib().effsc_hip_synchronize(0))
    t0 = time.perf_counter()
    for h in handles:
        bp.prove_raw(h, polys, chals, fins)
    T_batch = (time.perf_counter() - t0) / args.steps
    # the Python wrapper (proof objects built per instance) for comparison
    t0 = time.perf_counter()
    out = bp.prove(ts())
    T_wrapper = time.perf_counter() - t0
    assert all(p.round_polys == q.round_polys and p.final_value == q.final_value for p, q in zip(out, ref))
    # per-instance baseline, also at the C ABI: one device-resident proof per
    # instance (effsc_hip_sumcheck), the same inputs and outputs
    provers = [effsc.InnerProductProver(C.cast(x.ptr, C.c_void_p).value, C.cast(y.ptr, C.c_void_p).value,
                                        borrow=True, n=x.n)
               for x, y in zip(a, b)]
    L = N.lib()
    pbuf = [np.zeros(max(2 * v, 1), dtype=np.uint64) for v in vs]
    cbuf = [np.zeros(max(v, 1), dtype=np.uint64) for v in vs]
    fbuf = np.zeros(len(vs), dtype=np.uint64)
    nohook = C.cast(None, N.HOOK)

    def single_all(tset):
        for i, (p, v) in enumerate(zip(provers, vs)):
            N.check(L.effsc_hip_reset(p.h))
            N.check(L.effsc_hip_sumcheck(p.h, v, tset[i].h, nohook, None, pbuf[i].ctypes.data_as(N.u64p),
                                         cbuf[i].ctypes.data_as(N.u64p), fbuf[i:].ctypes.data_as(N.u64p)))
    single_all(ts())
    reps = max(1, args.steps // 2)
    sets = [ts() for _ in range(reps)]
    N.check(L.effsc_hip_synchronize(0))
    t1 = time.perf_counter()
    for tset in sets:
        single_all(tset)
    T_single = (time.perf_counter() - t1) / reps
    same = all(list(pbuf[i][:2 * v]) == [x for r in q.round_polys for x in r] and int(fbuf[i]) == q.final_value
               for i, (v, q) in enumerate(zip(vs, out)))
    alg = sum(algorithmic_bytes(v, tables=2) for v in vs)
    res = {"metric": "batched proofs/s (512 inner-product instances, v ~ U[12,24])",
                      "value": len(vs) / T_batch, "unit": "proofs/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": T_batch * 1e3, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
                      "config": {"workload": "batched_ip", "instances": len(vs), "vmin": args.vmin,
                                 "vmax": args.vmax, "table_bytes": 2 * total * 8, "seed": "0x5EED0004"},
                      "per_instance_launches_proofs_per_s": len(vs) / T_single,
                      "per_instance_ms": T_single * 1e3, "speedup": T_single / T_batch,
                      "python_wrapper_ms": T_wrapper * 1e3,
                      "bit_identical_to_per_instance": same,
                      "autotune": {"chosen_units_per_workgroup": tile,
                                   "ms_per_batch_by_tile": {str(c): t for c, t in tile_ms.items()}},
                      "gbps_reference_bytes": alg / T_batch / 1e9}
    for p in provers:
        p.close()
    bp.close()
    for t in a + b:
        t.free()
    return res


def bench_stream(args):
    # The v=28 inner-product proof with its tables on the host: two FileStream
    # files (serialized canonical words) mapped, moved to the device in chunks
    # through pinned buffers while the first two-round pass runs on the landed
    # chunks, then the device-resident proof.  The PCIe-inclusive rate, beside
    # the resident proof over the same tables.
    import tempfile

    import numpy as np

    import effsc
    from effsc import _native as N
    from effsc.streams import FileStream, inner_product_from_stream

    v = args.v if args.v is not None else 28
    n = 1 << v
    L = N.lib()
    tmpdir = tempfile.mkdtemp(prefix="effsc_stream_")
    paths = []
    for seed in (SEED_A, SEED_B):
        p = N.u64p()
        N.check(L.effsc_hip_device_alloc(0, n, C.byref(p)))
        N.check(L.effsc_hip_fill_uniform(0, p, n, seed, 0))
        host = np.zeros(n, dtype=np.uint64)
        N.check(L.effsc_hip_copy(0, host.ctypes.data_as(N.u64p), p, n))
        N.check(L.effsc_hip_device_free(0, p))
        path = os.path.join(tmpdir, "t%x.bin" % seed)
        FileStream.write_to_file(path, host)
        paths.append(path)
        del host
    log(f"bench: stream files written ({2 * n * 8 / 2**30:.1f} GiB) under {tmpdir}")
    streams = [FileStream(p) for p in paths]
    # warm the page cache and the code paths
    pr = inner_product_from_stream(*streams)
    ref = effsc.sumcheck(pr, v, effsc.Transcript(DOMSEP))
    pr.close()
    times_up, times_total = [], []
    for _ in range(args.steps):
        N.check(L.effsc_hip_synchronize(0))
        t0 = time.perf_counter()
        pr = inner_product_from_stream(*streams)
        t1 = time.perf_counter()
        proof = effsc.sumcheck(pr, v, effsc.Transcript(DOMSEP))
        t2 = time.perf_counter()
        times_up.append(t1 - t0)
        times_total.append(t2 - t0)
        pr.close()
    assert proof.round_polys == ref.round_polys and proof.final_value == ref.final_value
    # the same proof over the same tables already resident
    pr = inner_product_from_stream(*streams)
    pr.reset()
    effsc.sumcheck(pr, v, effsc.Transcript(DOMSEP))
    N.check(L.effsc_hip_synchronize(0))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pr.reset()
        res = effsc.sumcheck(pr, v, effsc.Transcript(DOMSEP))
    T_res = (time.perf_counter() - t0) / args.steps
    assert res.round_polys == ref.round_polys
    pr.close()
    # out of core: the tables stay in the files; construction streams them
    # for rounds 0-1 and the proof streams them again for rounds 2-3, with
    # only the level-n/4 tables resident
    from effsc.streams import OUT_OF_CORE
    times_ooc = []
    for _ in range(args.steps):
        N.check(L.effsc_hip_synchronize(0))
        t0 = time.perf_counter()
        pr = inner_product_from_stream(*streams, mode=OUT_OF_CORE)
        ooc = effsc.sumcheck(pr, v, effsc.Transcript(DOMSEP))
        times_ooc.append(time.perf_counter() - t0)
        pr.close()
        assert ooc.round_polys == ref.round_polys and ooc.final_value == ref.final_value
    T_ooc = min(times_ooc)
    for s_ in streams:
        s_.close()
    for p in paths:
        os.remove(p)
    os.rmdir(tmpdir)
    T_up, T_tot = min(times_up), min(times_total)
    return {
        "metric": "streamed Goldilocks sumcheck evals/s (InnerProductProver MSB, v=%d, tables in host files)" % v,
        "value": n / T_tot, "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": T_tot * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"stream_ip_v{v}", "v": v, "input": "two FileStream files, page-cached",
                   "seeds": "0x5EED0002/3"},
        "upload_and_first_pass_ms": T_up * 1e3, "host_to_device_gbps": 2 * n * 8 / T_up / 1e9,
        # repeated proofs of the same stream prover reuse the first pass
        # accumulated during the upload (the input never changes)
        "resident_repeat_ms_per_proof_first_pass_reused": T_res * 1e3, "bit_identical_to_resident": True,
        "out_of_core": {"ms_per_proof": T_ooc * 1e3, "evals_per_s": n / T_ooc,
                        "host_bytes_streamed_per_proof": 2 * 2 * n * 8, "resident_device_bytes": 2 * (n // 4) * 8,
                        "bit_identical_to_resident": True}}


def bench_matrix(args):
    # The reference's criterion matrix (benches/sumcheck.rs:11-190): full
    # MultilinearProver / InnerProductProver proofs with the test transcript
    # (SanityTranscript there, TestTranscript here) and free folds, over F64,
    # F64Ext2 and F64Ext3 at 2^16, 2^20 and 2^24 elements; Throughput::
    # Elements(2^v).  GPU side: inputs resident in HBM (the reference builds
    # its Vec inside the timed routine but generates it in the untimed
    # set-up); the median of `steps` proofs after `warmup`.  CPU side: the
    # oracle's serial restatement of the same prover, 1 warm-up + 3 samples
    # at min(v, 18).
    import numpy as np

    import effsc
    from effsc import _native as N
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    L = N.lib()
    dev = 0
    steps, warmup = max(args.steps, 3), max(args.warmup, 1)
    cells = []

    def dev_table(words, seed):
        p = N.u64p()
        N.check(L.effsc_hip_device_alloc(dev, words, C.byref(p)))
        N.check(L.effsc_hip_fill_uniform(dev, p, words, seed, 0))
        return p

    def cpu_cell(run_factory, v):
        cv = min(v, 18)
        st = _sampled(run_factory(cv), 3)
        return {"v": cv, "evals_per_s": (1 << cv) / st["median_s"], "median_s": st["median_s"],
                "p10_s": st["p10_s"], "p90_s": st["p90_s"], "cores": 1}

    fields = {1: "F64", 2: "F64Ext2", 3: "F64Ext3"}
    for kind in ("multilinear", "inner_product"):
        for ext in (1, 2, 3):
            for v in (16, 20, 24):
                n = 1 << v
                k = 1 if kind == "multilinear" else 2
                tabs = [dev_table(n * ext, SEED_A + i) for i in range(k)]
                addrs = [C.cast(t, C.c_void_p).value for t in tabs]
                if k == 1:
                    pr = effsc.MultilinearProver(addrs[0], ext=ext, input_ext=ext, borrow=True, n=n)
                else:
                    pr = effsc.InnerProductProver(addrs[0], addrs[1], ext=ext, input_ext=ext, borrow=True, n=n)

                def proof():
                    pr.reset()
                    return effsc.sumcheck(pr, v, effsc.TestTranscript(7, ext=ext))
                for _ in range(warmup):
                    proof()
                ts = []
                for _ in range(steps):
                    N.check(L.effsc_hip_synchronize(dev))
                    t0 = time.perf_counter()
                    proof()
                    ts.append(time.perf_counter() - t0)
                T = float(np.median(ts))
                pr.close()
                for t in tabs:
                    L.effsc_hip_device_free(dev, t)

                def factory(cv, kind=kind, ext=ext):
                    m = 1 << cv
                    tb = [O.uniform(m * ext, SEED_A + i) for i in range(1 if kind == "multilinear" else 2)]

                    def run():
                        p = (O.Prover.multilinear(tb[0], ext=ext) if kind == "multilinear"
                             else O.Prover.inner_product(tb[0], tb[1], ext=ext))
                        t0 = time.perf_counter()
                        p.sumcheck(cv, O.Transcript.rng(7))
                        return time.perf_counter() - t0
                    return run
                cpu = None if args.no_cpu else cpu_cell(factory, v)
                alg = algorithmic_bytes(v, tables=k, in_words=ext, work_words=ext)
                cell = {"group": f"{kind}/{fields[ext]}", "size": f"2^{v}", "gpu_ms": T * 1e3,
                        "gpu_evals_per_s": n / T, "gpu_gbs_reference_bytes": alg / T / 1e9,
                        "hbm_frac": alg / T / 1e9 / HBM_PEAK_GBS, "cpu": cpu}
                if cpu:
                    cell["gpu_over_cpu"] = cell["gpu_evals_per_s"] / cpu["evals_per_s"]
                log(f"bench: matrix {cell['group']} {cell['size']}: {T * 1e3:.3f} ms, "
                    f"{n / T / 1e9:.2f} G evals/s" + (f", CPU {cpu['evals_per_s'] / 1e6:.1f} M evals/s" if cpu else ""))
                cells.append(cell)
    for ext in (1, 2, 3):
        for v in (16, 20, 24):
            n = 1 << v
            src = dev_table(n * ext, SEED_ML)
            work = N.u64p()
            N.check(L.effsc_hip_device_alloc(dev, n * ext, C.byref(work)))
            wt = np.array([(0x1234567 + 99 * i) % O.P for i in range(ext)], dtype=np.uint64)
            newlen = C.c_size_t()

            def fold_once():
                N.check(L.effsc_hip_copy(dev, work, src, n * ext))
                N.check(L.effsc_hip_synchronize(dev))
                t0 = time.perf_counter()
                if ext == 1:
                    N.check(L.effsc_hip_fold(dev, work, n, int(wt[0]), C.byref(newlen)))
                else:
                    N.check(L.effsc_hip_fold_ext(dev, ext, W_FROM_RAW7, work, n, wt.ctypes.data_as(N.u64p),
                                                 C.byref(newlen)))
                return time.perf_counter() - t0
            for _ in range(warmup):
                fold_once()
            T = float(np.median([fold_once() for _ in range(steps)]))
            L.effsc_hip_device_free(dev, src)
            L.effsc_hip_device_free(dev, work)

            def factory(cv, ext=ext, wt=wt):
                m = 1 << cv
                tb = O.uniform(m * ext, SEED_ML)

                def run():
                    p = O.Prover.multilinear(tb, ext=ext)
                    t0 = time.perf_counter()
                    p.finalize([int(x) for x in wt])  # fold(&mut evals, w)
                    return time.perf_counter() - t0
                return run
            cpu = None if args.no_cpu else cpu_cell(factory, v)
            alg = 8 * ext * (n + n // 2)
            cell = {"group": f"fold/{fields[ext]}", "size": f"2^{v}", "gpu_ms": T * 1e3, "gpu_evals_per_s": n / T,
                    "gpu_gbs_reference_bytes": alg / T / 1e9, "hbm_frac": alg / T / 1e9 / HBM_PEAK_GBS, "cpu": cpu,
                    "timing": "one call including its launch and device synchronisation"}
            if cpu:
                cell["gpu_over_cpu"] = cell["gpu_evals_per_s"] / cpu["evals_per_s"]
            log(f"bench: matrix {cell['group']} {cell['size']}: {T * 1e3:.3f} ms")
            cells.append(cell)
    head = next(c for c in cells if c["group"] == "inner_product/F64" and c["size"] == "2^24")
    return {"metric": "Goldilocks sumcheck evals/s, reference criterion matrix (headline cell inner_product/F64 2^24)",
            "value": head["gpu_evals_per_s"], "unit": "evals/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": head["gpu_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u64", "data": "synthetic",
            "config": {"workload": "reference_bench_matrix", "transcript": "test RNG transcript",
                       "sizes": [16, 20, 24], "fields": ["F64", "F64Ext2", "F64Ext3"]},
            "cells": cells, "cpu": cpu_info()}


def serve_mix(count, seed):
    """cfg5 request mix: 40 % multilinear, 40 % inner product, 20 % product of
    3 over F_p^3; v log-uniform in [10, 26] (ln v ~ U[ln 10, ln 26], rounded)."""
    import math
    import random

    from effsc.serve import INNER_PRODUCT, MULTILINEAR, PRODUCT3_EXT3
    rng = random.Random(seed)
    out = []
    for i in range(count):
        u = rng.random()
        kind = MULTILINEAR if u < 0.4 else INNER_PRODUCT if u < 0.8 else PRODUCT3_EXT3
        v = int(round(math.exp(rng.uniform(math.log(10), math.log(26)))))
        out.append((kind, v, [rng.getrandbits(48) for _ in range(3)]))
    return out


def bench_serve(args):
    # Served proofs at a fixed offered load: requests arrive as a Poisson
    # process (rate = --load x the measured saturated proofs/s) and join the
    # running batch as slots and HBM arena bytes free up.  Tables are made on
    # the device beforehand (resident inputs, borrowed by both sides); the
    # baseline proves the same requests one by one with per-request provers.
    import random
    import threading

    import numpy as np

    import effsc
    from effsc import _native as N
    from effsc.batch import DeviceTable
    from effsc.serve import INNER_PRODUCT, MULTILINEAR, PRODUCT3_EXT3, Server

    K = {MULTILINEAR: 1, INNER_PRODUCT: 2, PRODUCT3_EXT3: 3}
    EXT = {MULTILINEAR: 1, INNER_PRODUCT: 1, PRODUCT3_EXT3: 3}
    reqs = serve_mix(args.count, 0x5EED0005)
    tabs = [[DeviceTable.uniform(1 << v, s) for s in seeds[:K[kind]]] for kind, v, seeds in reqs]
    addrs = [[C.cast(t.ptr, C.c_void_p).value for t in ts] for ts in tabs]
    in_bytes = sum(t.n * 8 for ts in tabs for t in ts)
    log(f"bench: serve {len(reqs)} requests (40/40/20 ML/IP/P3-ext3, v log-uniform in [10,26]), "
        f"{in_bytes / 2**30:.1f} GiB of resident tables, budget {args.budget_gib} GiB, max_active {args.max_active}")

    def ts(i):
        return effsc.Transcript(b"effsc-hip/serve/%d" % i, ext=EXT[reqs[i][0]])

    srv = Server(hbm_budget=int(args.budget_gib * 2**30), max_active=args.max_active)
    # warm-up: a few requests of every kind
    for i in range(min(6, len(reqs))):
        srv.wait(srv.submit(reqs[i][0], reqs[i][1], ts(i), tables=addrs[i]))
    # 1. saturation: every request submitted at once
    t0 = srv.now()
    ids = [srv.submit(k, v, ts(i), tables=addrs[i]) for i, (k, v, _) in enumerate(reqs)]
    sat = [srv.wait(r) for r in ids]
    T_sat = max(r.completed for r in sat) - t0
    sat_rate = len(reqs) / T_sat
    log(f"bench: saturated {sat_rate:.0f} proofs/s ({T_sat * 1e3:.1f} ms for {len(reqs)})")
    # 2. Poisson arrivals at the offered load
    lam = args.load * sat_rate
    rng = random.Random(0x5EED0006)
    gaps = [rng.expovariate(lam) for _ in reqs]
    arrivals = np.cumsum(gaps)
    order = list(range(len(reqs)))
    rng.shuffle(order)
    ids = [None] * len(reqs)

    def feeder():
        start = srv.now()
        for slot, i in enumerate(order):
            target = start + arrivals[slot]
            while True:
                d = target - srv.now()
                if d <= 0:
                    break
                time.sleep(min(d, 0.0005) if d > 0.0002 else 0)
            k, v, _ = reqs[i]
            ids[i] = srv.submit(k, v, ts(i), tables=addrs[i])

    th = threading.Thread(target=feeder)
    t1 = srv.now()
    th.start()
    th.join()
    served = [srv.wait(r) for r in ids]
    T_poi = max(r.completed for r in served) - t1
    lat = np.array([r.latency for r in served]) * 1e3
    by_kind = {}
    for (k, v, _), r in zip(reqs, served):
        by_kind.setdefault(k, []).append(r.latency * 1e3)
    st = srv.stats()
    srv.close()
    same_as_sat = all(a.proof == b.proof for a, b in zip(served, sat))
    # 2b. learned admission: a service-time network trained on the GPU from
    # requests measured alone; a server limited to --learn-slots requests at
    # once admits a burst of every request shortest-predicted first, against
    # first come first served
    from effsc.learn import ServiceTimeModel, measure_service_times, work_priority
    grid = [(k, v) for k in (MULTILINEAR, INNER_PRODUCT, PRODUCT3_EXT3) for v in range(10, 27)]
    cal = Server(hbm_budget=int(args.budget_gib * 2**30), max_active=1)
    samples = measure_service_times(cal, grid, lambda i, k: effsc.Transcript(b"cal", ext=EXT[k]))
    cal.close()
    model = ServiceTimeModel(hidden=32, seed=3)
    t_fit = time.perf_counter()
    fit_loss = model.fit(samples, steps=3000, lr=1e-2)
    t_fit = time.perf_counter() - t_fit
    prio = [model.priority(k, v) for k, v, _ in reqs]
    # the non-learned ordering on the same features: shortest work first,
    # work = 2^v x tables x words per element
    heur = [work_priority(k, v) for k, v, _ in reqs]
    lim = Server(hbm_budget=int(args.budget_gib * 2**30), max_active=args.learn_slots)

    def burst(prios):
        s0 = lim.now()
        rs = [lim.submit(k, v, ts(i), tables=addrs[i], priority=pr) for i, ((k, v, _), pr) in enumerate(zip(reqs, prios))]
        out = [lim.wait(r) for r in rs]
        return out, max(r.completed for r in out) - s0
    burst([0.0] * len(reqs))  # warm-up
    fifo, T_fifo = burst([0.0] * len(reqs))
    lrn, T_lrn = burst(prio)
    swf, T_swf = burst(heur)
    lim.close()

    def lat_stats(rs):
        x = np.array([r.latency for r in rs]) * 1e3
        return {"mean_ms": float(x.mean()), "p50_ms": float(np.percentile(x, 50)), "p99_ms": float(np.percentile(x, 99))}
    learned = {"slots": args.learn_slots, "calibration_requests": len(samples), "fit_loss_log_ms": fit_loss,
               "train_s": t_fit, "fifo": dict(lat_stats(fifo), proofs_per_s=len(reqs) / T_fifo),
               "learned": dict(lat_stats(lrn), proofs_per_s=len(reqs) / T_lrn),
               "heuristic_shortest_work_first": dict(lat_stats(swf), proofs_per_s=len(reqs) / T_swf),
               "identical_proofs": all(a.proof == b.proof == c.proof for a, b, c in zip(fifo, lrn, swf))}
    log(f"bench: admission mean latency: learned {learned['learned']['mean_ms']:.2f} ms, shortest work first "
        f"{learned['heuristic_shortest_work_first']['mean_ms']:.2f} ms, FIFO {learned['fifo']['mean_ms']:.2f} ms")
    # 3. baseline: the same requests one by one, per-request device-resident proofs
    provers = []
    for (k, v, _), a in zip(reqs, addrs):
        n = 1 << v
        if k == MULTILINEAR:
            provers.append(effsc.MultilinearProver(a[0], borrow=True, n=n))
        elif k == INNER_PRODUCT:
            provers.append(effsc.InnerProductProver(a[0], a[1], borrow=True, n=n))
        else:
            provers.append(effsc.CoefficientProver(a, ext=3, input_ext=1, borrow=True, n=n))

    def one_by_one():
        out = []
        for i, (p, (k, v, _)) in enumerate(zip(provers, reqs)):
            p.reset()
            out.append(effsc.sumcheck(p, v, ts(i)))
        return out
    one_by_one()
    t2 = time.perf_counter()
    single = one_by_one()
    T_single = time.perf_counter() - t2
    for p in provers:
        p.close()

    def norm(x):
        return x if isinstance(x, int) else tuple(x)
    identical = all(
        [[norm(y) for y in r] for r in q.round_polys] == [[norm(y) for y in r] for r in s.proof.round_polys]
        and norm(q.final_value) == norm(s.proof.final_value) for q, s in zip(single, served))
    alg = sum(algorithmic_bytes(v, tables=K[k], in_words=1, work_words=EXT[k]) for k, v, _ in reqs)
    res = {
        "metric": "served proofs/s and p50/p99 latency (Poisson arrivals, 40% ML / 40% IP / 20% P3-ext3, "
                  "v log-uniform [10,26])",
        "value": sat_rate, "unit": "proofs/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": T_sat * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": "served_mix", "requests": len(reqs), "budget_gib": args.budget_gib,
                   "max_active": args.max_active, "offered_load": args.load, "seed": "0x5EED0005",
                   "v_sampler": "round(exp(U(ln 10, ln 26)))", "inputs": "resident device tables, borrowed"},
        "saturated_proofs_per_s": sat_rate,
        "poisson": {"rate_offered": lam, "proofs_per_s": len(reqs) / T_poi, "p50_ms": float(np.percentile(lat, 50)),
                    "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                    "p50_ms_by_kind": {str(k): float(np.percentile(x, 50)) for k, x in by_kind.items()},
                    "p99_ms_by_kind": {str(k): float(np.percentile(x, 99)) for k, x in by_kind.items()}},
        "per_instance_launches_proofs_per_s": len(reqs) / T_single, "speedup_vs_per_instance": T_single / T_sat,
        "bit_identical_to_per_instance": identical, "poisson_equals_saturated": same_as_sat,
        "server": st, "gbps_reference_bytes": alg / T_sat / 1e9, "learned_admission": learned}
    for tsl in tabs:
        for t in tsl:
            t.free()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--v", type=int, default=None, help="number of variables (default: the workload's)")
    ap.add_argument("--cpu-v", type=int, default=None, help="CPU baseline sample size")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="ip: skip the batched and served parts of the line")
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["batch", "serve", "stream", "matrix"], default="ip",
                    help="ip: the v=28 inner-product proof (default, the headline line)")
    ap.add_argument("--count", type=int, default=512, help="batch instances / served requests")
    ap.add_argument("--load", type=float, default=0.7, help="serve: offered load, fraction of saturated proofs/s")
    ap.add_argument("--budget-gib", type=float, default=16.0, help="serve: HBM arena of the server")
    ap.add_argument("--max-active", type=int, default=256, help="serve: requests on the device at once")
    ap.add_argument("--learn-slots", type=int, default=16,
                    help="serve: max_active of the server comparing learned and FIFO admission")
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
        out = bench_batch(args)
    elif args.workload == "serve":
        out = bench_serve(args)
    elif args.workload == "stream":
        out = bench_stream(args)
    elif args.workload == "matrix":
        out = bench_matrix(args)
    else:
        out = bench_single(args, args.workload)
        if args.workload == "ip" and out["config"]["v"] == 28 and not args.no_extras:
            # the rest of the headline metric: batched proofs/s (512
            # instances, autotuned tiles) and served p50/p99 + proofs/s
            out["batched"] = summarize_batch(bench_batch(extras_args(args)))
            out["served"] = summarize_serve(bench_serve(extras_args(args)))
    print(json.dumps(out), flush=True)


def extras_args(args):
    a = argparse.Namespace(**vars(args))
    a.steps, a.warmup, a.count = 5, 1, 512
    return a


def summarize_batch(r):
    keep = ("metric", "value", "unit", "ms_per_step", "steps", "config", "per_instance_launches_proofs_per_s",
            "per_instance_ms", "speedup", "bit_identical_to_per_instance", "autotune", "gbps_reference_bytes")
    return {k: r[k] for k in keep if k in r}


def summarize_serve(r):
    keep = ("metric", "value", "unit", "config", "saturated_proofs_per_s", "poisson",
            "per_instance_launches_proofs_per_s", "speedup_vs_per_instance", "bit_identical_to_per_instance",
            "poisson_equals_saturated", "learned_admission")
    return {k: r[k] for k in keep if k in r}


if __name__ == "__main__":
    main()
