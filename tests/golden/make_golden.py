"""Writes tests/golden/golden.json: known-answer vectors for the sumcheck path.

Every expected value here comes from oracle/brute.py (pure-Python big-integer
arithmetic written from the definitions, no folding) and from hashlib's
SHAKE256; none comes from the C oracle or the HIP library, which the tests
check against these vectors.  Re-run with `python tests/golden/make_golden.py`.
"""
import hashlib
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import brute as B  # noqa: E402

P = B.P


def field_vectors():
    rng = random.Random(0x5EED)
    specials = [0, 1, 2, P - 1, P - 2, B.W_FROM_RAW7, 0xFFFFFFFF, 1 << 32, (1 << 32) + 1, 1 << 63, P >> 1]
    pairs = [(a, b) for a in specials for b in specials]
    pairs += [(rng.randrange(P), rng.randrange(P)) for _ in range(200)]
    mul = [[a, b, a * b % P] for a, b in pairs]
    add = [[a, b, (a + b) % P] for a, b in pairs[:60]]
    sub = [[a, b, (a - b) % P] for a, b in pairs[:60]]
    inv = [[a, pow(a, P - 2, P)] for a, _ in pairs[:40] if a]
    ext = {}
    for d in (2, 3):
        for w in (7, B.W_FROM_RAW7):
            F = B.Ext(d, w)
            cases = []
            for _ in range(20):
                a = tuple(rng.randrange(P) for _ in range(d))
                b = tuple(rng.randrange(P) for _ in range(d))
                cases.append([list(a), list(b), list(F.mul(a, b))])
            ext[f"d{d}_w{w:x}"] = cases
    return {"mul": mul, "add": add, "sub": sub, "inv": inv, "ext": ext,
            "mont_one": 0xFFFFFFFF, "mont_neg_one": 0xFFFFFFFE00000002}


def shake_vectors():
    msgs = [b"", b"abc", bytes(range(135)), bytes(range(136)), bytes(range(137)), b"effsc" * 100]
    return [[m.hex(), hashlib.shake_256(m).hexdigest(200)] for m in msgs]


def transcript_vectors():
    out = []
    for ds, ext in ((b"", 1), (b"effsc-hip/test", 1), (b"x" * 150, 3), (b"effsc-hip/fp2", 2)):
        t = B.ShakeTranscript(ds)
        ops = []
        for i in range(6):
            v = tuple(B.uniform(99, 3 * i + j) for j in range(ext))
            t.send(v)
            ops.append(["send", list(v)])
            if i % 2:
                ops.append(["challenge", list(t.challenge(ext))])
        ops.append(["challenge", list(t.challenge(ext))])
        out.append({"domsep": ds.hex(), "ext": ext, "ops": ops})
    # block-boundary cases of the padded stream: after a challenge the stream
    # is block-aligned; eight base-field sends (16 bytes each) and the
    # challenge word end exactly on the next boundary (a whole padding block
    # follows), and twelve F_p^3 sends cross two boundaries inside one round
    for ext, sends in ((1, (1, 8, 8, 7)), (3, (1, 12, 2))):
        t = B.ShakeTranscript(b"effsc-hip/blocks")
        ops = []
        for k, ns in enumerate(sends):
            for i in range(ns):
                v = tuple(B.uniform(98, 100 * k + 3 * i + j) for j in range(ext))
                t.send(v)
                ops.append(["send", list(v)])
            ops.append(["challenge", list(t.challenge(ext))])
        out.append({"domsep": b"effsc-hip/blocks".hex(), "ext": ext, "ops": ops})
    rng_t = B.RngTranscript(42)
    out.append({"rng_seed": 42, "challenges": [list(rng_t.challenge(1)) for _ in range(5)]})
    return out


def proof_vectors():
    cases = []

    def add(kind, ext, v, seeds, n=None, transcript="shake"):
        F = B.Ext(ext)
        n = (1 << v) if n is None else n
        t = B.ShakeTranscript(b"effsc-hip/golden") if transcript == "shake" else B.RngTranscript(7)
        if kind == "ml":
            tab = B.uniform_table(n, seeds[0])
            polys, chals, fin = B.prove_multilinear(F, tab, v, t)
        elif kind == "ip":
            ta, tb = B.uniform_table(n, seeds[0]), B.uniform_table(n, seeds[1])
            polys, chals, fin = B.prove_product(F, [ta, tb], v, t)
        elif kind.startswith("product"):
            k = int(kind[-1])
            polys, chals, fin = B.prove_product(F, [B.uniform_table(n, s) for s in seeds[:k]], v, t)
        elif kind == "eq":
            wpt = B.uniform_table(v * ext, seeds[1])
            wpt = [tuple(wpt[i * ext:(i + 1) * ext]) for i in range(v)] if ext > 1 else wpt
            pe = B.uniform_table(n, seeds[0])
            polys, chals, fin = B.prove_eq(F, wpt, pe, t, lift=(ext == 1))
        cases.append({"kind": kind, "ext": ext, "v": v, "n": n, "seeds": seeds, "transcript": transcript,
                      "domsep": "effsc-hip/golden", "round_polys": [[list(x) for x in rp] for rp in polys],
                      "challenges": [list(c) for c in chals], "final": list(fin)})

    for v in (1, 2, 3, 5):
        add("ml", 1, v, [11])
        add("ip", 1, v, [12, 13])
    add("ml", 1, 4, [11], n=13)
    add("ip", 1, 4, [12, 13], n=11)
    add("ip", 1, 3, [12, 13], n=5, transcript="rng")
    add("ml", 3, 3, [14])
    add("product1", 3, 4, [15])
    add("product2", 3, 4, [15, 16])
    add("product3", 1, 4, [15, 16, 17])
    add("product3", 3, 4, [15, 16, 17])
    add("product4", 1, 3, [15, 16, 17, 18])
    add("eq", 1, 4, [19, 20])
    add("eq", 1, 5, [19, 20])
    # F_p^2 = F_p[X]/(X^2 - w), w = from_raw(7): base inputs lifted
    add("ml", 2, 3, [21])
    add("ml", 2, 4, [21], n=11)
    add("ip", 2, 4, [22, 23])
    add("product2", 2, 3, [22, 23], transcript="rng")
    add("product3", 2, 4, [24, 25, 26])
    return cases


def main():
    data = {"field": field_vectors(), "shake256": shake_vectors(), "transcript": transcript_vectors(),
            "proofs": proof_vectors()}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(data, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
