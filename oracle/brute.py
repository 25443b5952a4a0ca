"""Pure-Python brute-force sumcheck, written from the mathematics.

TEST INFRASTRUCTURE ONLY.  It shares no code with the C oracle and folds
nothing: every round polynomial is evaluated from the definition

    g_j(t) = sum_{b in {0,1}^(v-j-1)} F(r_0, ..., r_{j-1}, t, b)

where F is the summand written with multilinear extensions in the reference's
MSB (half-split) variable order (src/multilinear_sumcheck.rs:1-8: round j binds
the top remaining index bit), and the Fiat-Shamir transcript is recomputed
from its whole byte stream with hashlib's SHAKE256.  Intended for v <= 6.

Wire format (EvalsInfty, src/sumcheck_prover.rs:40-59): degree 1 sends [g(0)];
degree d >= 2 sends [g(0), lead(g), g(2), ..., g(d-1)].
"""
from __future__ import annotations

import hashlib
from math import factorial

P = 0xFFFFFFFF00000001
W_FROM_RAW7 = 0xFFFFFFF800000001
MASK = (1 << 64) - 1


# ----------------------------------------------------------------- field
def inv(a):
    return pow(a % P, P - 2, P)


class Ext:
    """F_p[X]/(X^d - w); elements are tuples of d ints."""

    def __init__(self, d=1, w=W_FROM_RAW7):
        self.d = d
        self.w = w % P

    def zero(self):
        return (0,) * self.d

    def one(self):
        return (1,) + (0,) * (self.d - 1)

    def lift(self, x):
        return (x % P,) + (0,) * (self.d - 1)

    def add(self, a, b):
        return tuple((x + y) % P for x, y in zip(a, b))

    def sub(self, a, b):
        return tuple((x - y) % P for x, y in zip(a, b))

    def mul(self, a, b):
        d = self.d
        prod = [0] * (2 * d - 1)
        for i in range(d):
            for j in range(d):
                prod[i + j] += a[i] * b[j]
        for k in range(2 * d - 2, d - 1, -1):  # X^k = w X^(k-d)
            prod[k - d] += self.w * prod[k]
        return tuple(x % P for x in prod[:d])

    def scale(self, a, s):
        return tuple(x * s % P for x in a)


# ----------------------------------------------------------------- inputs
def _mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def uniform(seed, index):
    """Counter-based uniform element of [0, p): first candidate below p."""
    base = (_mix64(seed) + index * 0x9E3779B97F4A7C15) & MASK
    a = 0
    while True:
        c = _mix64((base + a * 0xD1B54A32D192ED03) & MASK)
        if c < P:
            return c
        a += 1


def uniform_table(n, seed):
    return [uniform(seed, i) for i in range(n)]


# ----------------------------------------------------------------- transcripts
FS_TAG = b"effsc-hip/fs/v3\x00"  # 16 bytes
SHAKE_RATE = 136  # SHAKE256 block size in bytes


def _pad8(bs):
    return bs + b"\x00" * (-len(bs) % 8)


class ShakeTranscript:
    """Byte-stream definition of the Fiat-Shamir transcript; every item is a
    whole number of little-endian 64-bit words.  A challenge is the SHAKE256
    of the stream so far; the stream is then padded to the next block
    boundary with SHAKE's own padding (0x1F, zeros, 0x80), so the sponge that
    produced the challenge is the sponge of the padded stream."""

    def __init__(self, domsep: bytes):
        self.stream = bytearray(FS_TAG + len(domsep).to_bytes(8, "little") + _pad8(domsep))

    def send(self, value):
        self.stream += (1).to_bytes(8, "little") + b"".join((x % P).to_bytes(8, "little") for x in value)

    def challenge(self, d):
        self.stream += (2).to_bytes(8, "little")
        out = hashlib.shake_256(bytes(self.stream)).digest(16 * d)
        pad = bytearray(SHAKE_RATE - len(self.stream) % SHAKE_RATE)
        pad[0] ^= 0x1F
        pad[-1] ^= 0x80
        self.stream += pad
        return tuple(int.from_bytes(out[16 * i:16 * i + 16], "little") % P for i in range(d))


class RngTranscript:
    """Test transcript: challenge = from_u64(splitmix64 next), lifted."""

    def __init__(self, seed):
        self.state = seed & MASK

    def send(self, value):
        pass

    def challenge(self, d):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        return (_mix64(self.state) % P,) + (0,) * (d - 1)


# ----------------------------------------------------------------- MLE
def mle(F: Ext, table, point):
    """sum_x table[x] * prod_j (z_j if bit_j(x) else 1 - z_j), MSB-first bits."""
    v = len(point)
    assert len(table) == 1 << v
    acc = F.zero()
    for x, val in enumerate(table):
        term = val
        for j, z in enumerate(point):
            bit = (x >> (v - 1 - j)) & 1
            term = F.mul(term, z if bit else F.sub(F.one(), z))
        acc = F.add(acc, term)
    return acc


def eq_point(F: Ext, w, z):
    acc = F.one()
    for wj, zj in zip(w, z):
        t = F.add(F.mul(wj, zj), F.mul(F.sub(F.one(), wj), F.sub(F.one(), zj)))
        acc = F.mul(acc, t)
    return acc


def _bits(b, m, F):
    return [F.lift((b >> (m - 1 - i)) & 1) for i in range(m)]


def _lead_coeff(F, evals, d):
    acc = F.zero()
    for i, g in enumerate(evals):
        c = inv(factorial(i) * factorial(d - i))
        if (d - i) % 2:
            c = (P - c) % P
        acc = F.add(acc, F.scale(g, c))
    return acc


def prove(F: Ext, summand, v, degree, transcript, num_rounds=None):
    """Generic brute-force prover for a summand F(z) given as a callable on a
    full point in F^v.  Returns (round_polys, challenges, final_value) with
    every element a tuple of d ints."""
    num_rounds = v if num_rounds is None else num_rounds
    chals, polys = [], []
    for j in range(num_rounds):
        m = v - j - 1
        evals = []
        for t in range(degree + 1):
            acc = F.zero()
            for b in range(1 << m):
                acc = F.add(acc, summand(chals + [F.lift(t)] + _bits(b, m, F)))
            evals.append(acc)
        if degree <= 1:
            wire = [evals[0]]
        else:
            wire = [evals[0], _lead_coeff(F, evals, degree)] + [evals[i] for i in range(2, degree)]
        for x in wire:
            transcript.send(x)
        polys.append(wire)
        chals.append(transcript.challenge(F.d))
    final = summand(chals) if num_rounds == v else None
    return polys, chals, final


def _pad(table, v, F, lift):
    t = [F.lift(x) if lift else tuple(x) for x in table]
    return t + [F.zero()] * ((1 << v) - len(t))


def prove_multilinear(F, evals, v, transcript, lift=True):
    T = _pad(evals, v, F, lift)
    return prove(F, lambda z: mle(F, T, z), v, 1, transcript)


def prove_product(F, tables, v, transcript, lift=True):
    Ts = [_pad(t, v, F, lift) for t in tables]

    def summand(z):
        acc = F.one()
        for T in Ts:
            acc = F.mul(acc, mle(F, T, z))
        return acc

    return prove(F, summand, v, len(tables), transcript)


def prove_eq(F, w, p_evals, transcript, lift=True):
    v = len(w)
    T = _pad(p_evals, v, F, lift)
    W = [F.lift(x) if lift else tuple(x) for x in w]
    return prove(F, lambda z: F.mul(eq_point(F, W, z), mle(F, T, z)), v, 2, transcript)


def claimed_sum_product(F, tables, lift=True):
    Ts = [[F.lift(x) if lift else tuple(x) for x in t] for t in tables]
    acc = F.zero()
    for i in range(len(Ts[0])):
        term = F.one()
        for T in Ts:
            term = F.mul(term, T[i])
        acc = F.add(acc, term)
    return acc


def prove_gkr(F, add, mult, wit, transcript, lift=True):
    """GKR round sumcheck from its definition: the summand is
    A(z) (Wb(z) + Wc(z)) + M(z) Wb(z) Wc(z) over the multilinear extensions
    of the (zero-padded) tables add, mult, w_b[b n + c] = W(b) and
    w_c[b n + c] = W(c); the wire's second value is the leading coefficient
    of the degree-2 interpolant through q(0), q(1), q(2)."""
    n = len(wit)
    nbc = n * n
    v = (nbc - 1).bit_length() if nbc > 1 else 0
    wb = [wit[i // n] for i in range(nbc)]
    wc = [wit[i % n] for i in range(nbc)]
    A, M, B, Cc = [_pad(t, v, F, lift) for t in (add, mult, wb, wc)]

    def summand(z):
        a, m, b, c = mle(F, A, z), mle(F, M, z), mle(F, B, z), mle(F, Cc, z)
        return F.add(F.mul(a, F.add(b, c)), F.mul(m, F.mul(b, c)))
    return prove(F, summand, v, 2, transcript)


def prove_lsb(F, tables, v, transcript, degree, lift=True):
    """An LSB (pair-split) prover binds the least significant index bit
    first: the brute force of the product of the tables' extensions with the
    point's coordinates taken in reverse order."""
    Ts = [_pad(t, v, F, lift) for t in tables]

    def summand(z):
        zr = list(reversed(z))
        acc = F.one()
        for T in Ts:
            acc = F.mul(acc, mle(F, T, zr))
        return acc
    return prove(F, summand, v, degree, transcript)


def _wire(F, evals, degree):
    if degree <= 1:
        return [evals[0]]
    return [evals[0], _lead_coeff(F, evals, degree)] + [evals[i] for i in range(2, degree)]


def prove_halving(F, tables, num_rounds, transcript, lift=True):
    """CoefficientProver with the product evaluator over tables of any length
    (src/provers/coefficient.rs:66-140, 272-329), from its definition: round
    j pairs k with k + h, h = floor(L_j / 2), for k < h (an odd last element
    takes no part); the round polynomial is
        g_j(t) = sum_{k < h} prod_i ((1 - t) T_i[k] + t T_i[k + h]),
    evaluated at t = 0..K and sent as g(0), its leading coefficient, g(2), ...;
    binding r keeps T_i[k] <- (1 - r) T_i[k] + r T_i[k + h] for k < h."""
    Ts = [[F.lift(x) if lift else tuple(x) for x in t] for t in tables]
    degree = len(Ts)
    polys, chals = [], []
    for _ in range(num_rounds):
        h = len(Ts[0]) // 2
        evals = []
        for t in range(degree + 1):
            tt = F.lift(t)
            omt = F.sub(F.one(), tt)
            acc = F.zero()
            for k in range(h):
                term = F.one()
                for T in Ts:
                    term = F.mul(term, F.add(F.mul(omt, T[k]), F.mul(tt, T[k + h])))
                acc = F.add(acc, term)
            evals.append(acc)
        wire = _wire(F, evals, degree)
        for x in wire:
            transcript.send(x)
        polys.append(wire)
        r = transcript.challenge(F.d)
        chals.append(r)
        omr = F.sub(F.one(), r)
        if len(Ts[0]) > 1:  # a table of one element is left as it is
            Ts = [[F.add(F.mul(omr, T[k]), F.mul(r, T[k + h])) for k in range(h)] for T in Ts]
    final = F.zero()
    if len(Ts[0]) == 1:
        final = F.one()
        for T in Ts:
            final = F.mul(final, T[0])
    return polys, chals, final


def prove_ml_pairs(F, table, num_rounds, transcript, lift=True):
    """MultilinearProverLSB over a table of any length
    (src/provers/multilinear_lsb.rs:59-126, 168-192), from its definition:
    the round value is the sum of the even-indexed elements (an unpaired last
    element included), and binding r keeps T[k] <- (1 - r) T[2k] + r T[2k+1]
    for k < floor(L / 2) (a one-element table stays)."""
    T = [F.lift(x) if lift else tuple(x) for x in table]
    polys, chals = [], []
    for _ in range(num_rounds):
        s0 = F.zero()
        for i in range(0, len(T), 2):
            s0 = F.add(s0, T[i])
        transcript.send(s0)
        polys.append([s0])
        r = transcript.challenge(F.d)
        chals.append(r)
        omr = F.sub(F.one(), r)
        if len(T) > 1:  # a table of one element is left as it is
            T = [F.add(F.mul(omr, T[2 * k]), F.mul(r, T[2 * k + 1])) for k in range(len(T) // 2)]
    final = T[0] if len(T) == 1 else F.zero()
    return polys, chals, final, T
