"""Goldilocks field helpers for the host side (src/field.rs:151, src/tests/fields.rs:26-100).

Elements: an int in [0, p) for the base field; a tuple (c0, c1) for
F_p[X]/(X^2 - w) and (c0, c1, c2) for F_p[X]/(X^3 - w).  Words crossing the C ABI are canonical unless a prover
was built with repr=MONT64.
"""
from __future__ import annotations

P = 0xFFFFFFFF00000001
EPS = 0xFFFFFFFF
# F64::from_raw(7): canonical value 7 * 2^-64 mod p (src/tests/fields.rs:49,80)
W_FROM_RAW7 = 0xFFFFFFF800000001
CANONICAL, MONT64 = 0, 1


def to_mont(x: int) -> int:
    """Canonical value -> arkworks Montgomery word (R = 2^64)."""
    return (x % P) * (1 << 64) % P


def from_mont(m: int) -> int:
    return m * pow(1 << 64, P - 2, P) % P


class Field:
    """Arithmetic of F_p (ext = 1), F_p[X]/(X^2 - w) (ext = 2) or
    F_p[X]/(X^3 - w) (ext = 3)."""

    def __init__(self, ext: int = 1, w: int = W_FROM_RAW7):
        if ext not in (1, 2, 3):
            raise ValueError("ext must be 1, 2 or 3")
        self.ext = ext
        self.w = w % P

    # element <-> words
    def words(self, x):
        if self.ext == 1:
            return [int(x) % P]
        return [int(c) % P for c in x]

    def elem(self, words):
        if self.ext == 1:
            return int(words[0])
        return tuple(int(c) for c in words[:self.ext])

    def zero(self):
        return 0 if self.ext == 1 else (0,) * self.ext

    def one(self):
        return 1 if self.ext == 1 else (1,) + (0,) * (self.ext - 1)

    def from_u64(self, x):
        return x % P if self.ext == 1 else (x % P,) + (0,) * (self.ext - 1)

    def add(self, a, b):
        if self.ext == 1:
            return (a + b) % P
        return tuple((x + y) % P for x, y in zip(a, b))

    def sub(self, a, b):
        if self.ext == 1:
            return (a - b) % P
        return tuple((x - y) % P for x, y in zip(a, b))

    def neg(self, a):
        return self.sub(self.zero(), a)

    def mul(self, a, b):
        if self.ext == 1:
            return a * b % P
        if self.ext == 2:
            (a0, a1), (b0, b1) = a, b
            return ((a0 * b0 + self.w * a1 * b1) % P, (a0 * b1 + a1 * b0) % P)
        a0, a1, a2 = a
        b0, b1, b2 = b
        w = self.w
        return ((a0 * b0 + w * (a1 * b2 + a2 * b1)) % P,
                (a0 * b1 + a1 * b0 + w * a2 * b2) % P,
                (a0 * b2 + a1 * b1 + a2 * b0) % P)

    def inv(self, a):
        if self.ext == 1:
            return pow(a, P - 2, P) if a % P else 0
        if self.ext == 2:
            a0, a1 = a
            n = (a0 * a0 - self.w * a1 * a1) % P
            ni = pow(n, P - 2, P) if n else 0
            return (a0 * ni % P, (-a1) * ni % P)
        a0, a1, a2 = a
        w = self.w
        t0 = (a0 * a0 - w * a1 * a2) % P
        t1 = (w * a2 * a2 - a0 * a1) % P
        t2 = (a1 * a1 - a0 * a2) % P
        n = (a0 * t0 + w * (a2 * t1 + a1 * t2)) % P
        ni = pow(n, P - 2, P) if n else 0
        return (t0 * ni % P, t1 * ni % P, t2 * ni % P)

    def is_zero(self, a):
        return a == self.zero()
