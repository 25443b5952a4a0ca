"""CPU model of the device Keccak-f[1600] lane layout (keccak_lanes in
efficient-sumcheck_amd/csrc/prove.hpp).

64 lanes of 32-bit halves, the gfx950 lane operations written out as
documented (DPP row shifts within 16-lane rows, out-of-row reads giving zero;
v_permlane32_swap / v_permlane16_swap; ds_bpermute; v_alignbit_b32;
v_bitop3_b32 with truth-table index 4 src0 + 2 src1 + src2) and the same
sequence per round as the kernel.  It checks the layout's algebra (positions
and copies, pi sources, the packed column parity, the truth tables) against a
plain Keccak-f[1600], itself pinned by hashlib's SHAKE256.  The GPU transcript
tests check the kernel.
"""
import hashlib
import random

M32 = 0xFFFFFFFF
M64 = (1 << 64) - 1
# FIPS 202, 3.2: round constants and rotation offsets r[x + 5 y]
RC = [0x0000000000000001, 0x0000000000008082, 0x800000000000808A, 0x8000000080008000, 0x000000000000808B,
      0x0000000080000001, 0x8000000080008081, 0x8000000000008009, 0x000000000000008A, 0x0000000000000088,
      0x0000000080008009, 0x000000008000000A, 0x000000008000808B, 0x800000000000008B, 0x8000000000008089,
      0x8000000000008003, 0x8000000000008002, 0x8000000000000080, 0x000000000000800A, 0x800000008000000A,
      0x8000000080008081, 0x8000000000008080, 0x0000000080000001, 0x8000000080008008]
ROT = [0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14]


def rotl(x, r):
    return ((x << r) | (x >> (64 - r))) & M64 if r else x


def keccak_f(state):
    s = list(state)
    for rnd in range(24):
        c = [s[x] ^ s[x + 5] ^ s[x + 10] ^ s[x + 15] ^ s[x + 20] for x in range(5)]
        for x in range(5):
            d = c[(x + 4) % 5] ^ rotl(c[(x + 1) % 5], 1)
            for y in range(0, 25, 5):
                s[x + y] ^= d
        b = [0] * 25
        for x in range(5):
            for y in range(5):
                b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl(s[x + 5 * y], ROT[x + 5 * y])
        for y in range(0, 25, 5):
            for x in range(5):
                s[x + y] = b[x + y] ^ ((~b[(x + 1) % 5 + y]) & b[(x + 2) % 5 + y])
        s[0] ^= RC[rnd]
    return s


# ---- lane operations on lists of 64 32-bit values
def row_shl(v, n):  # lane l reads l + n within its 16-lane row
    return [v[l + n] if (l & 15) + n < 16 else 0 for l in range(64)]


def row_shr(v, n):  # lane l reads l - n within its row
    return [v[l - n] if (l & 15) - n >= 0 else 0 for l in range(64)]


def row_ror(v, n):
    return [v[(l & ~15) | (((l & 15) - n) % 16)] for l in range(64)]


def _rows(v, r):
    return v[16 * r:16 * r + 16]


def permlane32_swap(a, b):  # a' = (a0, a1, b0, b1), b' = (a2, a3, b2, b3)
    return (_rows(a, 0) + _rows(a, 1) + _rows(b, 0) + _rows(b, 1),
            _rows(a, 2) + _rows(a, 3) + _rows(b, 2) + _rows(b, 3))


def permlane16_swap(a, b):  # a' = (a0, b0, a2, b2), b' = (a1, b1, a3, b3)
    return (_rows(a, 0) + _rows(b, 0) + _rows(a, 2) + _rows(b, 2),
            _rows(a, 1) + _rows(b, 1) + _rows(a, 3) + _rows(b, 3))


def bpermute(v, src):
    return [v[src[l]] for l in range(64)]


def xor(a, b):
    return [x ^ y for x, y in zip(a, b)]


def alignbit(a, b, s):
    return [(((x << 32) | y) >> t) & M32 for x, y, t in zip(a, b, s)]


def select(m, a, b):
    return [x if c else y for c, x, y in zip(m, a, b)]


def bitop3(a, b, c, table):
    out = []
    for x, y, z in zip(a, b, c):
        r = 0
        for bit in range(32):
            i = (((x >> bit) & 1) << 2) | (((y >> bit) & 1) << 1) | ((z >> bit) & 1)
            r |= ((table >> i) & 1) << bit
        out.append(r)
    return out


def keccak_lanes(state):
    """The kernel's rounds on the lane layout: group g = y of 8 lanes, position
    p holding x = (p + 4) mod 5, groups 5..7 zero."""
    lanes = range(64)
    g = [l >> 3 for l in lanes]
    p = [l & 7 for l in lanes]
    active = [gg < 5 for gg in g]
    X = [(pp + 4) % 5 for pp in p]
    Y = [gg if a else 0 for gg, a in zip(g, active)]
    sx = [(3 * (y - 3 * x)) % 5 for x, y in zip(X, Y)]
    src = [8 * sy + s + 1 for s, sy in zip(sx, X)]
    r = [ROT[s + 5 * sy] for s, sy in zip(sx, X)]
    tsh = [(32 - (rr & 31)) & 31 for rr in r]
    swp = [rr == 0 or rr > 32 for rr in r]
    m_iota = [M32 if l == 1 else 0 for l in lanes]
    copy = [pp == 6 for pp in p]
    lo = [state[X[l] + 5 * Y[l]] & M32 if active[l] else 0 for l in lanes]
    hi = [state[X[l] + 5 * Y[l]] >> 32 if active[l] else 0 for l in lanes]
    for rnd in range(24):
        # theta
        tl, th = xor(lo, row_ror(lo, 8)), xor(hi, row_ror(hi, 8))
        u = xor(*permlane32_swap(tl, th))
        w = xor(*permlane16_swap(u, u))
        cl, ch = permlane32_swap(w, w)
        pl, ph = row_shl(cl, 1), row_shl(ch, 1)
        lo = bitop3(lo, row_shr(cl, 1), alignbit(pl, ph, [31] * 64), 0x96)
        hi = bitop3(hi, row_shr(ch, 1), alignbit(ph, pl, [31] * 64), 0x96)
        # rho + pi
        sl, sh = bpermute(lo, src), bpermute(hi, src)
        P, Q = select(swp, sl, sh), select(swp, sh, sl)
        bh, bl = alignbit(P, Q, tsh), alignbit(Q, P, tsh)
        # chi, iota, copy refresh, inactive groups
        lo = bitop3(bl, row_shl(bl, 1), row_shl(bl, 2), 0xD2)
        hi = bitop3(bh, row_shl(bh, 1), row_shl(bh, 2), 0xD2)
        lo = bitop3(lo, [RC[rnd] & M32] * 64, m_iota, 0x78)
        hi = bitop3(hi, [RC[rnd] >> 32] * 64, m_iota, 0x78)
        lo = select(copy, row_shr(lo, 5), lo)
        hi = select(copy, row_shr(hi, 5), hi)
        lo = [v if a else 0 for v, a in zip(lo, active)]
        hi = [v if a else 0 for v, a in zip(hi, active)]
    out = []
    for i in range(25):
        l = 8 * (i // 5) + i % 5 + 1
        out.append(lo[l] | (hi[l] << 32))
    return out


def test_reference_keccak_matches_hashlib():
    # SHAKE256 of the empty message: one permutation of the padded state
    st = [0] * 25
    st[0] ^= 0x1F
    st[16] ^= 0x80 << 56
    out = b"".join(x.to_bytes(8, "little") for x in keccak_f(st))[:32]
    assert out == hashlib.shake_256(b"").digest(32)


def test_lane_layout_matches_reference_keccak():
    rng = random.Random(1)
    for _ in range(3):
        s = [rng.getrandbits(64) for _ in range(25)]
        assert keccak_lanes(s) == keccak_f(s)
    assert keccak_lanes([0] * 25) == keccak_f([0] * 25)


def test_a_wrong_chi_table_fails():
    # the model distinguishes truth tables: chi with src1 and src2 swapped
    # (0xB4 is b ^ (b1 & ~b2)) is a different permutation
    assert bitop3([0xF0], [0xCC], [0xAA], 0xD2) == [0xF0 ^ (~0xCC & 0xAA & M32)]
    assert bitop3([0xF0], [0xCC], [0xAA], 0xB4) != bitop3([0xF0], [0xCC], [0xAA], 0xD2)
