/*
 * effsc_oracle.c — CPU oracle for the Goldilocks MSB sumcheck path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * baseline.  It is never linked into, loaded by, or called from the product.
 *
 * A plain-C restatement of the reference crate `effsc` 0.0.2 (Rust).  The
 * reference cannot be compiled here (no Rust toolchain, crates not vendored),
 * so each function below cites the reference file:line whose behaviour it
 * follows.  Arithmetic is arkworks-style Montgomery form (R = 2^64), which is
 * the reference's in-memory contract (src/simd_fields/goldilocks/avx512.rs:1-35,
 * src/field.rs:99-145).  Every value crossing this file's interface is
 * canonical.
 *
 * What pins each part (see tests/test_oracle.py):
 *   field      - Python big-integer arithmetic, the reference's constants and
 *                edge cases (avx512.rs:19-35, 689-701; tests/fields.rs:61-64)
 *   extensions - schoolbook formulas (avx512.rs:749-751, 776-779), a*inv(a)=1
 *   keccak     - hashlib.shake_256 / sha3_256 vectors
 *   provers    - closed forms of the reference's tests (round-0 half sums and
 *                difference form, MLE evaluation at the challenges, verifier
 *                acceptance, EqFactored == InnerProduct(p, eq table), d=1
 *                coefficient == multilinear) and a pure-Python brute force
 */
#include "effsc_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* Base field, Montgomery form.  Constants: avx512.rs:19-35.                 */
/* ------------------------------------------------------------------------ */
#define P 0xFFFFFFFF00000001ULL
#define EPS 0xFFFFFFFFULL              /* 2^64 mod p */
#define INV 0xFFFFFFFEFFFFFFFFULL      /* -p^-1 mod 2^64 */
#define MONT_ONE EPS                   /* R mod p */
#define R2 0xFFFFFFFE00000001ULL       /* R^2 mod p = eps^2 */

/* scalar_add: avx512.rs:145-155 */
static inline uint64_t m_add(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    if (s < a) return s + EPS;
    if (s >= P) return s - P;
    return s;
}
/* scalar_sub: avx512.rs:157-164 */
static inline uint64_t m_sub(uint64_t a, uint64_t b) {
    if (a >= b) return a - b;
    return a - b + P;
}
static inline uint64_t m_neg(uint64_t a) { return a == 0 ? 0 : P - a; }
/* mont_mul, CIOS with N = 1 as arkworks' MontBackend: avx512.rs:521-547 */
static inline uint64_t m_mul(uint64_t a, uint64_t b) {
    u128 full = (u128)a * b;
    uint64_t lo = (uint64_t)full, hi = (uint64_t)(full >> 64);
    uint64_t k = lo * INV;
    u128 t = (u128)k * P;
    uint64_t t_lo = (uint64_t)t, t_hi = (uint64_t)(t >> 64);
    uint64_t s0 = lo + t_lo;
    int carry = s0 < lo;
    uint64_t r = hi + t_hi;
    int carry2 = r < hi;
    uint64_t r2 = r + (uint64_t)carry;
    int carry3 = r2 < r;
    r = r2;
    if (carry2 || carry3 || r >= P) r -= P;
    return r;
}
static inline uint64_t m_to(uint64_t canon) { return m_mul(canon % P, R2); }
static inline uint64_t m_from(uint64_t mont) { return m_mul(mont, 1); }
static uint64_t m_pow(uint64_t a, uint64_t e) {
    uint64_t r = MONT_ONE;
    while (e) {
        if (e & 1) r = m_mul(r, a);
        a = m_mul(a, a);
        e >>= 1;
    }
    return r;
}
static inline uint64_t m_inv(uint64_t a) { return m_pow(a, P - 2); }

uint64_t or_gl_add(uint64_t a, uint64_t b) { return m_add(a % P, b % P); }
uint64_t or_gl_sub(uint64_t a, uint64_t b) { return m_sub(a % P, b % P); }
uint64_t or_gl_mul(uint64_t a, uint64_t b) { return m_from(m_mul(m_to(a), m_to(b))); }
uint64_t or_gl_mul_i128(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % P); }
uint64_t or_gl_inv(uint64_t a) { return a % P == 0 ? 0 : m_from(m_inv(m_to(a))); }
uint64_t or_mont_mul(uint64_t a, uint64_t b) { return m_mul(a, b); }
uint64_t or_to_mont(uint64_t a) { return m_to(a); }
uint64_t or_from_mont(uint64_t a) { return m_from(a); }

/* ------------------------------------------------------------------------ */
/* Extension fields F_p[X]/(X^d - w), d <= 3.  The reference's F64Ext2/3 are   */
/* arkworks Quad/CubicExtField with NONRESIDUE = F64::from_raw(7)           */
/* (src/tests/fields.rs:41-100); w is a runtime parameter here.             */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t c[3]; } el;
typedef struct { int d; uint64_t wm; } fctx;

static inline el e_zero(void) { el r = {{0, 0, 0}}; return r; }
static inline el e_one(void) { el r = {{MONT_ONE, 0, 0}}; return r; }
static inline el e_from_u64(uint64_t x) { el r = {{m_to(x % P), 0, 0}}; return r; }
static inline el e_add(const fctx *f, el a, el b) {
    el r = e_zero();
    for (int i = 0; i < f->d; i++) r.c[i] = m_add(a.c[i], b.c[i]);
    return r;
}
static inline el e_sub(const fctx *f, el a, el b) {
    el r = e_zero();
    for (int i = 0; i < f->d; i++) r.c[i] = m_sub(a.c[i], b.c[i]);
    return r;
}
static inline el e_mul(const fctx *f, el a, el b) {
    el r = e_zero();
    if (f->d == 1) {
        r.c[0] = m_mul(a.c[0], b.c[0]);
    } else if (f->d == 2) {
        /* (a0 + a1 X)(b0 + b1 X), X^2 = w */
        r.c[0] = m_add(m_mul(a.c[0], b.c[0]), m_mul(f->wm, m_mul(a.c[1], b.c[1])));
        r.c[1] = m_add(m_mul(a.c[0], b.c[1]), m_mul(a.c[1], b.c[0]));
    } else {
        /* schoolbook mod X^3 - w */
        uint64_t a0 = a.c[0], a1 = a.c[1], a2 = a.c[2];
        uint64_t b0 = b.c[0], b1 = b.c[1], b2 = b.c[2];
        r.c[0] = m_add(m_mul(a0, b0), m_mul(f->wm, m_add(m_mul(a1, b2), m_mul(a2, b1))));
        r.c[1] = m_add(m_add(m_mul(a0, b1), m_mul(a1, b0)), m_mul(f->wm, m_mul(a2, b2)));
        r.c[2] = m_add(m_add(m_mul(a0, b2), m_mul(a1, b1)), m_mul(a2, b0));
    }
    return r;
}
static el e_inv(const fctx *f, el a) {
    el r = e_zero();
    if (f->d == 1) {
        r.c[0] = a.c[0] ? m_inv(a.c[0]) : 0;
    } else if (f->d == 2) {
        /* (a0 - a1 X) / (a0^2 - w a1^2) */
        uint64_t n = m_sub(m_mul(a.c[0], a.c[0]), m_mul(f->wm, m_mul(a.c[1], a.c[1])));
        uint64_t ni = n ? m_inv(n) : 0;
        r.c[0] = m_mul(a.c[0], ni);
        r.c[1] = m_mul(m_neg(a.c[1]), ni);
    } else {
        /* adjugate of the multiplication matrix */
        uint64_t a0 = a.c[0], a1 = a.c[1], a2 = a.c[2], w = f->wm;
        uint64_t t0 = m_sub(m_mul(a0, a0), m_mul(w, m_mul(a1, a2)));
        uint64_t t1 = m_sub(m_mul(w, m_mul(a2, a2)), m_mul(a0, a1));
        uint64_t t2 = m_sub(m_mul(a1, a1), m_mul(a0, a2));
        uint64_t n = m_add(m_mul(a0, t0), m_mul(w, m_add(m_mul(a2, t1), m_mul(a1, t2))));
        uint64_t ni = n ? m_inv(n) : 0;
        r.c[0] = m_mul(t0, ni);
        r.c[1] = m_mul(t1, ni);
        r.c[2] = m_mul(t2, ni);
    }
    return r;
}
static inline int e_is_zero(const fctx *f, el a) {
    for (int i = 0; i < f->d; i++)
        if (a.c[i]) return 0;
    return 1;
}
static inline el e_load(const fctx *f, const uint64_t *canon) {
    el r = e_zero();
    for (int i = 0; i < f->d; i++) r.c[i] = m_to(canon[i]);
    return r;
}
static inline void e_store(const fctx *f, el a, uint64_t *canon) {
    for (int i = 0; i < f->d; i++) canon[i] = m_from(a.c[i]);
}
static fctx mk_ctx(int d, uint64_t w) {
    fctx f;
    f.d = d;
    f.wm = m_to(w);
    return f;
}

void or_ext_mul(int d, uint64_t w, const uint64_t *a, const uint64_t *b, uint64_t *out) {
    fctx f = mk_ctx(d, w);
    e_store(&f, e_mul(&f, e_load(&f, a), e_load(&f, b)), out);
}
void or_ext_inv(int d, uint64_t w, const uint64_t *a, uint64_t *out) {
    fctx f = mk_ctx(d, w);
    e_store(&f, e_inv(&f, e_load(&f, a)), out);
}

/* ------------------------------------------------------------------------ */
/* Synthetic inputs: x_i = first candidate < p of mix(mix(seed) + i*G + a*K) */
/* ------------------------------------------------------------------------ */
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
uint64_t or_uniform(uint64_t seed, uint64_t index) {
    uint64_t base = mix64(seed) + index * 0x9E3779B97F4A7C15ULL;
    for (uint64_t a = 0;; a++) {
        uint64_t c = mix64(base + a * 0xD1B54A32D192ED03ULL);
        if (c < P) return c;
    }
}
void or_fill_uniform(uint64_t *out, size_t n, uint64_t seed) {
    for (size_t i = 0; i < n; i++) out[i] = or_uniform(seed, i);
}

/* ------------------------------------------------------------------------ */
/* Keccak-f[1600] (FIPS 202) and SHAKE256                                   */
/* ------------------------------------------------------------------------ */
static const uint64_t KRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808AULL, 0x8000000080008000ULL,
    0x000000000000808BULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008AULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000AULL,
    0x000000008000808BULL, 0x800000000000008BULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800AULL, 0x800000008000000AULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};
static const int KROT[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                             25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};
static inline uint64_t rotl(uint64_t x, int r) { return r ? (x << r) | (x >> (64 - r)) : x; }

void or_keccak_f1600(uint64_t A[25]) {
    for (int rnd = 0; rnd < 24; rnd++) {
        uint64_t C[5], D[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
        for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rotl(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) A[i] ^= D[i % 5];
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++) B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl(A[x + 5 * y], KROT[x + 5 * y]);
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                A[x + 5 * y] = B[x + 5 * y] ^ (~B[(x + 1) % 5 + 5 * y] & B[(x + 2) % 5 + 5 * y]);
        A[0] ^= KRC[rnd];
    }
}

#define RATE 136
typedef struct { uint64_t st[25]; size_t pos; } sponge;
static void sp_init(sponge *s) { memset(s, 0, sizeof *s); }
static void sp_absorb_byte(sponge *s, uint8_t b) {
    s->st[s->pos / 8] ^= (uint64_t)b << (8 * (s->pos % 8));
    if (++s->pos == RATE) {
        or_keccak_f1600(s->st);
        s->pos = 0;
    }
}
static void sp_absorb(sponge *s, const uint8_t *m, size_t n) {
    for (size_t i = 0; i < n; i++) sp_absorb_byte(s, m[i]);
}
/* pad (SHAKE domain 0x1F) and squeeze outlen bytes; destroys the sponge */
static void sp_finish(sponge *s, uint8_t *out, size_t outlen) {
    s->st[s->pos / 8] ^= (uint64_t)0x1F << (8 * (s->pos % 8));
    s->st[(RATE - 1) / 8] ^= (uint64_t)0x80 << (8 * ((RATE - 1) % 8));
    or_keccak_f1600(s->st);
    size_t i = 0, blk = 0;
    while (i < outlen) {
        if (blk == RATE) {
            or_keccak_f1600(s->st);
            blk = 0;
        }
        out[i++] = (uint8_t)(s->st[blk / 8] >> (8 * (blk % 8)));
        blk++;
    }
}
void or_shake256(const uint8_t *msg, size_t len, uint8_t *out, size_t outlen) {
    sponge s;
    sp_init(&s);
    sp_absorb(&s, msg, len);
    sp_finish(&s, out, outlen);
}

/* ------------------------------------------------------------------------ */
/* Transcripts.  Traits: src/transcript/transcript.rs:11-43.  The RNG kind    */
/* mirrors TestTranscript (src/transcript/sanity.rs:30-61): sends are no-ops   */
/* and a challenge is F::from_u64(rng.next_u64()), lifted into extensions.    */
/* ------------------------------------------------------------------------ */
struct or_transcript {
    int kind; /* 0 shake, 1 rng */
    sponge sp;
    uint64_t rng;
};
static const char FS_TAG[16] = "effsc-hip/fs/v3"; /* 15 chars + NUL = 16 bytes */

static void sp_absorb_u64(sponge *s, uint64_t x) {
    uint8_t b[8];
    for (int j = 0; j < 8; j++) b[j] = (uint8_t)(x >> (8 * j));
    sp_absorb(s, b, 8);
}

or_transcript *or_tr_new_shake(const uint8_t *domsep, size_t len) {
    or_transcript *t = (or_transcript *)calloc(1, sizeof *t);
    t->kind = 0;
    sp_init(&t->sp);
    sp_absorb(&t->sp, (const uint8_t *)FS_TAG, 16);
    sp_absorb_u64(&t->sp, (uint64_t)len);
    sp_absorb(&t->sp, domsep, len);
    for (size_t i = len; i % 8; i++) sp_absorb_byte(&t->sp, 0);
    return t;
}
or_transcript *or_tr_new_rng(uint64_t seed) {
    or_transcript *t = (or_transcript *)calloc(1, sizeof *t);
    t->kind = 1;
    t->rng = seed;
    return t;
}
static uint64_t splitmix_next(uint64_t *s) {
    *s += 0x9E3779B97F4A7C15ULL;
    return mix64(*s);
}
void or_tr_send(or_transcript *t, int d, const uint64_t *value) {
    if (t->kind != 0) return;
    sp_absorb_u64(&t->sp, 1);
    for (int i = 0; i < d; i++) sp_absorb_u64(&t->sp, value[i] % P);
}
void or_tr_challenge(or_transcript *t, int d, uint64_t *out) {
    if (t->kind == 1) {
        out[0] = splitmix_next(&t->rng) % P;
        for (int i = 1; i < d; i++) out[i] = 0;
        return;
    }
    /* the stream so far is padded to the block boundary (pad10*1 with the
       SHAKE suffix 0x1F) and the padded block absorbed: the output is the
       first 16 d bytes of SHAKE256 of the stream, and the padding stays in
       the stream, so the next item starts a fresh block */
    sp_absorb_u64(&t->sp, 2);
    uint8_t buf[48];
    sp_finish(&t->sp, buf, (size_t)16 * d);
    t->sp.pos = 0;
    for (int i = 0; i < d; i++) {
        uint64_t lo = 0, hi = 0;
        for (int j = 0; j < 8; j++) lo |= (uint64_t)buf[16 * i + j] << (8 * j);
        for (int j = 0; j < 8; j++) hi |= (uint64_t)buf[16 * i + 8 + j] << (8 * j);
        out[i] = (uint64_t)((((u128)hi << 64) | lo) % P);
    }
}
void or_tr_free(or_transcript *t) { free(t); }

/* ------------------------------------------------------------------------ */
/* Provers                                                                   */
/* ------------------------------------------------------------------------ */
#define MAXTAB 8
struct or_prover {
    int kind;
    fctx f;
    int deg;
    int ntab;
    el *tab[MAXTAB];
    size_t len[MAXTAB];
    /* EqFactored: tab[0] = p, tab[1] = eq_l, tab[2] = eq_r */
    size_t v, v_l, rounds_elapsed;
    int lsb; /* pair-split layout, pairs (2k, 2k+1): the provers in src/provers/..._lsb.rs */
};

static size_t next_pow2(size_t n) {
    size_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

static el *load_table(const fctx *f, const uint64_t *canon, size_t n, int lift_base) {
    el *t = (el *)malloc(sizeof(el) * (n ? n : 1));
    for (size_t i = 0; i < n; i++) {
        if (lift_base) {
            t[i] = e_zero();
            t[i].c[0] = m_to(canon[i]);
        } else {
            t[i] = e_load(f, canon + i * (size_t)f->d);
        }
    }
    return t;
}

/* In-place half-split fold with implicit zero padding of the high half:
 * new[k] = v[k] + (v[k+h] - v[k])*w, unpaired tail new[k] = v[k]*(1 - w),
 * h = next_pow2(n)/2.  src/inner_product_sumcheck.rs:143-176 (ip::fold) and
 * src/multilinear_sumcheck.rs:132-178 (fold; its SIMD path
 * src/simd_sumcheck/reduce.rs:27-75 has the same tail rule). */
static void fold_msb(const fctx *f, el *t, size_t *len, el w) {
    size_t n = *len;
    if (n <= 1) return;
    size_t half = next_pow2(n) >> 1;
    size_t paired = n - half;
    for (size_t k = 0; k < paired; k++) t[k] = e_add(f, t[k], e_mul(f, e_sub(f, t[k + half], t[k]), w));
    el omw = e_sub(f, e_one(), w);
    for (size_t k = paired; k < half; k++) t[k] = e_mul(f, t[k], omw);
    *len = half;
}

/* CoefficientProver's fold: msb_fold_vec, src/provers/coefficient.rs:131-140
 * (half = n/2, no padding; the reference drops an odd trailing element). */
static void fold_vec(const fctx *f, el *t, size_t *len, el c) {
    size_t n = *len;
    if (n <= 1) return;
    size_t half = n / 2;
    for (size_t k = 0; k < half; k++) t[k] = e_add(f, t[k], e_mul(f, c, e_sub(f, t[k + half], t[k])));
    *len = half;
}

/* build_eq_table: src/provers/eq_factored.rs:217-239 (MSB, w[0] = top bit) */
static el *build_eq_table(const fctx *f, const el *w, size_t v, size_t *out_len) {
    size_t size = (size_t)1 << v;
    el *t = (el *)malloc(sizeof(el) * size);
    for (size_t i = 0; i < size; i++) t[i] = e_zero();
    t[0] = e_one();
    for (size_t j = 0; j < v; j++) {
        size_t stride = (size_t)1 << (v - 1 - j);
        size_t block = 2 * stride;
        size_t populated = (size_t)1 << j;
        for (size_t b = 0; b < populated; b++) {
            size_t base = b * block;
            el parent = t[base];
            t[base] = e_mul(f, parent, e_sub(f, e_one(), w[j]));
            t[base + stride] = e_mul(f, parent, w[j]);
        }
    }
    *out_len = size;
    return t;
}

static or_prover *alloc_prover(int kind, int ext, uint64_t w) {
    if (ext < 1 || ext > 3) return NULL;
    or_prover *p = (or_prover *)calloc(1, sizeof *p);
    p->kind = kind;
    p->f = mk_ctx(ext, ext == 1 ? 0 : w);
    return p;
}

/* MultilinearProver::new: src/provers/multilinear.rs:31-35 */
or_prover *or_prover_new_ml(int ext, uint64_t w, const uint64_t *evals, size_t n) {
    or_prover *p = alloc_prover(OR_ML, ext, w);
    if (!p) return NULL;
    p->deg = 1;
    p->ntab = 1;
    p->tab[0] = load_table(&p->f, evals, n, 0);
    p->len[0] = n;
    return p;
}
/* InnerProductProver::new: src/provers/inner_product.rs:33-36 */
or_prover *or_prover_new_ip(int ext, uint64_t w, const uint64_t *a, const uint64_t *b, size_t n) {
    or_prover *p = alloc_prover(OR_IP, ext, w);
    if (!p) return NULL;
    p->deg = 2;
    p->ntab = 2;
    p->tab[0] = load_table(&p->f, a, n, 0);
    p->tab[1] = load_table(&p->f, b, n, 0);
    p->len[0] = p->len[1] = n;
    return p;
}
static or_prover *new_product(int ext, uint64_t w, int k, const uint64_t *const *tables, size_t n, int lift) {
    if (k < 1 || k > MAXTAB) return NULL;
    or_prover *p = alloc_prover(OR_PRODUCT, ext, w);
    if (!p) return NULL;
    p->deg = k;
    p->ntab = k;
    for (int i = 0; i < k; i++) {
        p->tab[i] = load_table(&p->f, tables[i], n, lift);
        p->len[i] = n;
    }
    return p;
}
/* CoefficientProver::new(evaluator, [], pairwise): src/provers/coefficient.rs:28-40 */
or_prover *or_prover_new_product(int ext, uint64_t w, int k, const uint64_t *const *tables, size_t n) {
    return new_product(ext, w, k, tables, n, 1);
}
or_prover *or_prover_new_product_ext(int ext, uint64_t w, int k, const uint64_t *const *tables, size_t n) {
    return new_product(ext, w, k, tables, n, 0);
}
/* EqFactoredProver::new: src/provers/eq_factored.rs:90-114 */
or_prover *or_prover_new_eq(int ext, uint64_t w, const uint64_t *point, size_t v, const uint64_t *pe, size_t n) {
    if (v >= 40 || n > ((size_t)1 << v)) return NULL;
    or_prover *p = alloc_prover(OR_EQ, ext, w);
    if (!p) return NULL;
    p->deg = 2;
    p->ntab = 3;
    p->v = v;
    p->v_l = v / 2;
    size_t N = (size_t)1 << v;
    el *pt = load_table(&p->f, point, v, 0);
    el *tp = (el *)malloc(sizeof(el) * N);
    for (size_t i = 0; i < N; i++) tp[i] = i < n ? e_load(&p->f, pe + i * (size_t)ext) : e_zero();
    p->tab[0] = tp;
    p->len[0] = N;
    p->tab[1] = build_eq_table(&p->f, pt, p->v_l, &p->len[1]);
    p->tab[2] = build_eq_table(&p->f, pt + p->v_l, v - p->v_l, &p->len[2]);
    free(pt);
    return p;
}

/* GkrProver::new: src/provers/gkr.rs:58-87 (the witness is expanded into
 * broadcast tables w_b[b n + c] = W(b), w_c[b n + c] = W(c)) */
or_prover *or_prover_new_gkr(int ext, uint64_t w, const uint64_t *add, const uint64_t *mult, const uint64_t *wit,
                             size_t n) {
    or_prover *p = alloc_prover(OR_GKR, ext, w);
    if (!p) return NULL;
    const fctx *f = &p->f;
    size_t nbc = n * n;
    p->deg = 2;
    p->ntab = 4;
    p->tab[0] = load_table(f, add, nbc, 0);
    p->tab[1] = load_table(f, mult, nbc, 0);
    el *wv = load_table(f, wit, n, 0);
    p->tab[2] = (el *)malloc(sizeof(el) * (nbc ? nbc : 1));
    p->tab[3] = (el *)malloc(sizeof(el) * (nbc ? nbc : 1));
    for (size_t b = 0; b < n; b++)
        for (size_t c = 0; c < n; c++) {
            p->tab[2][b * n + c] = wv[b];
            p->tab[3][b * n + c] = wv[c];
        }
    free(wv);
    for (int t = 0; t < 4; t++) p->len[t] = nbc;
    return p;
}

/* LSB (pair-split) provers: MultilinearProverLSB::new
 * (src/provers/multilinear_lsb.rs:33-36), InnerProductProverLSB::new
 * (src/provers/inner_product_lsb.rs), CoefficientProverLSB::new with the
 * product-of-k evaluator (src/provers/coefficient_lsb.rs), base inputs lifted. */
or_prover *or_prover_new_ml_lsb(int ext, uint64_t w, const uint64_t *evals, size_t n) {
    or_prover *p = or_prover_new_ml(ext, w, evals, n);
    if (p) p->lsb = 1;
    return p;
}
or_prover *or_prover_new_ip_lsb(int ext, uint64_t w, const uint64_t *a, const uint64_t *b, size_t n) {
    or_prover *p = or_prover_new_ip(ext, w, a, b, n);
    if (p) p->lsb = 1;
    return p;
}
or_prover *or_prover_new_product_lsb(int ext, uint64_t w, int k, const uint64_t *const *tables, size_t n) {
    or_prover *p = or_prover_new_product(ext, w, k, tables, n);
    if (p) p->lsb = 1;
    return p;
}
or_prover *or_prover_new_product_lsb_ext(int ext, uint64_t w, int k, const uint64_t *const *tables, size_t n) {
    or_prover *p = or_prover_new_product_ext(ext, w, k, tables, n);
    if (p) p->lsb = 1;
    return p;
}

int or_prover_degree(const or_prover *p) { return p->deg; }

/* Horner: eval_poly_at, src/provers/coefficient.rs:249-259 */
static el horner(const fctx *f, const el *c, int n, el x) {
    if (n == 0) return e_zero();
    el r = c[n - 1];
    for (int i = n - 2; i >= 0; i--) r = e_add(f, e_mul(f, r, x), c[i]);
    return r;
}

static void round_ml(or_prover *p, el *out) {
    /* compute_sumcheck_polynomial: src/multilinear_sumcheck.rs:86-124; the
     * wire keeps s0 only (src/provers/multilinear.rs:60-68). */
    const fctx *f = &p->f;
    size_t n = p->len[0];
    el s0 = e_zero();
    if (n == 1) {
        s0 = p->tab[0][0];
    } else if (n > 1) {
        size_t half = next_pow2(n) >> 1;
        for (size_t k = 0; k < half; k++) s0 = e_add(f, s0, p->tab[0][k]);
    }
    out[0] = s0;
}

static void round_ip(or_prover *p, el *out) {
    /* src/provers/inner_product.rs:67-112 */
    const fctx *f = &p->f;
    size_t n = p->len[0];
    const el *a = p->tab[0], *b = p->tab[1];
    if (n <= 1) {
        out[0] = n == 1 ? e_mul(f, a[0], b[0]) : e_zero();
        out[1] = e_zero();
        return;
    }
    size_t half = next_pow2(n) >> 1;
    size_t paired = n - half;
    el q0 = e_zero(), qi = e_zero();
    for (size_t i = 0; i < paired; i++) {
        el al = a[i], ah = a[half + i], bl = b[i], bh = b[half + i];
        q0 = e_add(f, q0, e_mul(f, al, bl));
        qi = e_add(f, qi, e_mul(f, e_sub(f, ah, al), e_sub(f, bh, bl)));
    }
    el tail = e_zero();
    for (size_t i = paired; i < half; i++) tail = e_add(f, tail, e_mul(f, a[i], b[i]));
    out[0] = e_add(f, q0, tail);
    out[1] = e_add(f, qi, tail);
}

/* Product-of-k evaluator: accumulates the coefficients of
 * prod_i (l_i + x (h_i - l_i)), as the reference's TripleProductEval does for
 * k = 3 (tests/adversarial_verifier.rs:444-465). */
static void product_accumulate(const fctx *f, int k, const el *lo, const el *hi, el *coeffs) {
    el poly[MAXTAB + 1];
    poly[0] = e_one();
    for (int i = 1; i <= k; i++) poly[i] = e_zero();
    for (int i = 0; i < k; i++) {
        el d = e_sub(f, hi[i], lo[i]);
        /* poly *= (lo_i + d x) */
        for (int j = i + 1; j >= 1; j--) poly[j] = e_add(f, e_mul(f, poly[j], lo[i]), e_mul(f, poly[j - 1], d));
        poly[0] = e_mul(f, poly[0], lo[i]);
    }
    for (int j = 0; j <= k; j++) coeffs[j] = e_add(f, coeffs[j], poly[j]);
}

static void round_product(or_prover *p, el *out) {
    /* evaluate_coefficients + EvalsInfty: src/provers/coefficient.rs:76-103,272-294 */
    const fctx *f = &p->f;
    int k = p->ntab, d = p->deg;
    size_t half = p->len[0] / 2;
    el coeffs[MAXTAB + 1];
    for (int j = 0; j <= d; j++) coeffs[j] = e_zero();
    el lo[MAXTAB], hi[MAXTAB];
    for (size_t i = 0; i < half; i++) {
        for (int t = 0; t < k; t++) {
            lo[t] = p->tab[t][i];
            hi[t] = p->tab[t][i + half];
        }
        product_accumulate(f, k, lo, hi, coeffs);
    }
    if (d <= 1) {
        out[0] = horner(f, coeffs, d + 1, e_zero());
        return;
    }
    out[0] = horner(f, coeffs, d + 1, e_zero());
    out[1] = coeffs[d];
    for (int i = 2; i < d; i++) out[i] = horner(f, coeffs, d + 1, e_from_u64((uint64_t)i));
}

static void round_eq(or_prover *p, el *out) {
    const fctx *f = &p->f;
    const el *pp = p->tab[0], *el_ = p->tab[1], *er = p->tab[2];
    if (p->rounds_elapsed < p->v_l) {
        /* round_poly_left: src/provers/eq_factored.rs:139-172 */
        size_t eq_l_half = p->len[1] >> 1;
        size_t p_half = p->len[0] >> 1;
        size_t m = p->len[2];
        el q0 = e_zero(), qi = e_zero();
        for (size_t a = 0; a < eq_l_half; a++) {
            el in0 = e_zero(), ind = e_zero();
            for (size_t b = 0; b < m; b++) {
                el e = er[b], pl = pp[a * m + b], ph = pp[p_half + a * m + b];
                in0 = e_add(f, in0, e_mul(f, e, pl));
                ind = e_add(f, ind, e_mul(f, e, e_sub(f, ph, pl)));
            }
            el lo = el_[a], hi = el_[eq_l_half + a];
            q0 = e_add(f, q0, e_mul(f, lo, in0));
            qi = e_add(f, qi, e_mul(f, e_sub(f, hi, lo), ind));
        }
        out[0] = q0;
        out[1] = qi;
    } else {
        /* round_poly_right: src/provers/eq_factored.rs:178-209 */
        size_t n = p->len[0];
        el scalar = el_[0];
        if (n <= 1) {
            out[0] = n == 1 ? e_mul(f, e_mul(f, scalar, er[0]), pp[0]) : e_zero();
            out[1] = e_zero();
            return;
        }
        size_t half = n >> 1;
        el q0 = e_zero(), qi = e_zero();
        for (size_t i = 0; i < half; i++) {
            el e0 = er[i], e1 = er[half + i], pl = pp[i], ph = pp[half + i];
            q0 = e_add(f, q0, e_mul(f, e0, pl));
            qi = e_add(f, qi, e_mul(f, e_sub(f, e1, e0), e_sub(f, ph, pl)));
        }
        out[0] = e_mul(f, q0, scalar);
        out[1] = e_mul(f, qi, scalar);
    }
}

/* GKR gate polynomial at one point: add (wb + wc) + mult wb wc */
static inline el gkr_f(const fctx *f, el a, el m, el wb, el wc) {
    return e_add(f, e_mul(f, a, e_add(f, wb, wc)), e_mul(f, m, e_mul(f, wb, wc)));
}

static void round_gkr(or_prover *p, el *out) {
    /* src/provers/gkr.rs:104-182: q(0), q(1), q(2) over the half-split pairs
     * (unpaired tail: hi = 0), then q(inf) = (q(0) + q(2) - 2 q(1)) / 2 */
    const fctx *f = &p->f;
    size_t n = p->len[0];
    el *A = p->tab[0], *M = p->tab[1], *WB = p->tab[2], *WC = p->tab[3];
    if (n <= 1) {
        out[0] = n == 1 ? gkr_f(f, A[0], M[0], WB[0], WC[0]) : e_zero();
        out[1] = e_zero();
        return;
    }
    size_t half = next_pow2(n) >> 1;
    size_t paired = n - half;
    el q0 = e_zero(), q1 = e_zero(), q2 = e_zero();
    for (size_t i = 0; i < paired; i++) {
        el al = A[i], ah = A[half + i], ml = M[i], mh = M[half + i];
        el bl = WB[i], bh = WB[half + i], cl = WC[i], ch = WC[half + i];
        q0 = e_add(f, q0, gkr_f(f, al, ml, bl, cl));
        q1 = e_add(f, q1, gkr_f(f, ah, mh, bh, ch));
        el a2 = e_sub(f, e_add(f, ah, ah), al), m2 = e_sub(f, e_add(f, mh, mh), ml);
        el b2 = e_sub(f, e_add(f, bh, bh), bl), c2 = e_sub(f, e_add(f, ch, ch), cl);
        q2 = e_add(f, q2, gkr_f(f, a2, m2, b2, c2));
    }
    for (size_t i = paired; i < half; i++) {
        el al = A[i], ml = M[i], bl = WB[i], cl = WC[i];
        q0 = e_add(f, q0, gkr_f(f, al, ml, bl, cl));
        q2 = e_add(f, q2, e_sub(f, e_mul(f, al, e_add(f, bl, cl)), e_mul(f, ml, e_mul(f, bl, cl))));
    }
    el two = e_add(f, e_one(), e_one());
    out[0] = q0;
    out[1] = e_mul(f, e_sub(f, e_add(f, q0, q2), e_add(f, q1, q1)), e_inv(f, two));
}

/* fold_lsb: new[k] = f[2k] + w (f[2k+1] - f[2k]), k < n/2 (an odd trailing
 * element is dropped), src/provers/multilinear_lsb.rs:104-126,
 * inner_product_lsb.rs fold_lsb, src/reductions/pairwise.rs:38-73 */
static void fold_lsb(const fctx *f, el *t, size_t *len, el w) {
    size_t n = *len;
    if (n <= 1) return;
    size_t h = n / 2;
    for (size_t k = 0; k < h; k++) t[k] = e_add(f, t[2 * k], e_mul(f, w, e_sub(f, t[2 * k + 1], t[2 * k])));
    *len = h;
}

/* compute_lsb (src/provers/multilinear_lsb.rs:59-99): s0 = sum of the
 * even-indexed elements (an odd trailing element counts in s0) */
static void round_ml_lsb(or_prover *p, el *out) {
    const fctx *f = &p->f;
    el s0 = e_zero();
    for (size_t i = 0; i < p->len[0]; i += 2) s0 = e_add(f, s0, p->tab[0][i]);
    out[0] = s0;
}

/* compute_lsb (src/provers/inner_product_lsb.rs): q0 = sum a_even b_even,
 * q_inf = sum (a_odd - a_even)(b_odd - b_even); one element: [a0 b0, 0] */
static void round_ip_lsb(or_prover *p, el *out) {
    const fctx *f = &p->f;
    size_t n = p->len[0];
    const el *a = p->tab[0], *b = p->tab[1];
    out[0] = out[1] = e_zero();
    if (n == 1) out[0] = e_mul(f, a[0], b[0]);
    if (n <= 1) return;
    for (size_t i = 0; i + 1 < n; i += 2) {
        out[0] = e_add(f, out[0], e_mul(f, a[i], b[i]));
        out[1] = e_add(f, out[1], e_mul(f, e_sub(f, a[i + 1], a[i]), e_sub(f, b[i + 1], b[i])));
    }
}

/* evaluate_coefficients over pairs (2i, 2i+1), src/provers/coefficient_lsb.rs */
static void round_product_lsb(or_prover *p, el *out) {
    const fctx *f = &p->f;
    int k = p->ntab, d = p->deg;
    size_t npairs = p->len[0] / 2;
    el coeffs[MAXTAB + 1];
    for (int j = 0; j <= d; j++) coeffs[j] = e_zero();
    el lo[MAXTAB], hi[MAXTAB];
    for (size_t i = 0; i < npairs; i++) {
        for (int t = 0; t < k; t++) {
            lo[t] = p->tab[t][2 * i];
            hi[t] = p->tab[t][2 * i + 1];
        }
        product_accumulate(f, k, lo, hi, coeffs);
    }
    out[0] = horner(f, coeffs, d + 1, e_zero());
    if (d <= 1) return;
    out[1] = coeffs[d];
    for (int i = 2; i < d; i++) out[i] = horner(f, coeffs, d + 1, e_from_u64((uint64_t)i));
}

static void fold_all(or_prover *p, el c) {
    if (p->lsb) {
        for (int t = 0; t < p->ntab; t++) fold_lsb(&p->f, p->tab[t], &p->len[t], c);
        return;
    }
    switch (p->kind) {
    case OR_GKR: /* src/provers/gkr.rs:93-101, 185-190 */
    case OR_ML:
    case OR_IP:
        for (int t = 0; t < p->ntab; t++) fold_msb(&p->f, p->tab[t], &p->len[t], c);
        break;
    case OR_PRODUCT:
        /* reduce(): src/provers/coefficient.rs:106-125 */
        for (int t = 0; t < p->ntab; t++) fold_vec(&p->f, p->tab[t], &p->len[t], c);
        break;
    default:
        break;
    }
}

void or_prover_round(or_prover *p, const uint64_t *chal, uint64_t *out) {
    el res[MAXTAB + 1];
    el c = e_zero();
    if (chal) c = e_load(&p->f, chal);
    if (p->lsb) {
        if (chal) fold_all(p, c);
        if (p->kind == OR_ML)
            round_ml_lsb(p, res);
        else if (p->kind == OR_IP)
            round_ip_lsb(p, res);
        else
            round_product_lsb(p, res);
        int nw = p->deg < 1 ? 1 : p->deg;
        for (int i = 0; i < nw; i++) e_store(&p->f, res[i], out + (size_t)i * p->f.d);
        return;
    }
    switch (p->kind) {
    case OR_ML: /* src/provers/multilinear.rs:58-67 */
        if (chal) fold_all(p, c);
        round_ml(p, res);
        break;
    case OR_IP: /* src/provers/inner_product.rs:61-66 */
        if (chal) fold_all(p, c);
        round_ip(p, res);
        break;
    case OR_PRODUCT: /* src/provers/coefficient.rs:272-276 */
        if (chal) fold_all(p, c);
        round_product(p, res);
        break;
    case OR_GKR:
        if (chal) fold_all(p, c);
        round_gkr(p, res);
        break;
    case OR_EQ: /* src/provers/eq_factored.rs:249-270 */
        if (chal) {
            size_t prev_var = p->rounds_elapsed - 1;
            fold_msb(&p->f, p->tab[0], &p->len[0], c);
            if (prev_var < p->v_l)
                fold_msb(&p->f, p->tab[1], &p->len[1], c);
            else
                fold_msb(&p->f, p->tab[2], &p->len[2], c);
        }
        round_eq(p, res);
        p->rounds_elapsed++;
        break;
    }
    int nw = p->deg < 1 ? 1 : p->deg;
    for (int i = 0; i < nw; i++) e_store(&p->f, res[i], out + (size_t)i * p->f.d);
}

void or_prover_finalize(or_prover *p, const uint64_t *chal) {
    el c = e_load(&p->f, chal);
    if (p->kind == OR_EQ) {
        /* src/provers/eq_factored.rs:272-280 */
        size_t last_var = p->v - 1;
        fold_msb(&p->f, p->tab[0], &p->len[0], c);
        if (last_var < p->v_l)
            fold_msb(&p->f, p->tab[1], &p->len[1], c);
        else
            fold_msb(&p->f, p->tab[2], &p->len[2], c);
        return;
    }
    fold_all(p, c);
}

void or_prover_final_value(const or_prover *p, uint64_t *out) {
    const fctx *f = &p->f;
    el r = e_zero();
    switch (p->kind) {
    case OR_ML: /* src/provers/multilinear.rs:73-79 */
        if (p->len[0] == 1) r = p->tab[0][0];
        break;
    case OR_IP: /* src/provers/inner_product.rs:120-126 */
        if (p->len[0] == 1) r = e_mul(f, p->tab[0][0], p->tab[1][0]);
        break;
    case OR_PRODUCT: { /* src/provers/coefficient.rs:300-329 */
        int k = p->ntab;
        if (k == 1 && p->len[0] == 1) {
            r = p->tab[0][0];
            break;
        }
        el coeffs[MAXTAB + 1], lo[MAXTAB], hi[MAXTAB];
        for (int j = 0; j <= p->deg; j++) coeffs[j] = e_zero();
        for (int t = 0; t < k; t++) {
            lo[t] = p->len[t] == 1 ? p->tab[t][0] : e_zero();
            hi[t] = e_zero();
        }
        product_accumulate(f, k, lo, hi, coeffs);
        r = e_add(f, horner(f, coeffs, p->deg + 1, e_zero()), horner(f, coeffs, p->deg + 1, e_one()));
        break;
    }
    case OR_EQ: /* src/provers/eq_factored.rs:282-288 */
        if (p->len[0] == 1) r = e_mul(f, e_mul(f, p->tab[0][0], p->tab[1][0]), p->tab[2][0]);
        break;
    case OR_GKR: /* src/provers/gkr.rs:192-201 */
        if (p->len[0] == 1) r = gkr_f(f, p->tab[0][0], p->tab[1][0], p->tab[2][0], p->tab[3][0]);
        break;
    }
    e_store(f, r, out);
}

size_t or_prover_table_len(const or_prover *p, int t) { return (t >= 0 && t < p->ntab) ? p->len[t] : 0; }
void or_prover_read_table(const or_prover *p, int t, uint64_t *out) {
    if (t < 0 || t >= p->ntab) return;
    for (size_t i = 0; i < p->len[t]; i++) e_store(&p->f, p->tab[t][i], out + i * (size_t)p->f.d);
}
void or_prover_free(or_prover *p) {
    if (!p) return;
    for (int t = 0; t < p->ntab; t++) free(p->tab[t]);
    free(p);
}

/* ------------------------------------------------------------------------ */
/* Runner: src/runner.rs:29-69 (no hook: the reference's noop hook)           */
/* ------------------------------------------------------------------------ */
void or_sumcheck(or_prover *p, size_t num_rounds, or_transcript *t, uint64_t *round_polys,
                 uint64_t *challenges, uint64_t *fin) {
    int d = p->f.d;
    int nw = p->deg < 1 ? 1 : p->deg;
    uint64_t prev[3];
    int have_prev = 0;
    for (size_t r = 0; r < num_rounds; r++) {
        uint64_t *evals = round_polys + r * (size_t)nw * d;
        or_prover_round(p, have_prev ? prev : NULL, evals);
        for (int i = 0; i < nw; i++) or_tr_send(t, d, evals + (size_t)i * d);
        or_tr_challenge(t, d, prev);
        memcpy(challenges + r * (size_t)d, prev, sizeof(uint64_t) * d);
        have_prev = 1;
    }
    if (have_prev) or_prover_finalize(p, prev);
    or_prover_final_value(p, fin);
}

/* ------------------------------------------------------------------------ */
/* Verifier: sumcheck_verify, evaluate_from_evals, barycentric_weight:        */
/* src/verifier.rs:61-200                                                    */
/* ------------------------------------------------------------------------ */
static el barycentric_weight(const fctx *f, size_t i, size_t d) {
    el w = e_one();
    for (size_t j = 0; j < d; j++) {
        if (j == i) continue;
        long long diff = (long long)i - (long long)j;
        if (diff > 0)
            w = e_mul(f, w, e_from_u64((uint64_t)diff));
        else
            w = e_mul(f, w, e_sub(f, e_zero(), e_from_u64((uint64_t)(-diff))));
    }
    return w;
}
static el evaluate_from_evals(const fctx *f, const el *evals, size_t d, el r) {
    if (d == 0) return e_zero();
    if (d == 1) return evals[0];
    if (d == 2) return e_add(f, evals[0], e_mul(f, r, e_sub(f, evals[1], evals[0])));
    el rm[MAXTAB + 2], prefix[MAXTAB + 3], suffix[MAXTAB + 3];
    for (size_t j = 0; j < d; j++) rm[j] = e_sub(f, r, e_from_u64(j));
    prefix[0] = e_one();
    for (size_t i = 0; i < d; i++) prefix[i + 1] = e_mul(f, prefix[i], rm[i]);
    suffix[d] = e_one();
    for (size_t i = d; i-- > 0;) suffix[i] = e_mul(f, suffix[i + 1], rm[i]);
    el result = e_zero();
    for (size_t i = 0; i < d; i++) {
        el num = e_mul(f, prefix[i], suffix[i + 1]);
        el den = barycentric_weight(f, i, d);
        if (!e_is_zero(f, den)) result = e_add(f, result, e_mul(f, e_mul(f, evals[i], num), e_inv(f, den)));
    }
    return result;
}

int or_sumcheck_verify(int ext, uint64_t w, const uint64_t *claimed_sum, size_t degree, size_t num_rounds,
                       or_transcript *t, const uint64_t *proof, size_t proof_len, uint64_t *challenges,
                       uint64_t *final_claim) {
    if (ext < 1 || ext > 3 || degree > MAXTAB) return -1;
    fctx f = mk_ctx(ext, ext == 1 ? 0 : w);
    el claim = e_load(&f, claimed_sum);
    size_t d = degree, cursor = 0;
    for (size_t round = 0; round < num_rounds; round++) {
        size_t n_wire = d > 1 ? d : 1;
        el recv[MAXTAB + 1];
        for (size_t i = 0; i < n_wire; i++) {
            if (cursor + (size_t)ext > proof_len) return (int)(1 + round);
            recv[i] = e_load(&f, proof + cursor);
            or_tr_send(t, ext, proof + cursor);
            cursor += (size_t)ext;
        }
        el h0 = recv[0];
        el h1 = e_sub(&f, claim, h0);
        uint64_t rc[3];
        or_tr_challenge(t, ext, rc);
        memcpy(challenges + round * (size_t)ext, rc, sizeof(uint64_t) * ext);
        el r = e_load(&f, rc);
        if (d == 0) {
            claim = h0;
            continue;
        }
        el h_inf = d >= 2 ? recv[1] : e_sub(&f, h1, h0);
        el q[MAXTAB + 1];
        size_t nq = 0;
        q[nq++] = h0;
        q[nq++] = e_sub(&f, h1, h_inf);
        for (size_t i = 2; i < d; i++) {
            el i_f = e_from_u64(i), i_d = e_one();
            for (size_t j = 0; j < d; j++) i_d = e_mul(&f, i_d, i_f);
            q[nq++] = e_sub(&f, recv[i], e_mul(&f, h_inf, i_d));
        }
        el q_r = evaluate_from_evals(&f, q, nq, r);
        el r_d = e_one();
        for (size_t j = 0; j < d; j++) r_d = e_mul(&f, r_d, r);
        claim = e_add(&f, q_r, e_mul(&f, h_inf, r_d));
    }
    e_store(&f, claim, final_claim);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CPU baseline: the reference InnerProductProver round structure on base    */
/* Montgomery words.  Folds: ip::fold (src/inner_product_sumcheck.rs:143-176, */
/* rayon leaves of 4096 elements); evaluate: serial loop                     */
/* (src/provers/inner_product.rs:94-110).                                    */
/* ------------------------------------------------------------------------ */
static void bench_fold(uint64_t *v, size_t *len, uint64_t w, int threads) {
    size_t n = *len;
    if (n <= 1) return;
    size_t half = next_pow2(n) >> 1, paired = n - half;
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 4096) num_threads(threads) if (threads > 1 && paired > 4096)
#endif
    for (size_t k = 0; k < paired; k++) v[k] = m_add(v[k], m_mul(m_sub(v[k + half], v[k]), w));
    uint64_t omw = m_sub(MONT_ONE, w);
    for (size_t k = paired; k < half; k++) v[k] = m_mul(v[k], omw);
    *len = half;
}
static void bench_eval(const uint64_t *a, const uint64_t *b, size_t n, uint64_t *q0o, uint64_t *qio) {
    if (n <= 1) {
        *q0o = n == 1 ? m_mul(a[0], b[0]) : 0;
        *qio = 0;
        return;
    }
    size_t half = next_pow2(n) >> 1, paired = n - half;
    uint64_t q0 = 0, qi = 0;
    for (size_t i = 0; i < paired; i++) {
        uint64_t al = a[i], ah = a[half + i], bl = b[i], bh = b[half + i];
        q0 = m_add(q0, m_mul(al, bl));
        qi = m_add(qi, m_mul(m_sub(ah, al), m_sub(bh, bl)));
    }
    uint64_t tail = 0;
    for (size_t i = paired; i < half; i++) tail = m_add(tail, m_mul(a[i], b[i]));
    *q0o = m_add(q0, tail);
    *qio = m_add(qi, tail);
}
static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}
/* compute_sumcheck_polynomial (src/inner_product_sumcheck.rs:88-141) with
 * the rayon split of its pairs: q(0) = sum a_lo b_lo, q(inf) = sum
 * (a_hi - a_lo)(b_hi - b_lo) over the paired prefix, plus the unpaired tail
 * (hi = 0) on both values. */
static void bench_eval_par(const uint64_t *a, const uint64_t *b, size_t n, int threads, uint64_t *q0o, uint64_t *qio) {
    if (n <= 1 || threads <= 1) {
        bench_eval(a, b, n, q0o, qio);
        return;
    }
    size_t half = next_pow2(n) >> 1, paired = n - half;
    uint64_t q0 = 0, qi = 0;
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads) if (paired > 1024)
#endif
    {
        uint64_t s0 = 0, si = 0;
#ifdef _OPENMP
#pragma omp for schedule(static, 1024)
#endif
        for (size_t i = 0; i < paired; i++) {
            uint64_t al = a[i], ah = a[half + i], bl = b[i], bh = b[half + i];
            s0 = m_add(s0, m_mul(al, bl));
            si = m_add(si, m_mul(m_sub(ah, al), m_sub(bh, bl)));
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            q0 = m_add(q0, s0);
            qi = m_add(qi, si);
        }
    }
    uint64_t tail = 0;
    for (size_t i = paired; i < half; i++) tail = m_add(tail, m_mul(a[i], b[i]));
    *q0o = m_add(q0, tail);
    *qio = m_add(qi, tail);
}

/* fused_fold_and_compute_polynomial (src/inner_product_sumcheck.rs:199-293):
 * one sweep over the quadruples (k, k + L/4, k + L/2, k + 3L/4) folds both
 * tables with w into [0, L/2) and accumulates the next round's
 * (q(0), q(inf)); the reference splits the quadruples with rayon joins down
 * to leaves of 1024 (workload_size: 32 KiB of 8-byte elements / 4).  Small
 * or non-power-of-two tables take the unfused fold + evaluate. */
static void bench_fused_round(uint64_t *a, uint64_t *b, size_t *len, uint64_t w, int threads, uint64_t *q0o,
                              uint64_t *qio) {
    size_t l = *len;
    if (l < 4 || (l & (l - 1))) {
        size_t lb = l;
        bench_fold(a, len, w, threads);
        bench_fold(b, &lb, w, threads);
        bench_eval_par(a, b, *len, threads, q0o, qio);
        return;
    }
    size_t q = l / 4;
    uint64_t c0 = 0, c2 = 0;
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads) if (threads > 1 && q > 1024)
#endif
    {
        uint64_t s0 = 0, s2 = 0;
#ifdef _OPENMP
#pragma omp for schedule(static, 1024)
#endif
        for (size_t k = 0; k < q; k++) {
            uint64_t x0 = a[k], x1 = a[k + q], x2 = a[k + 2 * q], x3 = a[k + 3 * q];
            uint64_t y0 = b[k], y1 = b[k + q], y2 = b[k + 2 * q], y3 = b[k + 3 * q];
            uint64_t na_lo = m_add(x0, m_mul(m_sub(x2, x0), w)), na_hi = m_add(x1, m_mul(m_sub(x3, x1), w));
            uint64_t nb_lo = m_add(y0, m_mul(m_sub(y2, y0), w)), nb_hi = m_add(y1, m_mul(m_sub(y3, y1), w));
            a[k] = na_lo;
            a[k + q] = na_hi;
            b[k] = nb_lo;
            b[k + q] = nb_hi;
            s0 = m_add(s0, m_mul(na_lo, nb_lo));
            s2 = m_add(s2, m_mul(m_sub(na_hi, na_lo), m_sub(nb_hi, nb_lo)));
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            c0 = m_add(c0, s0);
            c2 = m_add(c2, s2);
        }
    }
    *len = l / 2;
    *q0o = c0;
    *qio = c2;
}

double or_bench_ip_ex(uint64_t *a, uint64_t *b, size_t n, size_t num_rounds, int threads, int fused,
                      const uint8_t *domsep, size_t domsep_len, uint64_t *round_polys, uint64_t *challenges,
                      uint64_t *fin) {
    for (size_t i = 0; i < n; i++) {
        a[i] = m_to(a[i]);
        b[i] = m_to(b[i]);
    }
    or_transcript *t = or_tr_new_shake(domsep, domsep_len);
    double t0 = now_s();
    size_t la = n, lb = n;
    uint64_t prev = 0;
    int have = 0;
    for (size_t r = 0; r < num_rounds; r++) {
        uint64_t q[2];
        if (fused) {
            /* inner_product_sumcheck_partial (src/inner_product_sumcheck.rs:305-360):
             * the round-(i-1) fold fused into the round-i evaluation */
            if (have) {
                bench_fused_round(a, b, &la, prev, threads, &q[0], &q[1]);
                lb = la;
            } else {
                bench_eval_par(a, b, la, threads, &q[0], &q[1]);
            }
        } else {
            /* InnerProductProver::round (src/provers/inner_product.rs:61-113):
             * two parallel folds, then the serial evaluate */
            if (have) {
                bench_fold(a, &la, prev, threads);
                bench_fold(b, &lb, prev, threads);
            }
            bench_eval(a, b, la, &q[0], &q[1]);
        }
        q[0] = m_from(q[0]);
        q[1] = m_from(q[1]);
        or_tr_send(t, 1, &q[0]);
        or_tr_send(t, 1, &q[1]);
        round_polys[2 * r] = q[0];
        round_polys[2 * r + 1] = q[1];
        uint64_t c;
        or_tr_challenge(t, 1, &c);
        challenges[r] = c;
        prev = m_to(c);
        have = 1;
    }
    if (have) {
        bench_fold(a, &la, prev, threads);
        bench_fold(b, &lb, prev, threads);
    }
    *fin = la == 1 ? m_from(m_mul(a[0], b[0])) : 0;
    double el_s = now_s() - t0;
    or_tr_free(t);
    return el_s;
}

/* MultilinearProver over the base field as the reference runs it with the
 * parallel feature (src/provers/multilinear.rs:38-58 over
 * src/multilinear_sumcheck.rs:84-254): round 0 is
 * compute_sumcheck_polynomial (the pairs split by rayon joins), later rounds
 * fused_fold_and_compute_polynomial (one sweep over the quadruples
 * (k, k + L/4, k + L/2, k + 3L/4), split down to leaves of 32 KiB), the last
 * challenge folds with fold; each round sends [s0].  OpenMP static blocks of
 * 1024 stand in for the joins (field addition is exact, so the split does
 * not change a bit).  Non-power-of-two tables take fold + evaluate. */
static void bench_ml_eval(const uint64_t *v, size_t n, int threads, uint64_t *s0o, uint64_t *s1o) {
    if (n == 0) {
        *s0o = *s1o = 0;
        return;
    }
    if (n == 1) {
        *s0o = v[0];
        *s1o = 0;
        return;
    }
    size_t half = next_pow2(n) >> 1, paired = n - half;
    uint64_t s0 = 0, s1 = 0;
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads) if (threads > 1 && paired > 1024)
#endif
    {
        uint64_t a0 = 0, a1 = 0;
#ifdef _OPENMP
#pragma omp for schedule(static, 1024)
#endif
        for (size_t i = 0; i < paired; i++) {
            a0 = m_add(a0, v[i]);
            a1 = m_add(a1, v[half + i]);
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            s0 = m_add(s0, a0);
            s1 = m_add(s1, a1);
        }
    }
    for (size_t i = paired; i < half; i++) s0 = m_add(s0, v[i]);
    *s0o = s0;
    *s1o = s1;
}
static void bench_ml_fused(uint64_t *v, size_t *len, uint64_t w, int threads, uint64_t *s0o, uint64_t *s1o) {
    size_t l = *len;
    if (l < 4 || (l & (l - 1))) {
        bench_fold(v, len, w, threads);
        bench_ml_eval(v, *len, threads, s0o, s1o);
        return;
    }
    size_t q = l / 4;
    uint64_t s0 = 0, s1 = 0;
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads) if (threads > 1 && q > 1024)
#endif
    {
        uint64_t a0 = 0, a1 = 0;
#ifdef _OPENMP
#pragma omp for schedule(static, 1024)
#endif
        for (size_t k = 0; k < q; k++) {
            uint64_t x0 = v[k], x1 = v[k + q], x2 = v[k + 2 * q], x3 = v[k + 3 * q];
            uint64_t lo = m_add(x0, m_mul(m_sub(x2, x0), w)), hi = m_add(x1, m_mul(m_sub(x3, x1), w));
            v[k] = lo;
            v[k + q] = hi;
            a0 = m_add(a0, lo);
            a1 = m_add(a1, hi);
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            s0 = m_add(s0, a0);
            s1 = m_add(s1, a1);
        }
    }
    *len = l / 2;
    *s0o = s0;
    *s1o = s1;
}

double or_bench_ml_ex(uint64_t *v, size_t n, size_t num_rounds, int threads, const uint8_t *domsep,
                      size_t domsep_len, uint64_t *round_polys, uint64_t *challenges, uint64_t *fin) {
    for (size_t i = 0; i < n; i++) v[i] = m_to(v[i]);
    or_transcript *t = or_tr_new_shake(domsep, domsep_len);
    double t0 = now_s();
    size_t len = n;
    uint64_t prev = 0;
    int have = 0;
    for (size_t r = 0; r < num_rounds; r++) {
        uint64_t s0, s1;
        if (have)
            bench_ml_fused(v, &len, prev, threads, &s0, &s1);
        else
            bench_ml_eval(v, len, threads, &s0, &s1);
        s0 = m_from(s0);
        or_tr_send(t, 1, &s0);
        round_polys[r] = s0;
        uint64_t c;
        or_tr_challenge(t, 1, &c);
        challenges[r] = c;
        prev = m_to(c);
        have = 1;
    }
    if (have) bench_fold(v, &len, prev, threads);
    *fin = len == 1 ? m_from(v[0]) : 0;
    double el_s = now_s() - t0;
    or_tr_free(t);
    return el_s;
}

double or_bench_ip(uint64_t *a, uint64_t *b, size_t n, size_t num_rounds, int threads, uint64_t *round_polys,
                   uint64_t *challenges, uint64_t *fin) {
    static const char dom[] = "effsc-hip/bench/ip";
    return or_bench_ip_ex(a, b, n, num_rounds, threads, 0, (const uint8_t *)dom, sizeof(dom) - 1, round_polys,
                          challenges, fin);
}
