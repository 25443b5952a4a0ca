// Keccak-f[1600] (FIPS 202) and the SHAKE256-based Fiat-Shamir transcript,
// usable from host and device code.
//
// Transcript stream (the reference's transcript is the external spongefish
// crate, src/transcript/spongefish.rs:17-44, whose wire encoding is not part
// of the reference; this library fixes its own).  Every item is a whole number
// of little-endian u64 words, so the sponge absorbs whole lanes:
//   "effsc-hip/fs/v3\0" (2 words) || len(domsep) || domsep zero-padded to 8 bytes
//   send(x)      : word 1, then each canonical component
//   challenge(d) : word 2 is appended; the first 16 d bytes of
//                  SHAKE256(stream) give d little-endian u128 values mod p;
//                  then the stream is padded to the next 136-byte boundary
//                  with SHAKE's padding (0x1F, zeros, 0x80).
// The sponge is kept incrementally.  Because the padding stays in the stream,
// a challenge pads and permutes the sponge itself and reads its output from
// the permuted state: one permutation per challenge, plus one per 136 bytes
// absorbed within a round.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field.hpp"

namespace effsc_hip {

#define SKHD __host__ __device__ __forceinline__

SKHD uint64_t rotl64(uint64_t x, int r) { return r ? (x << r) | (x >> (64 - r)) : x; }

// Round constants and rotation offsets of Keccak-f[1600] (FIPS 202, 3.2).
SKHD uint64_t keccak_rc(int i) {
  const uint64_t rc[24] = {
      0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
      0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
      0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
      0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
      0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
      0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
  return rc[i];
}

__host__ __device__ inline void keccak_f1600(uint64_t* s) {
  const int rot[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
  for (int round = 0; round < 24; round++) {
    uint64_t c[5], b[25];
#pragma unroll
    for (int x = 0; x < 5; x++) c[x] = s[x] ^ s[x + 5] ^ s[x + 10] ^ s[x + 15] ^ s[x + 20];
#pragma unroll
    for (int x = 0; x < 5; x++) {
      uint64_t d = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
#pragma unroll
      for (int y = 0; y < 25; y += 5) s[x + y] ^= d;
    }
    // rho + pi: lane (x, y) moves to (y, 2x + 3y)
#pragma unroll
    for (int x = 0; x < 5; x++)
#pragma unroll
      for (int y = 0; y < 5; y++) b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(s[x + 5 * y], rot[x + 5 * y]);
#pragma unroll
    for (int y = 0; y < 25; y += 5)
#pragma unroll
      for (int x = 0; x < 5; x++) s[x + y] = b[x + y] ^ (~b[(x + 1) % 5 + y] & b[(x + 2) % 5 + y]);
    s[0] ^= keccak_rc(round);
  }
}

constexpr int kShakeRate = 136;

struct Sponge {
  uint64_t st[25];
  uint32_t pos;

  SKHD void init() {
    for (int i = 0; i < 25; i++) st[i] = 0;
    pos = 0;
  }
  SKHD void absorb_byte(uint8_t b) {
    st[pos >> 3] ^= (uint64_t)b << (8 * (pos & 7));
    if (++pos == kShakeRate) {
      keccak_f1600(st);
      pos = 0;
    }
  }
  SKHD void absorb(const uint8_t* m, size_t n) {
    for (size_t i = 0; i < n; i++) absorb_byte(m[i]);
  }
  SKHD void absorb_u64(uint64_t x) {
    for (int j = 0; j < 8; j++) absorb_byte((uint8_t)(x >> (8 * j)));
  }
};

struct FsTranscript {
  Sponge sp;

  SKHD void absorb_word(uint64_t x) {
    sp.st[sp.pos >> 3] ^= x;
    sp.pos += 8;
    if (sp.pos == kShakeRate) {
      keccak_f1600(sp.st);
      sp.pos = 0;
    }
  }
  SKHD void init(const uint8_t* domsep, size_t len) {
    const char tag[16] = "effsc-hip/fs/v3";
    sp.init();
    sp.absorb(reinterpret_cast<const uint8_t*>(tag), 16);
    sp.absorb_u64((uint64_t)len);
    sp.absorb(domsep, len);
    for (size_t i = len; i % 8; i++) sp.absorb_byte(0);
  }
  SKHD void send(int d, const uint64_t* v) {
    absorb_word(1);
    for (int i = 0; i < d; i++) absorb_word(gl::canon(v[i]));
  }
  SKHD void challenge(int d, uint64_t* out) {
    absorb_word(2);
    sp.st[sp.pos >> 3] ^= (uint64_t)0x1F << (8 * (sp.pos & 7));
    sp.st[(kShakeRate - 1) >> 3] ^= (uint64_t)0x80 << (8 * ((kShakeRate - 1) & 7));
    keccak_f1600(sp.st);
    sp.pos = 0;
    for (int i = 0; i < d; i++) out[i] = gl::reduce128(sp.st[2 * i], sp.st[2 * i + 1]);
  }
};

}  // namespace effsc_hip
