/*
 * mont_ref.c -- plain-C restatement of the limb-level arithmetic the GPU
 * kernels perform. TEST INFRASTRUCTURE ONLY (the checker; never linked into
 * the product library).
 *
 * Numbers are little-endian arrays of 32-bit limbs, L limbs, R = 2^(32L).
 *  - mr_mont_mul: Montgomery product a*b*R^-1 mod N (Montgomery 1985), CIOS
 *    ordering (Koc, Acar, Kaliski 1996), with the final conditional
 *    subtraction, so outputs are canonical.
 *  - mr_modexp: left-to-right binary exponentiation in the Montgomery domain
 *    (HAC Alg. 14.79) -- deliberately a different exponent schedule from the
 *    product's sliding window, so a schedule mistake cannot cancel out.
 *  - mr_apply_schedule: evaluates an exponent schedule in the product's packed
 *    format ((squarings << 16) | (odd index + 1); entry 0 = first window
 *    index) with this file's Montgomery product, so the product's host-side
 *    schedule builder can be checked on the CPU against x^e.
 *  - mr_exact_div: q = X / N for X = q*N (q < R) by right-to-left (Hensel)
 *    division, the quotient the GPU recovers from REDC digits.
 * The reference itself uses CGBN mul_wide + rem_wide on the GPU
 * (paillier_GPU_client.cu:332-360) and GMP on the CPU; both compute the same
 * integers, which the tests check against CPython pow.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MR_MAXL 256

static uint32_t mr_n0inv_(uint32_t n0) {
  uint32_t inv = 1;
  for (int i = 0; i < 5; ++i) inv *= 2u - n0 * inv;
  return 0u - inv;
}

uint32_t mr_n0inv(uint32_t n0) { return mr_n0inv_(n0); }

/* returns -1, 0, 1 */
static int mr_cmp(const uint32_t* a, const uint32_t* b, int L) {
  for (int i = L - 1; i >= 0; --i)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}

static uint32_t mr_sub(uint32_t* d, const uint32_t* a, const uint32_t* b, int L) {
  uint64_t br = 0;
  for (int i = 0; i < L; ++i) {
    uint64_t t = (uint64_t)a[i] - b[i] - br;
    d[i] = (uint32_t)t;
    br = (t >> 63) & 1u;
  }
  return (uint32_t)br;
}

void mr_mont_mul(const uint32_t* a, const uint32_t* b, const uint32_t* N, uint32_t n0inv, int L,
                 uint32_t* out) {
  uint32_t t[MR_MAXL + 2];
  memset(t, 0, sizeof(uint32_t) * (L + 2));
  for (int i = 0; i < L; ++i) {
    uint64_t c = 0;
    for (int j = 0; j < L; ++j) {
      uint64_t s = (uint64_t)a[i] * b[j] + t[j] + c;
      t[j] = (uint32_t)s;
      c = s >> 32;
    }
    uint64_t s = (uint64_t)t[L] + c;
    t[L] = (uint32_t)s;
    t[L + 1] = (uint32_t)(s >> 32);
    uint32_t m = t[0] * n0inv;
    c = ((uint64_t)m * N[0] + t[0]) >> 32;
    for (int j = 1; j < L; ++j) {
      uint64_t s2 = (uint64_t)m * N[j] + t[j] + c;
      t[j - 1] = (uint32_t)s2;
      c = s2 >> 32;
    }
    s = (uint64_t)t[L] + c;
    t[L - 1] = (uint32_t)s;
    t[L] = t[L + 1] + (uint32_t)(s >> 32);
  }
  uint32_t d[MR_MAXL];
  uint32_t borrow = mr_sub(d, t, N, L);
  if (t[L] || !borrow)
    memcpy(out, d, sizeof(uint32_t) * L);
  else
    memcpy(out, t, sizeof(uint32_t) * L);
}

/* R^2 mod N by doubling (2*32L doublings from 1) */
static void mr_r2(const uint32_t* N, int L, uint32_t* r2) {
  uint32_t x[MR_MAXL];
  memset(x, 0, sizeof(uint32_t) * L);
  x[0] = 1;
  for (int k = 0; k < 64 * L; ++k) {
    uint32_t carry = 0;
    for (int i = 0; i < L; ++i) {
      uint32_t nc = x[i] >> 31;
      x[i] = (x[i] << 1) | carry;
      carry = nc;
    }
    if (carry || mr_cmp(x, N, L) >= 0) mr_sub(x, x, N, L);
  }
  memcpy(r2, x, sizeof(uint32_t) * L);
}

void mr_r2_mod(const uint32_t* N, int L, uint32_t* r2) { mr_r2(N, L, r2); }

/* base^e mod N, base < N, N odd. Returns 0 on success. */
int mr_modexp(const uint32_t* base, const uint32_t* e, int e_limbs, const uint32_t* N, int L, uint32_t* out) {
  if (L <= 0 || L > MR_MAXL || !(N[0] & 1u)) return 1;
  const uint32_t n0 = mr_n0inv_(N[0]);
  uint32_t r2[MR_MAXL], x[MR_MAXL], acc[MR_MAXL], one[MR_MAXL];
  mr_r2(N, L, r2);
  mr_mont_mul(base, r2, N, n0, L, x); /* x = base*R */
  memset(one, 0, sizeof(uint32_t) * L);
  one[0] = 1;
  mr_mont_mul(one, r2, N, n0, L, acc); /* acc = R mod N */
  for (int bit = 32 * e_limbs - 1; bit >= 0; --bit) {
    mr_mont_mul(acc, acc, N, n0, L, acc);
    if ((e[bit / 32] >> (bit % 32)) & 1u) mr_mont_mul(acc, x, N, n0, L, acc);
  }
  mr_mont_mul(acc, one, N, n0, L, out);
  return 0;
}

/* Evaluate a packed exponent schedule with window w on base (canonical out). */
int mr_apply_schedule(const uint32_t* base, const uint32_t* sched, int nsched, int w, const uint32_t* N, int L,
                      uint32_t* out) {
  if (L <= 0 || L > MR_MAXL || nsched <= 0 || w < 1 || w > 8) return 1;
  const uint32_t n0 = mr_n0inv_(N[0]);
  const int tentries = 1 << (w - 1);
  uint32_t r2[MR_MAXL], x2[MR_MAXL], acc[MR_MAXL], one[MR_MAXL];
  uint32_t* T = (uint32_t*)malloc(sizeof(uint32_t) * L * tentries);
  if (!T) return 2;
  mr_r2(N, L, r2);
  mr_mont_mul(base, r2, N, n0, L, T);
  mr_mont_mul(T, T, N, n0, L, x2);
  for (int j = 1; j < tentries; ++j) mr_mont_mul(T + (j - 1) * L, x2, N, n0, L, T + j * L);
  if (sched[0] >= (uint32_t)tentries) {
    free(T);
    return 3;
  }
  memcpy(acc, T + sched[0] * L, sizeof(uint32_t) * L);
  for (int s = 1; s < nsched; ++s) {
    const uint32_t sq = sched[s] >> 16, idx = sched[s] & 0xFFFFu;
    for (uint32_t i = 0; i < sq; ++i) mr_mont_mul(acc, acc, N, n0, L, acc);
    if (idx) {
      if (idx - 1 >= (uint32_t)tentries) {
        free(T);
        return 3;
      }
      mr_mont_mul(acc, T + (idx - 1) * L, N, n0, L, acc);
    }
  }
  memset(one, 0, sizeof(uint32_t) * L);
  one[0] = 1;
  mr_mont_mul(acc, one, N, n0, L, out);
  free(T);
  return 0;
}

/* q = X / N for X = q*N exactly, q < 2^(32L): q_i = x_i * N^-1 mod 2^32, X -= q_i N 2^(32i) */
void mr_exact_div(const uint32_t* X, const uint32_t* N, int L, uint32_t* q) {
  uint32_t x[MR_MAXL + 1];
  memcpy(x, X, sizeof(uint32_t) * L);
  x[L] = 0;
  const uint32_t ninv = 0u - mr_n0inv_(N[0]); /* N^-1 mod 2^32 */
  for (int i = 0; i < L; ++i) {
    const uint32_t qi = x[i] * ninv;
    q[i] = qi;
    uint64_t br = 0;
    for (int j = 0; i + j < L; ++j) {
      uint64_t p = (uint64_t)qi * N[j] + br;
      uint64_t d = (uint64_t)x[i + j] - (uint32_t)p;
      x[i + j] = (uint32_t)d;
      br = (p >> 32) + ((d >> 63) & 1u);
    }
  }
}
