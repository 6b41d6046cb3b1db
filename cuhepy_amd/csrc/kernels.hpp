// Batched Paillier kernels for gfx950 built on the lane-group Montgomery core.
//
// Device data layout ("limb-interleaved"): a batch of `count` bignums of L
// limbs is stored as L rows of `stride` u32 words, limb j of item i at
// [j*stride + i] (stride = count rounded up to 64). A wave that loads limb j
// of consecutive items touches consecutive words, so HBM loads coalesce.
//
// Per-modulus constants live in a ModDev block in device memory; every kernel
// reads the modulus into registers once and then works out of VGPRs/LDS.
#pragma once
#include "mont.hpp"

namespace chx {

constexpr int LMAX = 128;
constexpr int WAVES_PER_BLOCK = 2;
constexpr int BLOCK_THREADS = 64 * WAVES_PER_BLOCK;

// Montgomery constants of one odd modulus N of `L` limbs, R = 2^(32L).
struct ModDev {
  u32 L;
  u32 n0inv;  // -N^-1 mod 2^32
  u32 bits;   // bit length of N
  u32 pad;
  u32 N[LMAX];
  u32 R1[LMAX];  // R mod N      (Montgomery form of 1)
  u32 R2[LMAX];  // R^2 mod N    (to Montgomery form)
};

// A plain value of up to LMAX limbs (e.g. a constant already in Montgomery form).
struct ValDev {
  u32 v[LMAX];
};

// Everything a Paillier key needs on the device.
//   width W = limbs of n^2 (64 or 128); n is zero-padded to W limbs.
struct PaillierDev {
  ModDev n2;      // n^2
  ModDev n;       // n, padded to W limbs (exact division by n, arithmetic mod n)
  ValDev nR;      // n * R mod n^2     (m*n = MontMul(m, nR))
  ValDev muR;     // inv * R mod n     (decrypt: m = L(u) * inv mod n)
  // CRT part (half width H = W/2): p^2, q^2 and n at H limbs; p and q at
  // H/2 limbs (p^2 and q^2 fit in H limbs), R_h = 2^(16 H)
  ModDev p2, q2, nh, ph, qh;
  ValDev hpR;     // h_p * R_h mod p
  ValDev hqR;     // h_q * R_h mod q
  ValDev qinvR;   // q^-1 * R_h mod p
  ValDev qRn;     // q * R_H mod n     (q*h = MontMul_n(h, qRn))
};

// Constants of the textbook CRT route of a lookup key's decryption (half
// width H = W/2; p and q at H/2 limbs, R_h = 2^(16 H)): the layout of the
// PaillierDev CRT part, with h_p = q^-1 g_a_inv mod p and h_q = p^-1 g_a_inv
// mod q, and R_h^3 mod p, mod q for the test x_p = 1 mod p.
struct LookupCrtDev {
  ModDev nh, ph, qh;
  ValDev hpR;     // q^-1 g_a_inv R_h mod p
  ValDev hqR;     // p^-1 g_a_inv R_h mod q
  ValDev qinvR;   // q^-1 R_h mod p
  ValDev qRn;     // q R_H mod n
  ValDev pR3;     // R_h^3 mod p
  ValDev qR3;     // R_h^3 mod q
};

template <int L>
__device__ __forceinline__ size_t limb_index(int limb, int item, int stride) {
  return (size_t)limb * (size_t)stride + (size_t)item;
}

// load / store a lane's K limbs of item `item` from an interleaved buffer
template <int L, int TPI>
__device__ __forceinline__ void load_dist(u32 (&x)[L / TPI], const u32* __restrict__ src, int item, int stride) {
  constexpr int K = L / TPI;
  const int r = Grp<TPI>::rank();
#pragma unroll
  for (int k = 0; k < K; ++k) x[k] = src[limb_index<L>(r * K + k, item, stride)];
}

template <int L, int TPI>
__device__ __forceinline__ void store_dist(u32* __restrict__ dst, const u32 (&x)[L / TPI], int item, int stride) {
  constexpr int K = L / TPI;
  const int r = Grp<TPI>::rank();
#pragma unroll
  for (int k = 0; k < K; ++k) dst[limb_index<L>(r * K + k, item, stride)] = x[k];
}

// load a lane's slice of a constant bignum (same for all groups)
template <int L, int TPI>
__device__ __forceinline__ void load_const(u32 (&x)[L / TPI], const u32* __restrict__ src) {
  constexpr int K = L / TPI;
  const int r = Grp<TPI>::rank();
#pragma unroll
  for (int k = 0; k < K; ++k) x[k] = src[r * K + k];
}

template <int L, int TPI>
__device__ __forceinline__ void set_zero(u32 (&x)[L / TPI]) {
#pragma unroll
  for (int k = 0; k < L / TPI; ++k) x[k] = 0;
}

template <int L, int TPI>
__device__ __forceinline__ void set_one(u32 (&x)[L / TPI]) {
  set_zero<L, TPI>(x);
  x[0] = (Grp<TPI>::rank() == 0) ? 1u : 0u;
}

// per-wave LDS staging area
template <int L, int TPI>
__device__ __forceinline__ uint4* wave_lds(uint4* smem) {
  constexpr int slots = (L / 4) * (64 / TPI);
  return smem + (threadIdx.x / 64) * slots;
}

template <int L, int TPI>
__device__ __forceinline__ void mont_init(Mont<L, TPI>& M, const ModDev* __restrict__ md, uint4* smem) {
  M.lds = wave_lds<L, TPI>(smem);
  M.n0inv = md->n0inv;
  load_const<L, TPI>(M.N, md->N);
}

// item handled by the caller's group; groups past `count` work on a clamped
// item (so every lane of a wave takes part in the cross-lane steps) and skip
// their stores.
template <int TPI>
__device__ __forceinline__ int group_item() {
  return (int)(blockIdx.x * (BLOCK_THREADS / TPI) + threadIdx.x / TPI);
}

// ---------------------------------------------------------------------------
// Sliding-window exponentiation with a host-built schedule shared by all items
// ---------------------------------------------------------------------------
// sched[0] = index of the first window; sched[s] for s >= 1 packs
// (squarings << 16) | (odd-power index + 1), index + 1 == 0 meaning "square
// only". The table of odd powers x^1, x^3, ..., x^(2^w - 1) is kept per item
// in global scratch (interleaved like every other batch buffer).
template <int L, int TPI>
__device__ void modexp_sched(Mont<L, TPI>& M, u32 (&x)[L / TPI], const u32* __restrict__ sched, int nsched,
                             int tentries, u32* __restrict__ table, int item, int tstride) {
  constexpr int K = L / TPI;
  // table build: T0 = x, T_j = T_{j-1} * x^2
  u32 x2[K];
#pragma unroll
  for (int k = 0; k < K; ++k) x2[k] = x[k];
  M.sqr(x2);
  store_dist<L, TPI>(table, x, item, tstride);
  u32 cur[K];
#pragma unroll
  for (int k = 0; k < K; ++k) cur[k] = x[k];
  for (int j = 1; j < tentries; ++j) {
    M.mul_into(cur, x2);
    store_dist<L, TPI>(table + (size_t)j * L * tstride, cur, item, tstride);
  }
  // first window
  load_dist<L, TPI>(x, table + (size_t)sched[0] * L * tstride, item, tstride);
  for (int s = 1; s < nsched; ++s) {
    const u32 op = sched[s];
    const int sq = (int)(op >> 16);
    const int idx = (int)(op & 0xFFFFu);
    for (int i = 0; i < sq; ++i) M.sqr(x);
    if (idx) {
      u32 y[K];
      load_dist<L, TPI>(y, table + (size_t)(idx - 1) * L * tstride, item, tstride);
      M.mul_into(x, y);
    }
  }
}

// to / from Montgomery form
template <int L, int TPI>
__device__ __forceinline__ void to_mont(Mont<L, TPI>& M, u32 (&x)[L / TPI], const ModDev* __restrict__ md) {
  u32 r2[L / TPI];
  load_const<L, TPI>(r2, md->R2);
  M.mul_into(x, r2);
}

template <int L, int TPI>
__device__ __forceinline__ void from_mont(Mont<L, TPI>& M, u32 (&x)[L / TPI]) {
  u32 one[L / TPI];
  set_one<L, TPI>(one);
  M.mul_into(x, one);
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// out = a * b mod N (pointwise); a, b may be any values below R = 2^(32L)
template <int L, int TPI>
__global__ __launch_bounds__(BLOCK_THREADS) void k_mulmod(const ModDev* __restrict__ md, const u32* __restrict__ A,
                                                          const u32* __restrict__ B, u32* __restrict__ out,
                                                          int count, int stride) {
  extern __shared__ uint4 smem[];
  constexpr int K = L / TPI;
  Mont<L, TPI> M;
  mont_init<L, TPI>(M, md, smem);
  const int item = group_item<TPI>();
  const int it = item < count ? item : count - 1;
  u32 a[K], b[K], t[K], r2[K];
  load_dist<L, TPI>(a, A, it, stride);
  load_dist<L, TPI>(b, B, it, stride);
  load_const<L, TPI>(r2, md->R2);
  M.mul(t, b, r2);  // b*R mod N  (any b < R)
  M.mul(b, a, t);   // a*b mod N  (any a < R)
  if (item < count) store_dist<L, TPI>(out, b, item, stride);
}

// out_i = x_i * y mod N with one operand shared by the batch; yR = y*R mod N
template <int L, int TPI>
__global__ __launch_bounds__(BLOCK_THREADS) void k_mulmod_bcast(const ModDev* __restrict__ md,
                                                                const u32* __restrict__ X,
                                                                const u32* __restrict__ yR, u32* __restrict__ out,
                                                                int count, int stride) {
  extern __shared__ uint4 smem[];
  constexpr int K = L / TPI;
  Mont<L, TPI> M;
  mont_init<L, TPI>(M, md, smem);
  const int item = group_item<TPI>();
  const int it = item < count ? item : count - 1;
  u32 x[K], y[K], t[K];
  load_dist<L, TPI>(x, X, it, stride);
  load_const<L, TPI>(y, yR);
  M.mul(t, y, x);
  if (item < count) store_dist<L, TPI>(out, t, item, stride);
}

// m = ((u - 1) // n) * c mod n with Python floor-division semantics, as the
// reference computes it (scheme.py:101-104, lookup.py:275-278):
//   u >= 1:  w = u - 1, q = (w - (w mod n)) / n  (exact division by n)
//   u == 0:  (-1) // n = -1, so q = n - 1 (= -1 mod n)
// then MontMul_n(q, cR) = q * c mod n for cR = c*R mod n. For a valid
// ciphertext u = 1 + q*n and w mod n = 0.
template <int L, int TPI>
__device__ __forceinline__ void paillier_finish(Mont<L, TPI>& Mn, const ModDev* __restrict__ mdn, u32 (&u)[L / TPI],
                                                const u32* __restrict__ cR) {
  constexpr int K = L / TPI;
  u32 nz = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) nz |= u[k];
  const bool is_zero = Grp<TPI>::group_bits(__ballot(nz != 0)) == 0ull;
  (void)sub_small<K, TPI>(u, 1u);
  // w mod n = MontMul(MontMul(w, R^2), 1)
  u32 r0[K], r2[K];
  load_const<L, TPI>(r2, mdn->R2);
  Mn.mul(r0, u, r2);
  from_mont<L, TPI>(Mn, r0);
  u32 w[K];
  (void)sub_dist<K, TPI>(w, u, r0);
  u32 q[K];
  Mn.exact_div(q, w);
  u32 nm1[K];
#pragma unroll
  for (int k = 0; k < K; ++k) nm1[k] = Mn.N[k];
  (void)sub_small<K, TPI>(nm1, 1u);
#pragma unroll
  for (int k = 0; k < K; ++k) q[k] = is_zero ? nm1[k] : q[k];
  u32 c[K];
  load_const<L, TPI>(c, cR);
  Mn.mul(u, q, c);
}

template <int L, int TPI>
__device__ __forceinline__ u32 group_popcount(const u32 (&m)[L / TPI], const u32* __restrict__ mask) {
  constexpr int K = L / TPI;
  u32 mk[K];
  load_const<L, TPI>(mk, mask);
  u32 c = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) c += __popc(m[k] & mk[k]);
#pragma unroll
  for (int off = 1; off < TPI; off <<= 1) c += (u32)__shfl_xor((int)c, off, 64);
  return c;
}

// ---------------------------------------------------------------------------
// Pre/post-processing around the lazy exponentiation (canonical engine)
// ---------------------------------------------------------------------------

// m = L(u) * cR mod n for u = c^e mod n^2 (the exponentiation is done by
// k_lazy_modexp): plain decrypt (e = phi, cR = inv*R), lookup decrypt (e = a,
// cR = g_a_inv*R). popcnt != nullptr: also the odd-bit Hamming count of m.
template <int L, int TPI>
__global__ __launch_bounds__(BLOCK_THREADS) void k_decrypt_finish(const ModDev* __restrict__ nm,
                                                                  const u32* __restrict__ cR,
                                                                  const u32* __restrict__ U, u32* __restrict__ out,
                                                                  int count, int stride, const u32* __restrict__ mask,
                                                                  u32* __restrict__ popcnt) {
  extern __shared__ uint4 smem[];
  constexpr int K = L / TPI;
  const int item = group_item<TPI>();
  const int it = item < count ? item : count - 1;
  u32 x[K];
  load_dist<L, TPI>(x, U, it, stride);
  Mont<L, TPI> Mn;
  mont_init<L, TPI>(Mn, nm, smem);
  paillier_finish<L, TPI>(Mn, nm, x, cR);
  if (out != nullptr && item < count) store_dist<L, TPI>(out, x, item, stride);
  if (popcnt != nullptr) {
    const u32 c = group_popcount<L, TPI>(x, mask);
    if (item < count && Grp<TPI>::rank() == 0) popcnt[item] = c;
  }
}

// x = x_p e_p + x_q e_q mod n^2 (canonical) from the half-width residues in
// U (x_p in rows [0, H), x_q in rows [H, 2H), H = L/2) and eR = [e_p R |
// e_q R] mod n^2, e_p = 1 mod p^2 and 0 mod q^2, e_q the converse: the
// unique x below n^2 with those residues (CRT).
template <int L, int TPI>
__device__ __forceinline__ void crt_recombine(Mont<L, TPI>& M2, const u32* __restrict__ eR,
                                              const u32* __restrict__ U, int it, int stride, u32 (&x)[L / TPI]) {
  constexpr int K = L / TPI, H = L / 2;
  const int r0 = Grp<TPI>::rank() * K;
  u32 xp[K], xq[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {  // half-width residues zero-extended to L limbs
    const int l = r0 + k;
    xp[k] = l < H ? U[limb_index<L>(l, it, stride)] : 0u;
    xq[k] = l < H ? U[limb_index<L>(H + l, it, stride)] : 0u;
  }
  u32 e[K], y[K];
  load_const<L, TPI>(e, eR);
  M2.mul(x, xp, e);  // x_p e_p mod n^2 (x_p < R: the product is canonical)
  load_const<L, TPI>(e, eR + L);
  M2.mul(y, xq, e);  // x_q e_q mod n^2
  add_mod<K, TPI>(x, y, M2.N);
}

// Lookup decryption from half-width residues (the CRT form of
// paillier/lookup.py:257-278). U holds x_p = c^a mod p^2 in rows [0, H) and
// x_q = c^a mod q^2 in rows [H, 2H), H = L/2. With the idempotents
// e_p (1 mod p^2, 0 mod q^2) and e_q (0 mod p^2, 1 mod q^2),
// x = x_p e_p + x_q e_q mod n^2 is c^a mod n^2 for every c in [0, n^2), unit
// or not; eR = [e_p R | e_q R] mod n^2 (L limbs each). Then as
// k_decrypt_finish: m = L(x) * g_a_inv mod n, optionally the odd-bit count.
// list != nullptr: only the list[0] items list[1..] (k_lookup_crt_fast's
// flagged ones), one per group; blocks past the list exit at once.
template <int L, int TPI>
__global__ __launch_bounds__(BLOCK_THREADS) void k_lookup_crt_finish(const ModDev* __restrict__ n2m,
                                                                     const u32* __restrict__ eR,
                                                                     const ModDev* __restrict__ nm,
                                                                     const u32* __restrict__ cR,
                                                                     const u32* __restrict__ U, u32* __restrict__ out,
                                                                     int count, int stride,
                                                                     const u32* __restrict__ mask,
                                                                     u32* __restrict__ popcnt,
                                                                     const u32* __restrict__ list = nullptr) {
  extern __shared__ uint4 smem[];
  constexpr int K = L / TPI;
  int item = group_item<TPI>();
  int it;
  if (list != nullptr) {
    const int nl = (int)list[0];
    if ((int)(blockIdx.x * (BLOCK_THREADS / TPI)) >= nl) return;  // the whole block is past the list
    it = (int)list[1 + (item < nl ? item : nl - 1)];
    item = item < nl ? it : count;  // groups past the list compute a duplicate and skip their stores
  } else {
    it = item < count ? item : count - 1;
  }
  Mont<L, TPI> M2;
  mont_init<L, TPI>(M2, n2m, smem);
  u32 x[K];
  crt_recombine<L, TPI>(M2, eR, U, it, stride, x);
  Mont<L, TPI> Mn;
  mont_init<L, TPI>(Mn, nm, smem);
  paillier_finish<L, TPI>(Mn, nm, x, cR);
  if (out != nullptr && item < count) store_dist<L, TPI>(out, x, item, stride);
  if (popcnt != nullptr) {
    const u32 c = group_popcount<L, TPI>(x, mask);
    if (item < count && Grp<TPI>::rank() == 0) popcnt[item] = c;
  }
}

// Lookup decryption by the textbook CRT route (the form of
// paillier/lookup.py:257-278 for valid ciphertexts). U as k_lookup_crt_finish.
// When x = c^a mod n^2 is 1 mod n (every ciphertext the scheme makes: g^a and
// each noise^a are 1 mod n), x_p = 1 + p (qL mod p) with L = L(x), so L_p =
// (x_p - 1)/p is an exact quotient and m = L g_a_inv mod n has m_p = L_p h_p
// mod p (h_p = q^-1 g_a_inv), m_q likewise; Garner then gives m below n, all
// at half or quarter width (each group's halves of TPI/2 lanes work on p and
// q, as k_crt_finish). Test: x_p R = lo R + hi R^2 = R (mod p) with x_p = lo +
// hi R, two quarter-width products per half. An item that fails it for p or
// q (c invalid or a non-unit, where L(x) depends on all of x) appends its
// index to list[1..] and counts itself in list[0] (zeroed by the caller); its
// outputs here are placeholders, which k_lookup_crt_finish over the list
// replaces with the exact result.
template <int H, int TPI>
__global__ __launch_bounds__(BLOCK_THREADS) void k_lookup_crt_fast(const LookupCrtDev* __restrict__ key,
                                                                   const u32* __restrict__ U, u32* __restrict__ out,
                                                                   int count, int stride, u32* __restrict__ list,
                                                                   const u32* __restrict__ mask,
                                                                   u32* __restrict__ popcnt) {
  extern __shared__ uint4 smem[];
  constexpr int HH = H / 2, TH = TPI / 2, K = H / TPI;
  static_assert(TPI >= 2 && HH / TH == K, "two half-width halves per group");
  const int item = group_item<TPI>();
  const int it = item < count ? item : count - 1;
  const int half = Grp<TPI>::rank() / TH;  // 0: p, 1: q
  const ModDev* md = half == 0 ? &key->ph : &key->qh;
  Mont<HH, TH> M;
  mont_init<HH, TH>(M, md, smem);
  const u32* Uh = U + (size_t)half * H * stride;
  u32 u[K], hi[K];
  load_dist<HH, TH>(u, Uh, it, stride);
  load_dist<HH, TH>(hi, Uh + (size_t)HH * stride, it, stride);
  // x_p R mod p = MontMul(lo, R^2) + MontMul(hi, R^3), against R mod p
  u32 t1[K], t2[K], c[K];
  load_const<HH, TH>(c, md->R2);
  M.mul(t1, u, c);
  load_const<HH, TH>(c, half == 0 ? key->pR3.v : key->qR3.v);
  M.mul(t2, hi, c);
  add_mod<K, TH>(t1, t2, M.N);
  load_const<HH, TH>(c, md->R1);
  u32 diff = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) diff |= t1[k] ^ c[k];
  // the group's both halves must pass
  const bool ok = Grp<TPI>::group_bits(__ballot(diff != 0)) == 0ull;
  if (!ok && item < count && Grp<TPI>::rank() == 0) list[1 + atomicAdd(list, 1u)] = (u32)item;
  (void)sub_small<K, TH>(u, 1u);
  u32 lq[K];
  M.exact_div(lq, u);
  u32 m[K];
  load_const<HH, TH>(c, half == 0 ? key->hpR.v : key->hqR.v);
  M.mul(m, lq, c);  // m_p (low half) or m_q (high half), canonical
  // Garner, computed by both halves modulo p
  u32 mp[K], mq[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const u32 o = (u32)__shfl_xor((int)m[k], TH);
    mp[k] = half == 0 ? m[k] : o;
    mq[k] = half == 0 ? o : m[k];
  }
  Mont<HH, TH> P;
  mont_init<HH, TH>(P, &key->ph, smem);
  u32 qi[K], h1[K], h2[K];
  load_const<HH, TH>(qi, key->qinvR.v);
  P.mul(h1, mp, qi);  // m_p * q^-1 mod p
  P.mul(h2, mq, qi);  // m_q * q^-1 mod p   (m_q < R_h, so the product is reduced)
  sub_mod<K, TH>(h1, h2, P.N);
  // to the full-width distribution: the high half's lanes hold limbs >= H/2, zero
#pragma unroll
  for (int k = 0; k < K; ++k) {
    h1[k] = half == 0 ? h1[k] : 0u;
    mq[k] = half == 0 ? mq[k] : 0u;
  }
  Mont<H, TPI> Mn;
  mont_init<H, TPI>(Mn, &key->nh, smem);
  u32 qr[K], res[K];
  load_const<H, TPI>(qr, key->qRn.v);
  Mn.mul(res, h1, qr);  // q*h  (< n, exact)
  (void)add_dist<K, TPI>(res, mq);
  if (out != nullptr && item < count) {
    // the plaintext at W = 2H limbs, the upper H zero
    store_dist<H, TPI>(out, res, item, stride);
    const u32 z[K] = {};
    store_dist<H, TPI>(out + (size_t)H * stride, z, item, stride);
  }
  if (popcnt != nullptr) {
    const u32 pc = group_popcount<H, TPI>(res, mask);
    if (item < count && Grp<TPI>::rank() == 0) popcnt[item] = pc;
  }
}

// c = (1 + m*n) * y mod n^2 for y = r^n mod n^2 (cu:522-546 form of scheme.py:63-83)
template <int L, int TPI>
__global__ __launch_bounds__(BLOCK_THREADS) void k_encrypt_finish(const PaillierDev* __restrict__ key,
                                                                  const u32* __restrict__ Mv,
                                                                  const u32* __restrict__ Y, u32* __restrict__ out,
                                                                  int count, int stride) {
  extern __shared__ uint4 smem[];
  constexpr int K = L / TPI;
  Mont<L, TPI> M;
  mont_init<L, TPI>(M, &key->n2, smem);
  const int item = group_item<TPI>();
  const int it = item < count ? item : count - 1;
  u32 a[K], y[K], c[K];
  load_dist<L, TPI>(a, Mv, it, stride);
  load_dist<L, TPI>(y, Y, it, stride);
  load_const<L, TPI>(c, key->nR.v);
  M.mul_into(a, c);                  // m*n (< n^2, exact)
  (void)add_small<K, TPI>(a, 1u);    // 1 + m*n
  load_const<L, TPI>(c, key->n2.R2);
  M.mul_into(y, c);                  // y*R
  M.mul_into(a, y);                  // (1 + m n) * y
  if (item < count) store_dist<L, TPI>(out, a, item, stride);
}

// k_encrypt_finish from half-width residues of y = r^n (U as crt_recombine):
// c = (1 + m*n) * y mod n^2 with y recombined modulo n^2 first.
template <int L, int TPI>
__global__ __launch_bounds__(BLOCK_THREADS) void k_encrypt_crt_finish(const PaillierDev* __restrict__ key,
                                                                      const u32* __restrict__ eR,
                                                                      const u32* __restrict__ Mv,
                                                                      const u32* __restrict__ U, u32* __restrict__ out,
                                                                      int count, int stride) {
  extern __shared__ uint4 smem[];
  constexpr int K = L / TPI;
  Mont<L, TPI> M;
  mont_init<L, TPI>(M, &key->n2, smem);
  const int item = group_item<TPI>();
  const int it = item < count ? item : count - 1;
  u32 a[K], y[K], c[K];
  crt_recombine<L, TPI>(M, eR, U, it, stride, y);
  load_dist<L, TPI>(a, Mv, it, stride);
  load_const<L, TPI>(c, key->nR.v);
  M.mul_into(a, c);                  // m*n (< n^2, exact)
  (void)add_small<K, TPI>(a, 1u);    // 1 + m*n
  load_const<L, TPI>(c, key->n2.R2);
  M.mul_into(y, c);                  // y*R
  M.mul_into(a, y);                  // (1 + m n) * y
  if (item < count) store_dist<L, TPI>(out, a, item, stride);
}

// CRT finish. Each group of TPI lanes handles one item; its two halves of
// TPI/2 lanes work on the p- and the q-residue at once, at half width (H/2
// limbs, K limbs per lane as at full width, so the low half's lanes hold
// exactly the low H/2 limbs of the full-width distribution).
// u_p = c^(p-1) mod p^2 (U: [2][H][stride], u_p then u_q) is 0 when p | c,
// else 1 mod p (Fermat), so L_p(u_p) = (u_p - 1) / p is an exact quotient
// below p, the Hensel division of the low H/2 limbs of u_p - 1; u_p = 0
// gives floor(-1 / p) = -1, i.e. p - 1 (the floor semantics of scheme.py:86-104).
// m_p = L_p(u_p) h_p mod p, m_q likewise; then Garner
// m = m_q + q ((m_p - m_q) q^-1 mod p), q h < n exact at H limbs.
// bad != nullptr: each item with p | c or q | c (a non-unit: u_p or u_q is 0)
// appends its index to bad[1..] and counts itself in bad[0] (zeroed by the
// caller); such items are where this formula and the full-width one may
// differ. out may be nullptr; popcnt != nullptr: the
// odd-bit count of m under the low H limbs of mask.
template <int H, int TPI>
__global__ __launch_bounds__(BLOCK_THREADS) void k_crt_finish(const PaillierDev* __restrict__ key,
                                                              const u32* __restrict__ U, u32* __restrict__ out,
                                                              int count, int stride, u32* __restrict__ bad,
                                                              const u32* __restrict__ mask,
                                                              u32* __restrict__ popcnt) {
  extern __shared__ uint4 smem[];
  constexpr int HH = H / 2, TH = TPI / 2, K = H / TPI;
  static_assert(TPI >= 2 && HH / TH == K, "two half-width halves per group");
  const int item = group_item<TPI>();
  const int it = item < count ? item : count - 1;
  const int half = Grp<TPI>::rank() / TH;  // 0: p, 1: q
  Mont<HH, TH> M;
  mont_init<HH, TH>(M, half == 0 ? &key->ph : &key->qh, smem);
  const u32* Uh = U + (size_t)half * H * stride;
  u32 u[K], hi[K];
  load_dist<HH, TH>(u, Uh, it, stride);
  load_dist<HH, TH>(hi, Uh + (size_t)HH * stride, it, stride);
  u32 nz = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) nz |= u[k] | hi[k];
  const bool is_zero = Grp<TH>::group_bits(__ballot(nz != 0)) == 0ull;
  if (bad != nullptr) {  // non-units (p | c or q | c): append the item once to the list bad[1 + i]
    const bool zq = __shfl_xor(is_zero ? 1 : 0, TH) != 0;  // the other half's test
    if ((is_zero || zq) && item < count && Grp<TPI>::rank() == 0) bad[1 + atomicAdd(bad, 1u)] = (u32)item;
  }
  (void)sub_small<K, TH>(u, 1u);
  u32 lq[K];
  M.exact_div(lq, u);
  u32 nm1[K];
#pragma unroll
  for (int k = 0; k < K; ++k) nm1[k] = M.N[k];
  (void)sub_small<K, TH>(nm1, 1u);
#pragma unroll
  for (int k = 0; k < K; ++k) lq[k] = is_zero ? nm1[k] : lq[k];
  u32 c[K], m[K];
  load_const<HH, TH>(c, half == 0 ? key->hpR.v : key->hqR.v);
  M.mul(m, lq, c);  // m_p (low half) or m_q (high half), canonical
  // Garner, computed by both halves modulo p
  u32 mp[K], mq[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const u32 o = (u32)__shfl_xor((int)m[k], TH);
    mp[k] = half == 0 ? m[k] : o;
    mq[k] = half == 0 ? o : m[k];
  }
  Mont<HH, TH> P;
  mont_init<HH, TH>(P, &key->ph, smem);
  u32 qi[K], h1[K], h2[K];
  load_const<HH, TH>(qi, key->qinvR.v);
  P.mul(h1, mp, qi);  // m_p * q^-1 mod p
  P.mul(h2, mq, qi);  // m_q * q^-1 mod p   (m_q < R_h, so the product is reduced)
  sub_mod<K, TH>(h1, h2, P.N);
  // to the full-width distribution: the high half's lanes hold limbs >= H/2, zero
#pragma unroll
  for (int k = 0; k < K; ++k) {
    h1[k] = half == 0 ? h1[k] : 0u;
    mq[k] = half == 0 ? mq[k] : 0u;
  }
  Mont<H, TPI> Mn;
  mont_init<H, TPI>(Mn, &key->nh, smem);
  u32 qr[K], res[K];
  load_const<H, TPI>(qr, key->qRn.v);
  Mn.mul(res, h1, qr);  // q*h  (< n, exact)
  (void)add_dist<K, TPI>(res, mq);
  if (out != nullptr && item < count) store_dist<H, TPI>(out, res, item, stride);
  if (popcnt != nullptr) {
    const u32 c = group_popcount<H, TPI>(res, mask);
    if (item < count && Grp<TPI>::rank() == 0) popcnt[item] = c;
  }
}

// ---------------------------------------------------------------------------
// Layout conversion: AoS [count][L] <-> interleaved [L][stride]
// ---------------------------------------------------------------------------
__global__ void k_aos_to_il(const u32* __restrict__ aos, u32* __restrict__ il, int count, int L, int stride) {
  __shared__ u32 tile[32][33];
  const int i0 = blockIdx.x * 32, l0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: ty 0..7
  for (int y = ty; y < 32; y += 8) {
    const int i = i0 + y, l = l0 + tx;
    tile[y][tx] = (i < count && l < L) ? aos[(size_t)i * L + l] : 0u;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int l = l0 + y, i = i0 + tx;
    if (l < L && i < stride) il[(size_t)l * stride + i] = tile[tx][y];
  }
}

__global__ void k_il_to_aos(const u32* __restrict__ il, u32* __restrict__ aos, int count, int L, int stride) {
  __shared__ u32 tile[32][33];
  const int i0 = blockIdx.x * 32, l0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int y = ty; y < 32; y += 8) {
    const int l = l0 + y, i = i0 + tx;
    tile[y][tx] = (l < L && i < count) ? il[(size_t)l * stride + i] : 0u;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int i = i0 + y, l = l0 + tx;
    if (i < count && l < L) aos[(size_t)i * L + l] = tile[tx][y];
  }
}

// ---------------------------------------------------------------------------
// Calibration copy for the HBM byte counters: copies an interleaved batch
// with the access pattern of the exponentiation and mulmod kernels (TPI lanes
// per item, K32 consecutive limb rows per lane, so one load instruction touches
// 64/TPI consecutive items of TPI rows). Its byte count is known exactly, so
// FETCH_SIZE / WRITE_SIZE of this kernel give the counters' scale for requests
// of that shape.
// ---------------------------------------------------------------------------
template <int TPI, int L32>
__global__ __launch_bounds__(BLOCK_THREADS) void k_copy_groups(const u32* __restrict__ src, u32* __restrict__ dst,
                                                               int count, int stride) {
  constexpr int K32 = L32 / TPI;
  const int item = group_item<TPI>();
  if (item >= count) return;
  const int r = Grp<TPI>::rank();
#pragma unroll
  for (int k = 0; k < K32; ++k) dst[(size_t)(r * K32 + k) * stride + item] = src[(size_t)(r * K32 + k) * stride + item];
}

// ---------------------------------------------------------------------------
// Instruction-rate probes: `iters` rounds of 16 independent operations per
// lane (8 chains x 2); the host times them to pin issue rates on the device.
//   kind 0: v_mad_u64_u32   (u64)a*b + acc
//   kind 1: v_mul_lo_u32    a*b (low 32 bits)
//   kind 2: v_mul_hi_u32    high 32 bits of a*b
//   kind 3: 64-bit add      acc + v
//   kind 4: v_fma_f64
//   kind 5: v_mad_u32_u24   24x24 + 32
//   kind 6: v_mad_u64_u32 with an SGPR-pair carry-out (the compiler's form; kind 0 writes VCC)
//   kind 7: kind 6 plus one v_add_u32 per two MADs (ops counted = MADs)
// ---------------------------------------------------------------------------
// The operations are written as inline asm so the compiler can neither hoist
// nor fold them; each asm statement is one ALU instruction on registers (no
// memory access), 8 independent chains per lane.
template <int KIND>
__global__ __launch_bounds__(256) void k_probe(u64* __restrict__ sink, int iters, u32 seed) {
  u32 b = seed * 2654435761u + blockIdx.x;
  u64 acc[8];
  u32 w[8];
  double d[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    acc[c] = ((u64)(threadIdx.x + c) << 33) | (seed + c);
    w[c] = seed ^ (threadIdx.x * 7 + c);
    d[c] = (double)(threadIdx.x + c);
  }
  const u64 b64 = ((u64)b << 17) | 1u;
  const double x = 1.0000001, y = 0.25;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int rep = 0; rep < 2; ++rep) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        if constexpr (KIND == 0)
          asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(acc[c]) : "v"(w[c]), "v"(b) : "vcc");
        if constexpr (KIND == 1) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(w[c]) : "v"(b));
        if constexpr (KIND == 2) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(w[c]) : "v"(b));
        if constexpr (KIND == 3) asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(acc[c]) : "v"(b64));
        if constexpr (KIND == 4) asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(d[c]) : "v"(x), "v"(y));
        if constexpr (KIND == 5) asm volatile("v_mad_u32_u24 %0, %0, %1, %0" : "+v"(w[c]) : "v"(b));
        if constexpr (KIND == 6) {  // v_mad_u64_u32 with the carry-out in an SGPR pair (as compiled code does)
          u64 cy;
          asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(acc[c]), "=s"(cy) : "v"(w[c]), "v"(b));
        }
        if constexpr (KIND == 7) {  // 2 such MADs + 1 independent 32-bit add: issue-rate mix
          u64 cy;
          asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(acc[c]), "=s"(cy) : "v"(w[c]), "v"(b));
          if (c & 1) asm volatile("v_add_u32 %0, %0, %1" : "+v"(w[c]) : "v"(b));
        }
      }
    }
  }
  u64 s = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) s ^= acc[c] ^ (u64)w[c] ^ (u64)__double_as_longlong(d[c]);
  if (s == 0x1234567ull) sink[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace chx
