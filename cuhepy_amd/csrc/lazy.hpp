// Lazy-carry Montgomery exponentiation for gfx950 (the modexp hot loop).
//
// Numbers are held in LP positions of RB bits (RB = 28 or 29) inside 64-bit
// accumulators. A limb product is < 2^(2 RB), so every limb product is exactly
// one in-place v_mad_u64_u32 (64-bit accumulator as the addend) as long as no
// accumulator can pass 2^64; carries are resolved once per product (settle).
// The row step below splits each lane's bottom register every row, so a
// register accumulates for at most K = LP / TPI rows (see FMAX): 29-bit
// positions (72 at 2048-bit moduli, 144 at 4096-bit) need no carry folds at
// 4 or 8 lanes per modulus.
//
// Montgomery product with R' = 2^(RB*LP) > 4N ("almost Montgomery", Walter
// 2002): inputs and outputs stay below 2N, so no conditional subtraction is
// needed inside the exponentiation; the result is made canonical once at the
// end.
//
// A group of TPI lanes holds K = LP / TPI positions per lane. Each row i adds
// a_i * b + m_i * N and shifts down one position. The K rows of a block are
// fully unrolled and the accumulator registers are renamed instead of moved
// (position p of row s lives in register (p + s) mod K), so the shift costs no
// instructions inside a lane; across lanes each lane's bottom position passes
// its low RB bits one lane down (one masked DPP row_shl) and carries the rest
// into its own next position. a_i comes from LDS (one group-uniform read per
// row), m_i from rank 0 by a masked DPP broadcast.
#pragma once
#include "mont.hpp"

namespace chx {

__device__ __forceinline__ u64 make64(u32 lo, u32 hi) { return ((u64)hi << 32) | lo; }

// An opaque copy of v: the compiler cannot hoist what is computed from it out
// of an enclosing loop. The lane-dependent bit offsets and addresses of
// from32/to32/stage are derived from it, so a persistent loop around a whole
// exponentiation does not keep dozens of them live in VGPRs across every tile
// (with TPI = 1 the rank is the constant 0 and stays visible to the compiler).
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <int RB, int LP, int TPI, int L32>
struct Lazy {
  static constexpr int K = LP / TPI;
  static constexpr int G = 64 / TPI;
  static constexpr u32 MASK = (1u << RB) - 1u;
  static constexpr int K32 = L32 / TPI;
  static_assert(LP % TPI == 0, "LP must be a multiple of TPI");
  static_assert(L32 % TPI == 0 && K32 % 4 == 0, "32-bit limbs per lane must be a multiple of 4");
  static_assert(RB * LP >= 32 * L32 + 2, "R' must exceed 4N");
  // Accumulator bound. A register holds one position for at most K rows: it
  // becomes a lane's top position with the low RB bits passed down from the
  // lane above (< 2^RB), takes one carry from the position below (< 2^(64-RB))
  // and 2 products per row of a general product (at most 3 products' worth in
  // a squaring row, see row_sq), and leaves as the lane's bottom position,
  // whose bits above RB go to the lane's next register and whose low bits go
  // one lane down. Products of settled positions (below 2^RB + 2^(64-2RB) + 1)
  // are below P = 2^(2RB) + 2^(66-RB). FMAX (FMAX_SQ) is the most rows that
  // keep a register below 2^64.
  static constexpr double TWO64 = 18446744073709551616.0;
  static constexpr double PMAX = (double)(1ull << (2 * RB)) + (double)(1ull << (66 - RB));
  static constexpr double ROOM = TWO64 - 2.0 * (double)(1ull << (64 - RB));
  static constexpr int FMAX = (int)(ROOM / (2.0 * PMAX));
  static constexpr int FMAX_SQ = (int)(ROOM / (3.0 * PMAX));
  static_assert(K <= FMAX, "more rows per register than a 64-bit accumulator holds");
  // Symmetric squaring rows where K of them fit; plain rows otherwise.
  static constexpr bool SYM = K <= FMAX_SQ;
  // Long product chains may run modulo N' = kN with k = -N^-1 mod 2^RB
  // instead of N: N' = -1 mod 2^RB, so -N'^-1 mod 2^RB = 1 and each row's
  // digit is the bottom position's low RB bits, with no multiply. Only the
  // residue mod N matters, and N divides N', so operands need only be
  // congruent mod N to their Montgomery forms (tables and constants built mod
  // N serve unchanged) and below 2N'; one last product modulo N itself makes
  // the result canonical. N' has up to RB more bits than N; UNIT marks the
  // shapes where R' still exceeds 16 N' (bases below 4N' give products below 2N').
  static constexpr bool UNIT = RB * LP >= 32 * L32 + RB + 4;
  static constexpr int SLOTS = (LP > L32 ? LP : L32);  // u32 rows of the staging area
  // u32 words of one wave's staging area: squarings stage (x_i, 2 x_i) pairs
  static constexpr int AREA = (SYM && 2 * LP > SLOTS ? 2 * LP : SLOTS) * G;
  typedef Grp<TPI> grp;

  u32* lds;     // this wave's staging area: AREA u32 words
  u32 n0;       // -N^-1 mod 2^RB
  u32 maskv;    // MASK held in a VGPR: a DPP move folds into an AND with it
  u32 N[K];     // modulus positions (this lane's slice)

  __device__ __forceinline__ void init(u32* wave_area, const u32* __restrict__ npos, u32 n0inv) {
    lds = wave_area;
    n0 = n0inv;
    maskv = MASK;
    asm volatile("" : "+v"(maskv));
    const int r = grp::rank();
#pragma unroll
    for (int p = 0; p < K; ++p) N[p] = npos[r * K + p];
  }

  // stage this lane's positions as the streamed operand: row i = position i
  __device__ __forceinline__ void stage(const u32 (&x)[K]) {
    const int g = TPI == 1 ? grp::gid() : opaque(grp::gid()), r = TPI == 1 ? 0 : opaque(grp::rank());
#pragma unroll
    for (int p = 0; p < K; ++p) lds[(r * K + p) * G + g] = x[p];
  }

  // stage the operand of a squaring: row i = the pair (x_i, 2 x_i) (u64)
  __device__ __forceinline__ void stage_sq(const u32 (&x)[K]) {
    const int g = TPI == 1 ? grp::gid() : opaque(grp::gid()), r = TPI == 1 ? 0 : opaque(grp::rank());
    u64* l64 = reinterpret_cast<u64*>(lds);
#pragma unroll
    for (int p = 0; p < K; ++p) l64[(r * K + p) * G + g] = make64(x[p], x[p] << 1);
  }

  // Settle accumulators into positions for the next product (value < R'
  // guaranteed, so nothing leaves the group's top position). Each lane ripples
  // its own positions, then passes its carry (< 2^(64-RB)) one lane up, where
  // it is absorbed by positions 0 and 1 only: every position ends below
  // D = 2^RB + 2^(64-2RB) + 1 instead of 2^RB. Products of such positions stay
  // below 2^(2RB) + 2^(66-RB) (PMAX, which the row bounds above use).
  // renorm() finishes the job where exact RB-bit positions are needed (to32,
  // top32).
  __device__ __forceinline__ void settle(u32 (&X)[K], const u64 (&V)[K]) {
    static_assert(TPI == 1 || K >= 2, "settle needs two positions per lane");
    u64 c = 0;
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const u64 t = V[p] + c;
      X[p] = (u32)t & MASK;
      c = t >> RB;
    }
    if constexpr (TPI > 1) {
      const u64 cin = make64(grp::shift_up((u32)c), grp::shift_up((u32)(c >> 32)));
      const u64 t = (u64)X[0] + cin;
      X[0] = (u32)t & MASK;
      X[1] += (u32)(t >> RB);
    }
  }

  // Exact RB-bit positions from settled ones (value < R'), in place. Settled
  // positions are below 2^RB + 2^(64-2RB) + 1, so every carry of a lane's own
  // ripple is 0 or 1; the carries between lanes are then resolved at once with
  // a generate/propagate lookahead over the group (a lane propagates when all
  // its positions are MASK), and one more ripple adds each lane's carry-in.
  __device__ __forceinline__ void renorm(u32 (&X)[K]) {
    u32 c = 0;
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const u32 t = X[p] + c;
      X[p] = t & MASK;
      c = t >> RB;
    }
    if constexpr (TPI > 1) {
      u32 all = X[0];
#pragma unroll
      for (int p = 1; p < K; ++p) all &= X[p];
      u32 cout;
      u32 cin = grp::lookahead(c != 0, all == MASK, &cout);
#pragma unroll
      for (int p = 0; p < K; ++p) {
        const u32 t = X[p] + cin;
        X[p] = t & MASK;
        cin = t >> RB;
      }
    }
  }

  // b = a*b*R'^-1 mod N (below 2N), in place; a staged in LDS. The result
  // overwrites b only after the last row has read it. U: the modulus held in
  // N is = -1 mod 2^RB (digit without multiply, see UNIT).
  template <bool U = false>
  __device__ __forceinline__ void mul_staged(u32 (&b)[K]) {
    const int g = grp::gid();
    u64 V[K];
#pragma unroll
    for (int p = 0; p < K; ++p) V[p] = 0;
#pragma unroll 1
    for (int blk = 0; blk < TPI; ++blk) {
      const u32* arow = lds + (blk * K) * G + g;
      row<0, U>(V, b, arow);
    }
    settle(b, V);
  }

  // rows S..K-1 of one block (unrolled by recursion over S)
  template <int S, bool U = false>
  __device__ __forceinline__ void row(u64 (&V)[K], const u32 (&b)[K], const u32* arow) {
    if constexpr (S < K) {
      const u32 a = arow[S * G];
#pragma unroll
      for (int p = 0; p < K; ++p) {
        const int q = (p + S) % K;
        V[q] = (u64)a * b[p] + V[q];
      }
      reduce<S, U>(V);
      row<S + 1, U>(V, b, arow);
    }
  }

  // Second half of row S: add m*N for the digit m that clears relative
  // position 0, then shift down one position.
  template <int S, bool U = false>
  __device__ __forceinline__ void reduce(u64 (&V)[K]) {
    u32 m;
    if constexpr (U)
      m = grp::bcast0_and((u32)V[S], maskv);
    else
      m = grp::bcast0_and((u32)V[S] * n0, maskv);
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const int q = (p + S) % K;
      V[q] = (u64)m * N[p] + V[q];
    }
    // Every lane carries the bits of its bottom position above RB into its
    // next position and passes only the low RB bits down to the lane below
    // (value unchanged). Rank 0's low bits are 0 after m*N, so the group's
    // top rank receives 0 from the next group without a select.
    const u64 v0 = V[S];
    V[(S + 1) % K] += v0 >> RB;
    V[S] = (u64)(grp::shift_down_z0((u32)v0) & maskv);
  }

  // Squaring rows (operand x staged in LDS and held in b). Term x_i x_j of the
  // square belongs at absolute position i + j. The row loop meets it twice: in
  // row i at relative position j, and in row j at relative position i. In row
  // S of a block, lane position p meets the terms with i mod K = S and
  // j mod K = p, and its mirror has the two residues swapped. So each row adds
  // 2 x_i x_j where p > S, nothing where p < S (the mirror, which has p > S,
  // adds the doubled term), and the single term x_i x_j where p = S (the
  // mirror also has p = S and adds the other half; for i = j it is the square
  // term). The choice depends only on (S, p), so every lane of the wave runs
  // the same code: K(K+1)/2 product MADs per block instead of K^2. Each term
  // is still added in a row <= i + j, before the reduction of its position.
  // A position receives at most 2 x_i x_j + m N_p < 3 * 2^(2 RB) per row.
  template <int S, bool U = false>
  __device__ __forceinline__ void row_sq(u64 (&V)[K], const u32 (&b)[K], const u64* arow) {
    if constexpr (S < K) {
      const u64 pr = arow[S * G];
      const u32 a = (u32)pr, a2 = (u32)(pr >> 32);
      V[(2 * S) % K] = (u64)a * b[S] + V[(2 * S) % K];
#pragma unroll
      for (int p = S + 1; p < K; ++p) {
        const int q = (p + S) % K;
        V[q] = (u64)a2 * b[p] + V[q];
      }
      reduce<S, U>(V);
      row_sq<S + 1, U>(V, b, arow);
    }
  }

  // b = b^2 R'^-1 mod N (below 2N), in place; b also staged in LDS (stage_sq).
  template <bool U = false>
  __device__ __forceinline__ void sqr_staged(u32 (&b)[K]) {
    const int g = grp::gid();
    u64 V[K];
#pragma unroll
    for (int p = 0; p < K; ++p) V[p] = 0;
#pragma unroll 1
    for (int blk = 0; blk < TPI; ++blk) {
      const u64* arow = reinterpret_cast<const u64*>(lds) + (blk * K) * G + g;
      row_sq<0, U>(V, b, arow);
    }
    settle(b, V);
  }

  // x = x * y (both below 2N)
  template <bool U = false>
  __device__ __forceinline__ void mul_into(u32 (&x)[K], const u32 (&y)[K]) {
    stage(y);
    wave_lds_sync();
    mul_staged<U>(x);
    wave_lds_sync();
  }

  template <bool U = false>
  __device__ __forceinline__ void sqr(u32 (&x)[K]) {
    if constexpr (SYM) {
      stage_sq(x);
      wave_lds_sync();
      sqr_staged<U>(x);
    } else {
      stage(x);
      wave_lds_sync();
      mul_staged<U>(x);
    }
    wave_lds_sync();
  }

  // 32-bit limbs (lane slice of K32, value < 2^(32 L32)) -> positions
  __device__ __forceinline__ void from32(u32 (&x)[K], const u32 (&w)[K32]) {
    const int g = TPI == 1 ? grp::gid() : opaque(grp::gid()), r = TPI == 1 ? 0 : opaque(grp::rank());
#pragma unroll
    for (int k = 0; k < K32; ++k) lds[(r * K32 + k) * G + g] = w[k];
    wave_lds_sync();
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const int bit = RB * (r * K + p);
      const int j = bit >> 5, off = bit & 31;
      const u32 lo = j < L32 ? lds[j * G + g] : 0u;
      const u32 hi = j + 1 < L32 ? lds[(j + 1) * G + g] : 0u;
      x[p] = (u32)(make64(lo, hi) >> off) & MASK;
    }
    wave_lds_sync();
  }

  // bit 32*L32 of a value below 2^(32 L32 + 1) in exact positions (renorm),
  // the same in every lane of the group (the carry word for cond_sub)
  __device__ __forceinline__ u32 top32(const u32 (&x)[K]) {
    constexpr int q = (32 * L32) / RB, off = (32 * L32) % RB;
    static_assert(q < LP, "bit 32*L32 lies inside the positions");
    const bool mine = grp::rank() == q / K;
    const u32 b = mine ? (x[q % K] >> off) : 0u;
    return grp::group_bits(__ballot(b != 0)) != 0ull ? 1u : 0u;
  }

  // exact positions (renorm; value < 2^(32 L32)) -> 32-bit limbs (lane slice)
  __device__ __forceinline__ void to32(u32 (&w)[K32], const u32 (&x)[K]) {
    const int g = TPI == 1 ? grp::gid() : opaque(grp::gid()), r = TPI == 1 ? 0 : opaque(grp::rank());
    stage(x);
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < K32; ++k) {
      const int bit = 32 * (r * K32 + k);
      const int q = bit / RB, off = bit % RB;
      const u32 p0 = lds[q * G + g];
      const u32 p1 = q + 1 < LP ? lds[(q + 1) * G + g] : 0u;
      const u32 p2 = q + 2 < LP ? lds[(q + 2) * G + g] : 0u;
      // 32 bits starting at `off` inside position q span at most three positions
      u64 v = (u64)p0 >> off;
      v |= (u64)p1 << (RB - off);
      if (2 * RB - off < 64) v |= (u64)p2 << (2 * RB - off);
      w[k] = (u32)v;
    }
    wave_lds_sync();
  }
};

// Constants of one modulus in the lazy radix (device layout).
constexpr int LPMAX = 160;
struct LazyDev {
  u32 n0;           // -N^-1 mod 2^RB
  u32 rb, lp, l32;  // radix bits, positions, 32-bit limbs (checked by the host)
  u32 N[LPMAX];     // modulus, positions
  u32 R2[LPMAX];    // R'^2 mod N, positions
  u32 R1[LPMAX];    // R' mod N, positions (Montgomery form of 1)
  u32 RW[LPMAX];    // 2^(32 L32) R'^2 mod N, positions (upper half of a double-width input)
  u32 NK[LPMAX];    // N' = kN = -1 mod 2^RB (k = -N^-1 mod 2^RB), positions; Lazy::UNIT shapes only
  u32 N32[LMAX];    // modulus, 32-bit limbs (final canonicalisation)
};

// x^e mod N for a batch; e shared (host-built sliding-window schedule).
// blockIdx.y selects one of two (modulus, schedule, input, output) sets so the
// CRT halves run in one launch. Input: 32-bit limbs (any value < 2^(32 L32));
// with in_hi set, the input is c_lo + 2^(32 L32) c_hi, the upper L32 limbs read
// from in_hi, so a CRT half reduces the full-width ciphertext itself. Output:
// canonical 32-bit limbs. Window table: positions, per item, in global scratch
// ([entry][item][position]: a lane's positions are contiguous).
struct LazyJob {
  const LazyDev* md;
  const u32* sched;
  int nsched;
  const u32* in;
  const u32* in_hi;  // nullptr: single-width input
  u32* out;
  u32* table;
};

template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(BLOCK_THREADS) void k_lazy_modexp(LazyJob j0, LazyJob j1, int count, int stride,
                                                               int tentries, int tstride) {
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB, LP, TPI, L32> E;
  constexpr int K = E::K, K32 = E::K32;
  // U: exponentiate modulo N' = kN (digits without multiply); the result is
  // taken modulo N by the final product
  constexpr bool U = E::UNIT;
  const LazyJob jb = blockIdx.y == 0 ? j0 : j1;
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * E::AREA;
  E M;
  if constexpr (U)
    M.init(area, jb.md->NK, 1u);
  else
    M.init(area, jb.md->N, jb.md->n0);
  const int item = group_item<TPI>();
  const int it = item < count ? item : count - 1;
  const int r = Grp<TPI>::rank();

  u32 w[K32];
#pragma unroll
  for (int k = 0; k < K32; ++k) w[k] = jb.in[(size_t)(r * K32 + k) * stride + it];
  u32 x[K];
  M.from32(x, w);
  {
    u32 r2[K];
#pragma unroll
    for (int p = 0; p < K; ++p) r2[p] = jb.md->R2[r * K + p];
    M.template mul_into<U>(x, r2);  // = x R' mod N, below 2N' (2N without U)
  }
  if (jb.in_hi != nullptr) {  // block-uniform
    // + c_hi 2^(32 L32) R' (below 2N, or 2N' with U): the sum is below 4N
    // (4N'), and the host admits this form only where R' > 16N, which keeps
    // every product of it below 2N (UNIT shapes have R' > 16N' by definition). Settling the position-wise sum leaves each position at most
    // 2^RB, and its value below R' leaves nothing above the top position.
#pragma unroll
    for (int k = 0; k < K32; ++k) w[k] = jb.in_hi[(size_t)(r * K32 + k) * stride + it];
    u32 xh[K], rw[K];
    M.from32(xh, w);
#pragma unroll
    for (int p = 0; p < K; ++p) rw[p] = jb.md->RW[r * K + p];
    M.template mul_into<U>(xh, rw);
    u64 v[K];
#pragma unroll
    for (int p = 0; p < K; ++p) v[p] = (u64)x[p] + xh[p];
    M.settle(x, v);
  }
  // odd powers x, x^3, ..., x^(2^w - 1)
  u32* tab = jb.table;
  auto tstore = [&](int e, const u32 (&v)[K]) {
#pragma unroll
    for (int p = 0; p < K; ++p) tab[((size_t)e * tstride + item) * LP + r * K + p] = v[p];
  };
  auto tload = [&](int e, u32 (&v)[K]) {
#pragma unroll
    for (int p = 0; p < K; ++p) v[p] = tab[((size_t)e * tstride + item) * LP + r * K + p];
  };
  tstore(0, x);
  if (tentries > 1) {
    // x^2 stays staged in LDS while the odd powers are built in place
    u32 cur[K];
#pragma unroll
    for (int p = 0; p < K; ++p) cur[p] = x[p];
    M.template sqr<U>(cur);
    M.stage(cur);
#pragma unroll
    for (int p = 0; p < K; ++p) cur[p] = x[p];
    wave_lds_sync();
    for (int e = 1; e < tentries; ++e) {
      M.template mul_staged<U>(cur);
      tstore(e, cur);
    }
    wave_lds_sync();
  }
  tload((int)jb.sched[0], x);
  for (int s = 1; s < jb.nsched; ++s) {
    const u32 op = jb.sched[s];
    const int sq = (int)(op >> 16);
    const int idx = (int)(op & 0xFFFFu);
    for (int i = 0; i < sq; ++i) M.template sqr<U>(x);
    if (idx) {
      u32 y[K];
      tload(idx - 1, y);
      M.template mul_into<U>(x, y);
    }
  }
  // out of Montgomery form: MontMul(x, 1) <= N, then canonical. With U the
  // product is taken modulo N itself: x < 2N' < R' gives (x + mN)/R' <= N.
  if constexpr (U) M.init(area, jb.md->N, jb.md->n0);
  {
    u32 one[K];
#pragma unroll
    for (int p = 0; p < K; ++p) one[p] = 0;
    one[0] = r == 0 ? 1u : 0u;
    M.mul_into(x, one);
  }
  M.renorm(x);
  M.to32(w, x);
  u32 n32[K32];
#pragma unroll
  for (int k = 0; k < K32; ++k) n32[k] = jb.md->N32[r * K32 + k];
  cond_sub<K32, TPI>(w, 0u, n32);
  if (item < count) {
#pragma unroll
    for (int k = 0; k < K32; ++k) jb.out[(size_t)(r * K32 + k) * stride + item] = w[k];
  }
}

// Pointwise out[i] = a[i] b[i] mod N for any a, b < 2^(32 L32), interleaved
// [L32][stride] (the Hamming evaluator's ciphertext multiply,
// encode_hamming_server: cu:672-732, lcu:1505-1561). t = MontMul(a, b) =
// a b R'^-1 (modulo N' = kN on UNIT shapes) is below a b / R' + N' < R';
// MontMul(t, R'^2 mod N), modulo N itself, is then below t N / R' + N < 2N,
// and one conditional subtraction (with the bit above the limbs, as 2N may
// pass 2^(32 L32)) makes it canonical.
template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(BLOCK_THREADS) void k_lazy_mulmod(const LazyDev* __restrict__ md,
                                                               const u32* __restrict__ A, const u32* __restrict__ B,
                                                               u32* __restrict__ out, int count, int stride) {
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB, LP, TPI, L32> E;
  constexpr int K = E::K, K32 = E::K32;
  constexpr bool U = E::UNIT;
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * E::AREA;
  E M;
  if constexpr (U)
    M.init(area, md->NK, 1u);
  else
    M.init(area, md->N, md->n0);
  const int item = group_item<TPI>();
  const int it = item < count ? item : count - 1;
  const int r = Grp<TPI>::rank();
  u32 w[K32], x[K], y[K];
#pragma unroll
  for (int k = 0; k < K32; ++k) w[k] = A[(size_t)(r * K32 + k) * stride + it];
  M.from32(x, w);
#pragma unroll
  for (int k = 0; k < K32; ++k) w[k] = B[(size_t)(r * K32 + k) * stride + it];
  M.from32(y, w);
  M.template mul_into<U>(x, y);
  if constexpr (U) M.init(area, md->N, md->n0);
#pragma unroll
  for (int p = 0; p < K; ++p) y[p] = md->R2[r * K + p];
  M.mul_into(x, y);
  M.renorm(x);
  const u32 hi = M.top32(x);
  M.to32(w, x);
  u32 n32[K32];
#pragma unroll
  for (int k = 0; k < K32; ++k) n32[k] = md->N32[r * K32 + k];
  cond_sub<K32, TPI>(w, hi, n32);
  if (item < count) {
#pragma unroll
    for (int k = 0; k < K32; ++k) out[(size_t)(r * K32 + k) * stride + item] = w[k];
  }
}

// Lookup tables to the lazy Montgomery domain: out[e] = in[e] * R' mod N
// (below 2N) as positions; in: AoS 32-bit limbs (L32 per entry), out: AoS
// positions (LP per entry). One group per entry.
template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(BLOCK_THREADS) void k_lazy_table(const LazyDev* __restrict__ md,
                                                              const u32* __restrict__ in, u32* __restrict__ out,
                                                              int entries) {
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB, LP, TPI, L32> E;
  constexpr int K = E::K, K32 = E::K32;
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * E::AREA;
  E M;
  M.init(area, md->N, md->n0);
  const int item = group_item<TPI>();
  const int it = item < entries ? item : entries - 1;
  const int r = Grp<TPI>::rank();
  u32 w[K32], x[K], r2[K];
#pragma unroll
  for (int k = 0; k < K32; ++k) w[k] = in[(size_t)it * L32 + r * K32 + k];
  M.from32(x, w);
#pragma unroll
  for (int p = 0; p < K; ++p) r2[p] = md->R2[r * K + p];
  M.mul_into(x, r2);
  if (item < entries) {
#pragma unroll
    for (int p = 0; p < K; ++p) out[(size_t)item * LP + r * K + p] = x[p];
  }
}

// Digits of the reference's bit packing ("0 b_0 0 b_1 ...", hamming/paillier.py:81-95):
// every data bit sits at an even position, so an 8-bit digit is one of the 16
// values with bits 1, 3, 5, 7 clear. spread4 maps 4 data bits to that digit,
// squeeze4 back.
__device__ __forceinline__ int spread4(int v) { return (v & 1) | ((v & 2) << 1) | ((v & 4) << 2) | ((v & 8) << 3); }
__device__ __forceinline__ int squeeze4(int d) { return (d & 1) | ((d >> 1) & 2) | ((d >> 2) & 4) | ((d >> 3) & 8); }
constexpr int PACKED_ZERO_BITS = 0xAA;

// Base-2^8 digits of interleaved plaintexts (limb j of item i at
// m[j * stride + i]; limbs at or past `limbs` read as zero) into item-major
// rows of `chunks` bytes, little-endian: the digit extraction of the fused
// re-encryption (lcu:1102-1104, lookup.py:239). One thread per item; the
// loads of a wave are coalesced per limb. *packed is set to `token` when
// some item has every digit in the bit-packing alphabet (a fresh token per
// call: the flag needs no clearing).
__global__ __launch_bounds__(BLOCK_THREADS) void k_il_digits(const u32* __restrict__ m, int limbs, int stride,
                                                             int count, unsigned char* __restrict__ digits,
                                                             int chunks, u32* __restrict__ packed, u32 token) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  constexpr u32 Z = 0x01010101u * PACKED_ZERO_BITS;
  unsigned char* row = digits + (size_t)i * chunks;
  u32 zbits = 0;
  if ((chunks & 15) == 0) {  // rows of whole 16-byte words (rows start 16-byte aligned)
    for (int q = 0; q < chunks / 16; ++q) {
      uint4 w;
      w.x = 4 * q + 0 < limbs ? m[(size_t)(4 * q + 0) * stride + i] : 0u;
      w.y = 4 * q + 1 < limbs ? m[(size_t)(4 * q + 1) * stride + i] : 0u;
      w.z = 4 * q + 2 < limbs ? m[(size_t)(4 * q + 2) * stride + i] : 0u;
      w.w = 4 * q + 3 < limbs ? m[(size_t)(4 * q + 3) * stride + i] : 0u;
      zbits |= (w.x | w.y | w.z | w.w) & Z;
      *reinterpret_cast<uint4*>(row + 16 * q) = w;
    }
  } else {
    for (int b = 0; b < chunks; ++b) {
      const int j = b >> 2;
      const u32 v = j < limbs ? m[(size_t)j * stride + i] : 0u;
      const unsigned char d = (unsigned char)(v >> (8 * (b & 3)));
      zbits |= d & PACKED_ZERO_BITS;
      row[b] = d;
    }
  }
  if (zbits == 0) *packed = token;
}

// Row-group table for lookup encryption, R = 2 or 3 rows per entry: entry e
// of group j is
//   prod_{t < R} g_table[R j + t][spread4((e >> 4t) & 15)] mod n^2
//   = g^(sum_t spread4((e >> 4t) & 15) 256^t 2^(8 R j)),
// so one product covers R 8-bit digits of a packed message (a wider digit,
// HAC 14.109 with a wider window). In and out: AoS positions in lazy
// Montgomery form (the product of such values is the form of the product).
// One group per entry; entries = groups * 16^R.
template <int RB, int LP, int TPI, int L32, int R>
__global__ __launch_bounds__(BLOCK_THREADS) void k_lazy_group_table(const LazyDev* __restrict__ md,
                                                                    const u32* __restrict__ gtab,
                                                                    u32* __restrict__ out, int entries) {
  static_assert(R >= 2 && R <= 5, "2 to 5 rows per entry");
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB, LP, TPI, L32> E;
  constexpr int K = E::K, SH = 4 * R;
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * E::AREA;
  E M;
  M.init(area, md->N, md->n0);
  const int item = group_item<TPI>();
  const int it = item < entries ? item : entries - 1;
  const int r = Grp<TPI>::rank();
  const int j = it >> SH, e = it & ((1 << SH) - 1);
  u32 x[K];
  {
    const u32* src = gtab + ((size_t)(R * j) * 256 + spread4(e & 15)) * LP + r * K;
#pragma unroll
    for (int p = 0; p < K; ++p) x[p] = src[p];
  }
#pragma unroll
  for (int t = 1; t < R; ++t) {
    const u32* src = gtab + ((size_t)(R * j + t) * 256 + spread4((e >> (4 * t)) & 15)) * LP + r * K;
    u32 y[K];
#pragma unroll
    for (int p = 0; p < K; ++p) y[p] = src[p];
    M.mul_into(x, y);
  }
  if (item < entries) {
#pragma unroll
    for (int p = 0; p < K; ++p) out[(size_t)item * LP + r * K + p] = x[p];
  }
}

// Noise-pair table: entry a * nn + b = noise[a] * noise[b] mod n^2 (nn
// entries in the noise table), so two noise factors of an encryption take one
// product. AoS positions in lazy Montgomery form, one group per entry.
template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(BLOCK_THREADS) void k_lazy_noise_pairs(const LazyDev* __restrict__ md,
                                                                    const u32* __restrict__ ntab, int nn,
                                                                    u32* __restrict__ out) {
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB, LP, TPI, L32> E;
  constexpr int K = E::K;
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * E::AREA;
  E M;
  M.init(area, md->N, md->n0);
  const int entries = nn * nn;
  const int item = group_item<TPI>();
  const int it = item < entries ? item : entries - 1;
  const int r = Grp<TPI>::rank();
  const u32* a = ntab + (size_t)(it / nn) * LP + r * K;
  const u32* b = ntab + (size_t)(it % nn) * LP + r * K;
  u32 x[K], y[K];
#pragma unroll
  for (int p = 0; p < K; ++p) x[p] = a[p], y[p] = b[p];
  M.mul_into(x, y);
  if (item < entries) {
#pragma unroll
    for (int p = 0; p < K; ++p) out[(size_t)item * LP + r * K + p] = x[p];
  }
}

// Noise-triple table: entry (a * nn + b) * nn + c = noise[a] * noise[b] *
// noise[c] mod n^2, from the noise-pair table (entry a * nn + b) and the noise
// table, one product per entry, so three noise factors of an encryption take
// one product. AoS positions in lazy Montgomery form, one group per entry.
template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(BLOCK_THREADS) void k_lazy_noise_triples(const LazyDev* __restrict__ md,
                                                                      const u32* __restrict__ nptab,
                                                                      const u32* __restrict__ ntab, int nn,
                                                                      u32* __restrict__ out) {
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB, LP, TPI, L32> E;
  constexpr int K = E::K;
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * E::AREA;
  E M;
  M.init(area, md->N, md->n0);
  const int entries = nn * nn * nn;
  const int item = group_item<TPI>();
  const int it = item < entries ? item : entries - 1;
  const int r = Grp<TPI>::rank();
  const u32* a = nptab + (size_t)(it / nn) * LP + r * K;
  const u32* b = ntab + (size_t)(it % nn) * LP + r * K;
  u32 x[K], y[K];
#pragma unroll
  for (int p = 0; p < K; ++p) x[p] = a[p], y[p] = b[p];
  M.mul_into(x, y);
  if (item < entries) {
#pragma unroll
    for (int p = 0; p < K; ++p) out[(size_t)item * LP + r * K + p] = x[p];
  }
}

// Plain residues of a table in lazy Montgomery form: out[e] = MontMul(in[e],
// 1) = the entry's value v (at most N) instead of v R'. AoS positions, one
// group per entry.
template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(BLOCK_THREADS) void k_lazy_plain(const LazyDev* __restrict__ md,
                                                              const u32* __restrict__ in, u32* __restrict__ out,
                                                              int entries) {
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB, LP, TPI, L32> E;
  constexpr int K = E::K;
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * E::AREA;
  E M;
  M.init(area, md->N, md->n0);
  const int item = group_item<TPI>();
  const int it = item < entries ? item : entries - 1;
  const int r = Grp<TPI>::rank();
  u32 x[K], one[K];
#pragma unroll
  for (int p = 0; p < K; ++p) x[p] = in[(size_t)it * LP + r * K + p], one[p] = 0;
  one[0] = r == 0 ? 1u : 0u;
  M.mul_into(x, one);
  if (item < entries) {
#pragma unroll
    for (int p = 0; p < K; ++p) out[(size_t)item * LP + r * K + p] = x[p];
  }
}

// Lookup-table encryption (lookup.py:213-252, lcu:839-948) in the lazy domain:
//   c = prod_i g_table[i][d_i] * prod_k noise[idx_k] mod n^2
// Tables are AoS positions in lazy Montgomery form. Rows where every item of
// the wave has digit 0 are skipped (g_table[i][0] = 1); other items with
// digit 0 multiply by their row's j = 0 entry (the Montgomery one), so the
// wave stays convergent and the product is unchanged. With a row-group table
// of R = prow rows per entry (ptab, k_lazy_group_table, the first pgroups
// groups), rows R j .. R j + R - 1 take one product from it when every item
// of the wave has all R digits in the packing alphabet; otherwise the wave
// multiplies the rows one at a time (so do rows past the last group of the
// table). Either way the result is the same residue. With a noise-pair table
// (nptab, k_lazy_noise_pairs) the noise factors go two per product, with a
// noise-triple table (n3tab, k_lazy_noise_triples) three, as long as more
// than three are left. The first product of a wave with an active row
// is a copy (the accumulator starts at the Montgomery one). With plain-residue
// copies of the noise tables (ntab_plain, nptab_plain: v instead of v R'),
// the last noise product also leaves the Montgomery form, so no separate
// conversion product is needed.
template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_lazy_lookup_encrypt(const LazyDev* __restrict__ md,
                                                                       const u32* __restrict__ gtab,
                                                                       const u32* __restrict__ ptab, int prow,
                                                                       int pgroups,
                                                                       const u32* __restrict__ ntab,
                                                                       const u32* __restrict__ nptab,
                                                                       const u32* __restrict__ n3tab,
                                                                       const u32* __restrict__ ntab_plain,
                                                                       const u32* __restrict__ nptab_plain,
                                                                       const unsigned char* __restrict__ digits,
                                                                       int chunks,
                                                                       const unsigned char* __restrict__ nidx,
                                                                       int nmult, int nnoise, u32* __restrict__ out,
                                                                       int count, int stride) {
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB, LP, TPI, L32> E;
  constexpr int K = E::K, K32 = E::K32;
  constexpr bool U = E::UNIT;  // products modulo N' = kN, the last one modulo N
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * E::AREA;
  E M;
  if constexpr (U)
    M.init(area, md->NK, 1u);
  else
    M.init(area, md->N, md->n0);
  const int item = group_item<TPI>();
  const int it = item < count ? item : count - 1;
  const int r = Grp<TPI>::rank();
  u32 acc[K];
#pragma unroll
  for (int p = 0; p < K; ++p) acc[p] = md->R1[r * K + p];
  bool first = true;  // wave-uniform: acc still holds the Montgomery one
  const unsigned char* dg = digits + (size_t)it * chunks;
  for (int i = 0; i < chunks;) {
    // one step: a row group (i a multiple of prow, wave-uniform test) or a single row
    const int d0 = dg[i];
    const u32* e = gtab + ((size_t)i * 256 + d0) * LP;
    bool act;
    int step = 1;
    if (ptab != nullptr && i % prow == 0 && i + prow <= chunks && i / prow < pgroups) {
      int dor = d0, idx = squeeze4(d0);
      for (int t = 1; t < prow; ++t) {
        const int dt = dg[i + t];
        dor |= dt;
        idx |= squeeze4(dt) << (4 * t);
      }
      if (__ballot((dor & PACKED_ZERO_BITS) != 0) == 0ull) {
        step = prow;
        e = ptab + ((size_t)(i / prow) * ((size_t)1 << (4 * prow)) + idx) * LP;
        act = __ballot(dor != 0) != 0ull;
      } else {
        act = __ballot(d0 != 0) != 0ull;
      }
    } else {
      act = __ballot(d0 != 0) != 0ull;
    }
    i += step;
    if (!act) continue;  // every item multiplies by one: uniform skip
    u32 y[K];
    e += r * K;
#pragma unroll
    for (int p = 0; p < K; ++p) y[p] = e[p];
    if (first) {
#pragma unroll
      for (int p = 0; p < K; ++p) acc[p] = y[p];
      first = false;
    } else {
      M.template mul_into<U>(acc, y);
    }
  }
  const unsigned char* ni = nidx + (size_t)it * nmult;
  const bool plain = ntab_plain != nullptr && (nptab == nullptr || nptab_plain != nullptr);
  bool out_of_form = false;
  for (int j = 0; j < nmult;) {
    // index mod table size, as lcu:933-934
    const int a = ni[j] % nnoise;
    const bool triple = n3tab != nullptr && j + 3 < nmult;  // never the last product
    const bool pair = !triple && nptab != nullptr && j + 1 < nmult;
    const int step = triple ? 3 : pair ? 2 : 1;
    const bool last = plain && j + step >= nmult;
    const u32* e = triple ? n3tab + (((size_t)a * nnoise + ni[j + 1] % nnoise) * nnoise + ni[j + 2] % nnoise) * LP
                   : pair ? (last ? nptab_plain : nptab) + ((size_t)a * nnoise + ni[j + 1] % nnoise) * LP
                          : (last ? ntab_plain : ntab) + (size_t)a * LP;
    j += step;
    u32 y[K];
    e += r * K;
#pragma unroll
    for (int p = 0; p < K; ++p) y[p] = e[p];
    if (last) {
      // acc (below 2N', or 2N) times a plain residue v (at most N), modulo N
      // itself: the result acc v R'^-1 = the plain product, below 2N
      if constexpr (U) M.init(area, md->N, md->n0);
      M.mul_into(acc, y);
      out_of_form = true;
    } else if (first) {
#pragma unroll
      for (int p = 0; p < K; ++p) acc[p] = y[p];
      first = false;
    } else {
      M.template mul_into<U>(acc, y);
    }
  }
  if (!out_of_form) {
    if constexpr (U) M.init(area, md->N, md->n0);  // acc < 2N' < R': MontMul_N(acc, 1) <= N
    u32 one[K];
#pragma unroll
    for (int p = 0; p < K; ++p) one[p] = 0;
    one[0] = r == 0 ? 1u : 0u;
    M.mul_into(acc, one);
  }
  // canonical: acc is below 2N (below N after MontMul(acc, 1)); 2N can pass
  // 2^(32 L32) when N is near it, so the bit above the limbs takes part
  u32 w[K32];
  M.renorm(acc);
  const u32 hi = M.top32(acc);
  M.to32(w, acc);
  u32 n32[K32];
#pragma unroll
  for (int k = 0; k < K32; ++k) n32[k] = md->N32[r * K32 + k];
  cond_sub<K32, TPI>(w, hi, n32);
  if (item < count) {
#pragma unroll
    for (int k = 0; k < K32; ++k) out[(size_t)(r * K32 + k) * stride + item] = w[k];
  }
}

// base_i^(e_i) mod N with a per-item exponent (fixed-window method, HAC
// Alg. 14.82): every item runs the same sequence -- `w` squarings then one
// table multiply per window, `nwin` windows from the top -- so a wave never
// diverges; the table holds x^0 .. x^(2^w - 1) per item. Exponents: e_limbs
// 32-bit limbs per item, interleaved like the other batch buffers; nwin
// covers the widest exponent of the batch.
// One item per group: out[item] = in[item]^exps[item] mod N, fixed window
// `w`, `nwin` windows. Table: 2^w entries of LP positions at
// table[(e * tstride + slot) * LP + pos]. `it` is the clamped item used for
// loads (groups past `count` compute a duplicate and skip the store).
template <int RB, int LP, int TPI, int L32>
__device__ __forceinline__ void modexp_each_item(Lazy<RB, LP, TPI, L32>& M, const LazyDev* __restrict__ md,
                                                 const u32* __restrict__ in, const u32* __restrict__ exps,
                                                 int e_limbs, int nwin, int w, u32* __restrict__ out, int count,
                                                 int stride, u32* __restrict__ table, int tstride, int item, int it,
                                                 int slot) {
  typedef Lazy<RB, LP, TPI, L32> E;
  constexpr int K = E::K, K32 = E::K32;
  constexpr bool U = E::UNIT;  // products modulo N' = kN, the last one modulo N
  if constexpr (U) M.init(M.lds, md->NK, 1u);
  const int r = TPI == 1 ? 0 : opaque(Grp<TPI>::rank());
  auto tstore = [&](int e, const u32 (&v)[K]) {
#pragma unroll
    for (int p = 0; p < K; ++p) table[((size_t)e * tstride + slot) * LP + r * K + p] = v[p];
  };
  auto tload = [&](int e, u32 (&v)[K]) {
#pragma unroll
    for (int p = 0; p < K; ++p) v[p] = table[((size_t)e * tstride + slot) * LP + r * K + p];
  };
  auto digit = [&](int win) -> int {
    const int bit = win * w;
    const int j = bit >> 5, off = bit & 31;
    const u32 lo = j < e_limbs ? exps[(size_t)j * stride + it] : 0u;
    const u32 hi = j + 1 < e_limbs ? exps[(size_t)(j + 1) * stride + it] : 0u;
    return (int)((u32)(make64(lo, hi) >> off) & ((1u << w) - 1u));
  };
  u32 x[K];
  {
    u32 wv[K32];
#pragma unroll
    for (int k = 0; k < K32; ++k) wv[k] = in[(size_t)(r * K32 + k) * stride + it];
    M.from32(x, wv);
    u32 r2[K];
#pragma unroll
    for (int p = 0; p < K; ++p) r2[p] = md->R2[r * K + p];
    M.template mul_into<U>(x, r2);
  }
  {
    u32 cur[K];
#pragma unroll
    for (int p = 0; p < K; ++p) cur[p] = md->R1[r * K + p];
    tstore(0, cur);
    tstore(1, x);
    M.stage(x);
#pragma unroll
    for (int p = 0; p < K; ++p) cur[p] = x[p];
    wave_lds_sync();
    for (int e = 2; e < (1 << w); ++e) {
      M.template mul_staged<U>(cur);
      tstore(e, cur);
    }
    wave_lds_sync();
  }
  tload(digit(nwin - 1), x);
  for (int win = nwin - 2; win >= 0; --win) {
    for (int i = 0; i < w; ++i) M.template sqr<U>(x);
    const int d = digit(win);
    if (__ballot(d != 0) == 0ull) continue;  // whole wave multiplies by one: skip
    u32 y[K];
    tload(d, y);
    M.template mul_into<U>(x, y);
  }
  if constexpr (U) M.init(M.lds, md->N, md->n0);  // x < 2N' < R': MontMul_N(x, 1) <= N
  {
    u32 one[K];
#pragma unroll
    for (int p = 0; p < K; ++p) one[p] = 0;
    one[0] = r == 0 ? 1u : 0u;
    M.mul_into(x, one);
  }
  u32 wv[K32];
  M.renorm(x);
  M.to32(wv, x);
  u32 n32[K32];
#pragma unroll
  for (int k = 0; k < K32; ++k) n32[k] = md->N32[r * K32 + k];
  cond_sub<K32, TPI>(wv, 0u, n32);
  if (item < count) {
#pragma unroll
    for (int k = 0; k < K32; ++k) out[(size_t)(r * K32 + k) * stride + item] = wv[k];
  }
}

// base_i^(e_i) mod N with a per-item exponent (fixed-window method, HAC
// Alg. 14.82): every item runs the same sequence -- `w` squarings then one
// table multiply per window, `nwin` windows from the top -- so a wave never
// diverges; the table holds x^0 .. x^(2^w - 1) per item. Exponents: e_limbs
// 32-bit limbs per item, interleaved like the other batch buffers; nwin
// covers the widest exponent of the batch.
template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(BLOCK_THREADS) void k_lazy_modexp_each(const LazyDev* __restrict__ md,
                                                                    const u32* __restrict__ in,
                                                                    const u32* __restrict__ exps, int e_limbs,
                                                                    int nwin, int w, u32* __restrict__ out,
                                                                    int count, int stride, u32* __restrict__ table,
                                                                    int tstride) {
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB, LP, TPI, L32> E;
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * E::AREA;
  E M;
  M.init(area, md->N, md->n0);
  const int item = group_item<TPI>();
  const int it = item < count ? item : count - 1;
  modexp_each_item<RB, LP, TPI, L32>(M, md, in, exps, e_limbs, nwin, w, out, count, stride, table, tstride, item,
                                     it, item);
}

// ---------------------------------------------------------------------------
// Served Hamming evaluation: many queries against one resident index
// ---------------------------------------------------------------------------
// encode_hamming_server (hamming/paillier.py:214-230; paillier_GPU_client.cu:986-991
// over gpu_pointwise_mulmod_hex, cu:672-732; paillier_GPU_lookup_client.cu:2312-2329
// over gpu_pointwise_mulmod_bigint, lcu:1506) multiplies a query ciphertext into every index
// ciphertext mod n^2. One launch serves up to MQ_MAX queries over one range of
// the index (k_fixed_mulmod below), in (query, tile) work units, query-major.
constexpr int MQ_MAX = 32;
struct MultiQuery {
  int nq;
  const u32* op[MQ_MAX];  // operand tables (k_fixed_operand), device
  u32* out[MQ_MAX];       // interleaved [L32][out_stride], indexed by the global entry
  u32* head;              // head of the queue of (query, tile) units, zeroed before the launch
  u32* done;              // optional (nullptr): finished units per query, zeroed before the launch
  u64* stamp;             // with done: device wall clock when each query's last unit finished (zeroed before the launch)
};

// ---------------------------------------------------------------------------
// Served Hamming evaluation with a fixed query operand
// ---------------------------------------------------------------------------
// For one query y and many index entries x, x * y mod N is computed without
// Montgomery reduction rows: with the query's table
//   Y_i = y * 2^(RB (i + 2)) mod N   (i < LP, lazy: below 2N)
// and x = sum_i x_i 2^(RB i) in RB-bit positions,
//   V = sum_i x_i Y_i  ==  x y 2^(2 RB)   (mod N),
// a product of an LP-vector by an LP x LP matrix: LP^2 multiply-adds instead
// of the 2 LP^2 of a Montgomery product. Two REDC rows (Montgomery 1985) then
// divide by 2^(2 RB): the result is x y mod N below
//   N (LP 2^(RB+1) / 2^(2 RB) + 1) < 2N,
// and one conditional subtraction makes it canonical. A position of V
// receives LP products below 2^RB * (2^RB + 2^(64-2RB) + 1) (settled table
// positions): below 2^64 for every built configuration.
//
// Work split of one launch (a workgroup of TPI waves, tiles of 64 entries):
//  A  wave w converts positions [w K, w K + K) of the tile's 64 entries (one
//     entry per lane) from the index limbs, stored position-major in LDS
//     (xs[i][e]);
//  B  wave w sums the positions [w K, w K + K) of V for all 64 entries, one
//     entry per lane: row i adds x_i[e] * Y_i[w K + p] for p < K. The table
//     chunk is the same for every lane of the wave, so it is read with scalar
//     loads and enters v_mad_u64_u32 as an SGPR operand;
//  C  each lane ripples its K positions (RB-bit) and passes them and its
//     carry through LDS;
//  D  TPI lanes per entry again: add the carry of the chunk below (the top
//     chunk keeps its own carry above the top position), two REDC rows,
//     settle, exact positions, 32-bit limbs, conditional subtraction, store.
// Table layout per query: LP rows of TPI chunks of CH words (K used).
template <int K>
struct FixedShape {
  static constexpr int CH = (K + 3) / 4 * 4;  // chunk words, a multiple of 4 (scalar loads)
};

// Query tables: out[q][i][r CH + p] = y_q * 2^(RB (i + 2)) mod N (positions),
// from y_q (AoS 32-bit limbs, L32 per query) and the modulus constants
// B_i = 2^(RB (i + 2)) R' mod N (AoS positions, LP per row): one lazy product
// y_q * B_i R'^-1 per group (q, i).
template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(BLOCK_THREADS) void k_fixed_operand(const LazyDev* __restrict__ md,
                                                                 const u32* __restrict__ B,
                                                                 const u32* __restrict__ y, u32* __restrict__ out,
                                                                 int count) {
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB, LP, TPI, L32> E;
  constexpr int K = E::K, K32 = E::K32, CH = FixedShape<K>::CH;
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * E::AREA;
  E M;
  M.init(area, md->N, md->n0);
  const int item = group_item<TPI>();
  const int total = count * LP;
  const int it = item < total ? item : total - 1;
  const int q = it / LP, i = it % LP;
  const int r = Grp<TPI>::rank();
  u32 w[K32], x[K], b[K];
#pragma unroll
  for (int k = 0; k < K32; ++k) w[k] = y[(size_t)q * L32 + r * K32 + k];
  M.from32(x, w);
#pragma unroll
  for (int p = 0; p < K; ++p) b[p] = B[(size_t)i * LP + r * K + p];
  M.mul_into(x, b);
  if (item < total) {
#pragma unroll
    for (int p = 0; p < K; ++p) out[((size_t)q * LP + i) * (TPI * CH) + r * CH + p] = x[p];
  }
}

constexpr int FIX_TILE = 64;  // index entries per workgroup tile

// One fold of a chunk's K accumulators (value unchanged): each keeps its low
// RB bits plus the bits of the one below above RB; the top one's high bits go
// to hi, in units of the position above the chunk.
template <int RB, int K>
__device__ __forceinline__ void chunk_fold(u64 (&acc)[K], u64& hi) {
  constexpr u64 MASK = (1ull << RB) - 1ull;
  hi += acc[K - 1] >> RB;
#pragma unroll
  for (int p = K - 1; p >= 1; --p) acc[p] = (acc[p] & MASK) + (acc[p - 1] >> RB);
  acc[0] &= MASK;
}
// table rows per scalar-load batch (FIX_ROWS x K words of SGPRs): 4, or the
// largest of 2, 1 that divides LP
template <int LP>
struct FixedRows {
  static constexpr int R = LP % 4 == 0 ? 4 : LP % 2 == 0 ? 2 : 1;
};

// Rows between accumulator folds of the served multiply. A row adds x_i Y_p
// with x_i < 2^RB and a settled table position Y_p < D = 2^RB + 2^(64-2RB) + 1;
// after a fold a position holds less than 2^RB + 2^(64-RB) (its low bits and
// the carry of the one below). SEG rows plus the finish's two REDC products
// (below 2^(2RB) each) stay below 2^64. SEG is a multiple of FIX_ROWS that
// divides LP, or LP itself when all rows fit (27- and 28-bit shapes).
template <int RB, int LP>
struct FixedSeg {
  static constexpr double TWO64 = 18446744073709551616.0;
  static constexpr double ROW = (double)(1ull << RB) * ((double)(1ull << RB) + (double)(1ull << (64 - 2 * RB)) + 1.0);
  static constexpr double BASE = (double)(1ull << RB) + (double)(1ull << (64 - RB)) + 2.0 * (double)(1ull << (2 * RB));
  static constexpr int MAXR = (int)((TWO64 - BASE) / ROW) - 1;  // one row of margin for the rounding of doubles
  static constexpr int pick(int s) {
    return s <= 0 ? 0 : (LP % s == 0 && s % FixedRows<LP>::R == 0) ? s : pick(s - 1);
  }
  static constexpr int SEG = MAXR >= LP ? LP : pick(MAXR);
  static_assert(SEG > 0 && LP % SEG == 0 && SEG % FixedRows<LP>::R == 0, "no fold interval fits the shape");
  static_assert((double)SEG * ROW + BASE < TWO64, "fold interval overflows the accumulators");
};

// LDS of one k_fixed_mulmod workgroup (u32 words): the tile's positions
// xs[LP][64], then the per-query exchange areas of the finish: REDC digits
// md[2][64], chunk carries cw[TPI][64] (u64), carry lookahead flags
// gp[TPI][64], each chunk's two lowest positions nb[TPI][2][64], borrow
// lookahead flags bp[TPI][64].
template <int LP, int TPI>
struct FixedLds {
  static constexpr int XS = LP * FIX_TILE;
  static constexpr int MD = XS;
  static constexpr int CW = MD + 2 * FIX_TILE;  // even: u64-aligned
  static constexpr int GP = CW + 2 * TPI * FIX_TILE;
  static constexpr int NB = GP + TPI * FIX_TILE;
  static constexpr int BP = NB + 2 * TPI * FIX_TILE;
  static constexpr int WORDS = BP + TPI * FIX_TILE;
};

// Positions [W K, W K + K) of one index entry (K = LP / TPI): position q is
// bits [RB q, RB q + RB) of the entry's 32-bit limbs xb[j * so]. With W a
// template argument every limb index and bit offset is a constant: the limbs
// the chunk spans are loaded once, each position is one funnel shift and a
// mask. Stored to out[q * FIX_TILE].
template <int RB, int LP, int TPI, int L32, int W>
__device__ __forceinline__ void fixed_positions_w(const u32* __restrict__ xb, size_t so, u32* out) {
  constexpr int K = LP / TPI;
  constexpr u32 MASK = (1u << RB) - 1u;
  constexpr int J0 = (RB * W * K) >> 5;                  // first limb of the chunk
  constexpr int J1 = (RB * (W * K + K) - 1) >> 5;        // last limb (may exceed L32 - 1)
  constexpr int NL = J1 - J0 + 2;                        // limbs read, plus one above
  u32 lim[NL];
#pragma unroll
  for (int t = 0; t < NL; ++t) lim[t] = J0 + t < L32 ? xb[(size_t)(J0 + t) * so] : 0u;
#pragma unroll
  for (int p = 0; p < K; ++p) {
    const int q = W * K + p;
    const int b = RB * q, t = (b >> 5) - J0, o = b & 31;
    out[q * FIX_TILE] = (u32)(make64(lim[t], lim[t + 1]) >> o) & MASK;
  }
}

template <int RB, int LP, int TPI, int L32, int W = 0>
__device__ __forceinline__ void fixed_positions(int wave, const u32* __restrict__ xb, size_t so, u32* out) {
  if constexpr (W < TPI) {
    if (wave == W)
      fixed_positions_w<RB, LP, TPI, L32, W>(xb, so, out);
    else
      fixed_positions<RB, LP, TPI, L32, W + 1>(wave, xb, so, out);
  }
}

typedef __attribute__((address_space(4))) const u32 cu32;  // read-only, wave-uniform: scalar loads

// Finish of one query for chunk W (wave W, one entry per lane; W a template
// argument, so every position and limb index is a constant). acc holds
// positions [W K, W K + K) of V = sum_i x_i Y_i. Steps (each exchange through
// LDS behind a workgroup barrier, five per query):
//  1  wave 0 computes the two REDC digits m1, m2 from positions 0 and 1;
//  2  every chunk adds (m1 + m2 2^RB) N: V is now divisible by 2^(2 RB);
//  3  each chunk ripples to RB-bit positions and hands its carry (< 2^37) to
//     the chunk above, which adds it at its position 0 and ripples again: the
//     carry out is then one bit, and a chunk propagates a carry only if all its
//     positions are MASK; a generate/propagate lookahead over the chunks
//     gives each its carry-in, and a last ripple makes every position exact.
//     The top chunk keeps what leaves position LP-1 (positions LP, LP+1);
//  4  R = V / 2^(2 RB) (below 2N): chunk W forms the 32-bit limbs whose lowest
//     bit lies in its positions, with the two lowest positions of the chunk
//     above for limbs that straddle the boundary;
//  5  R - N on those limbs; a borrow lookahead over the chunks decides R >= N
//     (or bit 32 L32 of R set), and each chunk stores R or R - N.
// `hi` is what the chunk's accumulator folds moved above its top position
// (in units of the chunk above's position 0; 0 when no fold ran): it joins the
// chunk's carry in step 3. `mid` runs right after the first barrier: every
// wave of the workgroup has then finished whatever it did before the finish
// (the served kernel converts its next tile into the position rows there).
struct NoMid {
  __device__ __forceinline__ void operator()() const {}
};
template <int RB, int LP, int TPI, int L32, int W, class Mid = NoMid>
__device__ __forceinline__ void fixed_finish_w(const LazyDev* __restrict__ md, u64 (&acc)[LP / TPI], u32* lds,
                                               int lane, u32* __restrict__ ob, size_t so, bool store,
                                               u64 hi = 0, const Mid& mid = Mid()) {
  constexpr int K = LP / TPI, T = TPI;
  constexpr u32 MASK = (1u << RB) - 1u;
  typedef FixedLds<LP, TPI> S;
  u32* mdg = lds + S::MD;
  u64* cw = reinterpret_cast<u64*>(lds + S::CW);
  u32* gp = lds + S::GP;
  u32* nb = lds + S::NB;
  u32* bp = lds + S::BP;
  const cu32* Nc = (const cu32*)md->N;
  const cu32* N32 = (const cu32*)md->N32;
  static_assert(K >= 2, "positions 0 and 1 must lie in chunk 0");
  // 1: REDC digits (positions 0 and 1 are complete sums: nothing lies below)
  if constexpr (W == 0) {
    const u32 n0 = *(const cu32*)&md->n0;
    const u32 m1 = ((u32)acc[0] * n0) & MASK;
    const u64 t0 = acc[0] + (u64)m1 * Nc[0];
    const u64 v1 = acc[1] + (u64)m1 * Nc[1] + (t0 >> RB);
    const u32 m2 = ((u32)v1 * n0) & MASK;
    mdg[lane] = m1;
    mdg[FIX_TILE + lane] = m2;
  }
  __syncthreads();
  mid();
  // 2: add (m1 + m2 2^RB) N on this chunk (the top chunk's m2 N_(LP-1) lies at
  // position LP, above every chunk: it joins the top carry in step 3)
  const u32 m2 = mdg[FIX_TILE + lane];
  {
    const u32 m1 = mdg[lane];
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const int q = W * K + p;
      acc[p] = (u64)m1 * Nc[q] + acc[p];
      if (q >= 1) acc[p] = (u64)m2 * Nc[q - 1] + acc[p];
    }
  }
  // 3: exact positions
  u32 X[K];
  u64 ctop = 0;
  {
    u64 c = 0;
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const u64 t = acc[p] + c;
      X[p] = (u32)t & MASK;
      c = t >> RB;
    }
    if constexpr (W < T - 1) cw[W * FIX_TILE + lane] = c + hi;  // below 2^37 (three shares of 2^(64-RB) at most)
    else ctop = c + hi + (u64)m2 * Nc[LP - 1];  // everything at positions >= LP (below 2^47: V < 2N 2^(2RB))
  }
  __syncthreads();
  {
    u64 cin = 0;
    if constexpr (W > 0) cin = cw[(W - 1) * FIX_TILE + lane];
    u32 all = MASK;
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const u64 t = (u64)X[p] + cin;
      X[p] = (u32)t & MASK;
      cin = t >> RB;
      all &= X[p];
    }
    if constexpr (W < T - 1) gp[W * FIX_TILE + lane] = (u32)cin | (all == MASK ? 2u : 0u);
    else ctop += cin;
  }
  __syncthreads();
  {
    u32 cin = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const u32 f = gp[w * FIX_TILE + lane];
      cin = (f & 1u) | ((f >> 1) & cin);
    }
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const u32 t = X[p] + cin;
      X[p] = t & MASK;
      cin = t >> RB;
    }
    if constexpr (W == T - 1) ctop += cin;
  }
  // 4: limbs of R whose lowest bit lies in this chunk
  nb[(2 * W) * FIX_TILE + lane] = X[0];
  nb[(2 * W + 1) * FIX_TILE + lane] = X[1];
  __syncthreads();
  u32 up0, up1;  // positions W K + K and W K + K + 1
  if constexpr (W < T - 1) {
    up0 = nb[(2 * W + 2) * FIX_TILE + lane];
    up1 = nb[(2 * W + 3) * FIX_TILE + lane];
  } else {
    up0 = (u32)ctop & MASK;
    up1 = (u32)(ctop >> RB);
  }
  auto pos = [&](int q) -> u32 {  // q in [W K, W K + K + 2); 0 above
    const int o = q - W * K;
    return o < K ? X[o] : o == K ? up0 : o == K + 1 ? up1 : 0u;
  };
  constexpr int B0 = RB * W * K - 2 * RB;        // R bits of this chunk: [B0, B1)
  constexpr int B1 = RB * (W * K + K) - 2 * RB;
  constexpr int J0 = B0 <= 0 ? 0 : (B0 + 31) / 32;
  constexpr int J1 = W == T - 1 ? L32 : ((B1 + 31) / 32 < L32 ? (B1 + 31) / 32 : L32);
  constexpr int NJ = J1 - J0 > 0 ? J1 - J0 : 1;
  u32 rl[NJ];
#pragma unroll
  for (int j = J0; j < J1; ++j) {
    const int b = 32 * j + 2 * RB, q = b / RB, o = b % RB;
    u64 v = (u64)pos(q) >> o;
    v |= (u64)pos(q + 1) << (RB - o);
    if (2 * RB - o < 64) v |= (u64)pos(q + 2) << (2 * RB - o);
    rl[j - J0] = (u32)v;
  }
  // 5: R - N, borrow lookahead, store R or R - N
  u32 bo = 0, eq = 1;
#pragma unroll
  for (int j = J0; j < J1; ++j) {
    const u64 d = (u64)rl[j - J0] - N32[j] - bo;
    bo = (u32)(d >> 63);
    eq &= rl[j - J0] == N32[j] ? 1u : 0u;
  }
  u32 tb = 0;
  if constexpr (W == T - 1) {
    constexpr int b = 32 * L32 + 2 * RB, q = b / RB, o = b % RB;
    tb = (pos(q) >> o) & 1u;
  }
  bp[W * FIX_TILE + lane] = bo | (eq << 1) | (tb << 2);
  __syncthreads();
  u32 bin = 0, chain = 0, top = 0;
#pragma unroll
  for (int w = 0; w < T; ++w) {
    const u32 f = bp[w * FIX_TILE + lane];
    if (w == W) bin = chain;
    chain = (f & 1u) | ((f >> 1) & 1u & chain);
    if (w == T - 1) top = (f >> 2) & 1u;
  }
  const bool sub = top != 0 || chain == 0;  // R >= N
  if (store) {
    u32 b = bin;
#pragma unroll
    for (int j = J0; j < J1; ++j) {
      const u64 d = (u64)rl[j - J0] - N32[j] - b;
      b = (u32)(d >> 63);
      ob[(size_t)j * so] = sub ? (u32)d : rl[j - J0];
    }
  }
}

template <int RB, int LP, int TPI, int L32, int W = 0>
__device__ __forceinline__ void fixed_finish(int wave, const LazyDev* __restrict__ md, u64 (&acc)[LP / TPI],
                                             u32* lds, int lane, u32* __restrict__ ob, size_t so, bool store,
                                             u64 hi = 0) {
  if constexpr (W < TPI) {
    if (wave == W)
      fixed_finish_w<RB, LP, TPI, L32, W>(md, acc, lds, lane, ob, so, store, hi);
    else
      fixed_finish<RB, LP, TPI, L32, W + 1>(wave, md, acc, lds, lane, ob, so, store, hi);
  }
}

// Index limbs of chunk W of one entry: the limbs [J0, J0 + NL) that positions
// [W K, W K + K) span, plus one above (zero past the top limb).
template <int RB, int LP, int TPI, int L32, int W>
struct FixedChunk {
  static constexpr int K = LP / TPI;
  static constexpr int J0 = (RB * W * K) >> 5;
  static constexpr int NL = ((RB * (W * K + K) - 1) >> 5) - J0 + 2;
};

// The units of one launch, run by wave W of every workgroup (W a template
// argument: the limb indices, bit offsets and the prefetch registers are
// constants). Two units are in hand at any time: the workgroup works on unit u
// while the index limbs of the next unit un are on their way into registers
// (loaded before the row loop, converted into the position rows by the
// finish's hook, after its first barrier, when every wave has left the row
// loop). The queue head is bumped mid-unit, so the grab of the unit after un
// is hidden behind the rest of the finish. A unit costs six barriers.
template <int RB, int LP, int TPI, int L32, int W>
__device__ __forceinline__ void fixed_units(const LazyDev* __restrict__ md, const u32* __restrict__ x, int begin,
                                            int count, int stride, const MultiQuery& mq, int out_stride,
                                            u32* lds, int* slot) {
  constexpr int K = LP / TPI, CH = FixedShape<K>::CH;
  constexpr u32 MASK = (1u << RB) - 1u;
  typedef FixedChunk<RB, LP, TPI, L32, W> C;
  u32* xs = lds;  // [LP][FIX_TILE]
  const int lane = lane_id();
  const int tiles = (count + FIX_TILE - 1) / FIX_TILE;
  // Work units in query-major order: the workgroups sweep the range for one
  // query before the next, so at any time they read one or two operand tables
  // (92 KB each at 4096-bit moduli) instead of all of them. Units come from a
  // queue, so a fast workgroup takes more units than a slow one (strided
  // units leave the launch waiting for its slowest workgroups, 14% slower
  // measured) and the queries complete one after another at the pace of the
  // grid.
  const int units = tiles * mq.nq;
  u32 lim[C::NL];
  // an opaque stride keeps the per-limb addresses from being hoisted out of
  // the unit loop (they would be held, or spilled, across it)
  auto fetch = [&](int unit) {
    const int base = (unit % tiles) * FIX_TILE;
    const int local = base + lane < count ? base + lane : count - 1;
    const size_t so = (size_t)opaque(stride);
    const u32* xb = x + (size_t)begin + local;
#pragma unroll
    for (int t = 0; t < C::NL; ++t) lim[t] = C::J0 + t < L32 ? xb[(size_t)(C::J0 + t) * so] : 0u;
  };
  auto convert = [&]() {
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const int q = W * K + p;
      const int b = RB * q, t = (b >> 5) - C::J0, o = b & 31;
      xs[q * FIX_TILE + lane] = __builtin_amdgcn_alignbit(lim[t + 1], lim[t], o) & MASK;
    }
  };
  int u = slot[0], un = slot[1];
  if (u < units) {
    fetch(u);
    convert();
  }
  if (un < units) fetch(un);
  __syncthreads();  // position rows of unit u
  while (u < units) {
    const int q = u / tiles, tile = u - q * tiles;
    const int base = tile * FIX_TILE;
    // positions [W K, W K + K) of V for entry `lane`; the table does not
    // change during the launch: read through the constant address space, so
    // the wave-uniform loads are scalar
    const cu32* Y = (const cu32*)(mq.op[q] + W * CH);
    u64 acc[K];
#pragma unroll
    for (int p = 0; p < K; ++p) acc[p] = 0;
    // Rows in batches of FIX_ROWS: the batch's scalar loads (and its x
    // reads) are issued together, so one wait covers them all; a scalar
    // load can return out of order, so every use of one waits for all.
    // Every SEG rows (29-bit shapes only) the accumulators are folded: each
    // keeps its low RB bits plus the bits of the one below above RB, and the
    // top one's high bits move to hi (the value is unchanged).
    constexpr int FIX_ROWS = FixedRows<LP>::R;
    constexpr int SEG = FixedSeg<RB, LP>::SEG;
    u64 hi = 0;
#pragma unroll 1
    for (int s0 = 0; s0 < LP; s0 += SEG) {
#pragma unroll 1
      for (int i0 = s0; i0 < s0 + SEG; i0 += FIX_ROWS) {
        u32 yv[FIX_ROWS][K], xv[FIX_ROWS];
#pragma unroll
        for (int j = 0; j < FIX_ROWS; ++j) {
          const cu32* yr = Y + (size_t)(i0 + j) * (TPI * CH);
#pragma unroll
          for (int p = 0; p < K; ++p) yv[j][p] = yr[p];
          xv[j] = xs[(i0 + j) * FIX_TILE + lane];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < FIX_ROWS; ++j) {
#pragma unroll
          for (int p = 0; p < K; ++p) acc[p] = (u64)xv[j] * yv[j][p] + acc[p];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      if constexpr (SEG < LP) {
        if (s0 + SEG < LP) chunk_fold<RB, K>(acc, hi);
      }
    }
    // reduction, exact positions, canonical 32-bit limbs, store; after the
    // finish's first barrier the position rows are free: the next unit's
    // positions go there, and the unit after it is taken from the queue
    const bool more = un < units;
    auto mid = [&]() {
      if (more) convert();
      if (threadIdx.x == 0) slot[0] = (int)atomicAdd(mq.head, 1u);
    };
    const size_t so = (size_t)opaque(out_stride);
    fixed_finish_w<RB, LP, TPI, L32, W>(md, acc, lds, lane, mq.out[q] + (size_t)begin + base + lane, so,
                                        base + lane < count, hi, mid);
    __syncthreads();  // every wave has issued the unit's stores; slot[0] and the next positions are set
    // a query is complete when its last unit is: each unit counts itself and
    // raises its query's stamp to the device clock at its end, so the stamp
    // ends as the clock of the query's last unit. Both atomics return nothing:
    // no unit waits for them. (Timing records, read after the launch, not a
    // hand-over of data: no fence, which would write the L2 back on every unit.)
    if (mq.done != nullptr && threadIdx.x == 0) {
      atomicAdd(mq.done + q, 1u);
      atomicMax(reinterpret_cast<unsigned long long*>(mq.stamp + q),
                (unsigned long long)__builtin_amdgcn_s_memrealtime());
    }
    u = un;
    un = slot[0];  // written before the barrier above, rewritten only after the next unit's first finish barrier
    if (u < units && un < units) fetch(un);
  }
}

template <int RB, int LP, int TPI, int L32, int W = 0>
__device__ __forceinline__ void fixed_units_for(int wave, const LazyDev* __restrict__ md, const u32* __restrict__ x,
                                                int begin, int count, int stride, const MultiQuery& mq,
                                                int out_stride, u32* lds, int* slot) {
  if constexpr (W < TPI) {
    if (wave == W)
      fixed_units<RB, LP, TPI, L32, W>(md, x, begin, count, stride, mq, out_stride, lds, slot);
    else
      fixed_units_for<RB, LP, TPI, L32, W + 1>(wave, md, x, begin, count, stride, mq, out_stride, lds, slot);
  }
}

template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(64 * TPI) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_fixed_mulmod(const LazyDev* __restrict__ md,
                                                           const u32* __restrict__ x, int begin, int count,
                                                           int stride, MultiQuery mq, int out_stride) {
  extern __shared__ uint4 smem4[];
  static_assert(LP % TPI == 0, "LP must be a multiple of TPI");
  __shared__ int slot[2];  // the unit in hand and the next one
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
  if (threadIdx.x == 0) {
    slot[0] = (int)atomicAdd(mq.head, 1u);
    slot[1] = (int)atomicAdd(mq.head, 1u);
  }
  __syncthreads();
  fixed_units_for<RB, LP, TPI, L32>(wave, md, x, begin, count, stride, mq, out_stride,
                                    reinterpret_cast<u32*>(smem4), slot);
}

// ---------------------------------------------------------------------------
// Pointwise multiply with one reduction: out[i] = a[i] b[i] mod N for any
// a, b < 2^(32 L32) (the evaluator's ciphertext multiply, encode_hamming_server:
// cu:672-732, lcu:1505-1561)
// ---------------------------------------------------------------------------
// A workgroup of TPI waves takes tiles of 64 items (one per lane), as
// k_fixed_mulmod; K = LP / TPI positions per chunk.
//  1  each wave converts its K positions of a and b into LDS (as, bs);
//  2  product T = a b in 2 LP positions: wave w sums chunks w and TPI + w
//     (the triangle of the schoolbook product gives them w + 1 and TPI - w
//     blocks of K rows, so every wave does LP^2 / TPI multiply-adds);
//  3  each chunk ripples to RB-bit positions and hands its carry to the chunk
//     above, which adds it at its bottom position (settled positions, below
//     2^RB + 2^(64 - 2RB) + 1);
//  4  fold: with 2^(RB h) <= N / 2, V = T_low 2^(2RB) + sum_{t >= h} T_t Z_t,
//     Z_t = 2^(RB (t + 2)) mod N (a per-modulus table, canonical positions,
//     read with scalar loads as the query table of k_fixed_mulmod), is
//     = a b 2^(2RB) (mod N) and below N 2^(2RB) (1/2 + nz 2^(1 - RB)):
//     T_low 2^(2RB) is T's positions below h shifted up by two, and
//     the nz rows t >= h add nz multiply-adds per position;
//  5  the finish of k_fixed_mulmod (two REDC rows, exact positions, limbs,
//     one conditional subtraction): R = (V + m N) / 2^(2RB) < 2N.
// About 2 LP^2 multiply-adds per item (LP^2 + nz LP), against the 4 LP^2 of
// two Montgomery products.
// LDS (u32 words, one region of 2 LP rows of 64): as rows [0, LP), bs rows
// [LP, 2 LP) for steps 1-2; then the folded positions ts rows [0, nz), the
// finish's exchange areas rows [FB, FB + 2 + 6 TPI), the chunk boundary
// positions bd rows [2 LP - 6 TPI, 2 LP - 4 TPI) and the chunk carries cx
// (u64) rows [2 LP - 4 TPI, 2 LP). The host guarantees nz <= FB, so the
// finish can start while other waves still fold.
// The fold table (nz rows of LP positions) is grouped for scalar loads: the
// two rows 2g, 2g + 1 of wave w's chunk are 2 CH consecutive words at
// ((g TPI + w) 2 CH) (CH = K rounded up to 4, zero padded).
template <int LP, int TPI>
struct PwLds {
  static constexpr int WORDS = 2 * LP * FIX_TILE;
  static constexpr int FB = 2 * LP - 12 * TPI - 2;  // first row of the finish's areas
  static constexpr int BD = (2 * LP - 6 * TPI) * FIX_TILE;
  static constexpr int CX = (2 * LP - 4 * TPI) * FIX_TILE;
  // fixed_finish addresses its areas from XS = LP rows past the pointer it gets
  static constexpr int FIN = (FB - LP) * FIX_TILE;
  static_assert(FB >= LP, "finish areas below the as rows' end");
  static_assert(FixedLds<LP, TPI>::WORDS - FixedLds<LP, TPI>::XS <= (2 * LP - 6 * TPI - FB) * FIX_TILE,
                "finish areas overlap the boundary rows");
};

// K rows of a product chunk (step 2). Row s multiplies a_i (i = the block's
// first row + s) by the window b_(j0 - s + p), p < K, held in a ring of K
// registers: slot (p - s) mod K; the value entering at row s (p = 0) is read
// from bq[(K - 1 - s) * FIX_TILE] (bq: the entering value of the block's last
// row). MODE 0: every entering value is read. MODE 1 (last block of a low
// chunk, j0 = 0): values entering after row 0 lie below b_0, so the terms
// p < s are skipped. MODE 2 (first block of a high chunk, j0 = LP): the value
// entering at row 0 lies above b_(LP-1), and terms p >= s are skipped.
template <int K, int MODE>
__device__ __forceinline__ void pw_block(u64 (&acc)[K], u32 (&ring)[K], const u32* ap, const u32* bq) {
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const u32 a = ap[s * FIX_TILE];
    const int slot = (K - s) % K;
    if (MODE == 0 || (MODE == 1 && s == 0) || (MODE == 2 && s > 0)) ring[slot] = bq[(K - 1 - s) * FIX_TILE];
#pragma unroll
    for (int p = 0; p < K; ++p) {
      if (MODE == 1 && p < s) continue;
      if (MODE == 2 && p >= s) continue;
      acc[p] = (u64)a * ring[(p - s + K) % K] + acc[p];
    }
  }
}

// Fold intervals of k_pw_mulmod (29-bit shapes; the 27- and 28-bit ones never
// fold). PB: product blocks of K rows between folds (a row adds a_i b_j <
// 2^(2RB) to a position; a folded position holds less than 2^RB +
// 2^(64-RB)). FS: fold-table rows between folds, a multiple of 4: a row adds a
// T position (below 2^RB + 3 2^(64-2RB) + 1 after the chunk carries, which
// include what the product folds moved above each chunk) times a canonical Z
// position, and the finish's two REDC products follow the last segment.
template <int RB, int LP, int TPI>
struct PwSeg {
  static constexpr int K = LP / TPI;
  static constexpr double TWO64 = 18446744073709551616.0;
  static constexpr double P2 = (double)(1ull << (2 * RB));
  static constexpr int PROWS = (int)((TWO64 - (double)(1ull << RB) - (double)(1ull << (64 - RB))) / P2) - 1;
  static constexpr int PB = PROWS >= TPI * K ? TPI : (PROWS / K > 0 ? PROWS / K : 1);
  static constexpr double ROW = (double)(1ull << RB) * ((double)(1ull << RB) + 3.0 * (double)(1ull << (64 - 2 * RB)) + 1.0);
  static constexpr double BASE = (double)(1ull << RB) + (double)(1ull << (64 - RB)) + 2.0 * P2;
  static constexpr int FS = ((int)((TWO64 - BASE) / ROW) - 1) / 4 * 4;
  static_assert(PB >= 1 && (double)(PB * K) * P2 + (double)(1ull << RB) + (double)(1ull << (64 - RB)) < TWO64,
                "product fold interval overflows");
  static_assert(FS >= 4 && (double)FS * ROW + BASE < TWO64, "fold-table fold interval overflows");
};

// Positions [c K, c K + K) of the product (chunk c; LOW: c < TPI) of the
// operands in LDS (as_l, bs_l: column of this lane, rows of FIX_TILE words);
// hi receives what the folds move above the chunk (0 without folds).
template <int RB, int LP, int TPI, bool LOW>
__device__ __forceinline__ void pw_chunk(u64 (&acc)[LP / TPI], u64& hi, const u32* as_l, const u32* bs_l, int c) {
  constexpr int K = LP / TPI;
  constexpr int PB = PwSeg<RB, LP, TPI>::PB;
  const int nblocks = LOW ? c + 1 : 2 * TPI - c;  // wave-uniform
  hi = 0;
#pragma unroll
  for (int p = 0; p < K; ++p) acc[p] = 0;
  u32 ring[K];
  // rows i from i0; the value entering at row i0 is b_(j0), j0 = c K - i0
  const int i0 = LOW ? 0 : (c - TPI) * K;
  const int j0 = LOW ? c * K : LP;
  const u32* ap = as_l + i0 * FIX_TILE;
  const u32* bq = bs_l + (j0 - (K - 1)) * FIX_TILE;
  if constexpr (LOW) {
#pragma unroll
    for (int p = 1; p < K; ++p) ring[p] = bs_l[(j0 + p) * FIX_TILE];  // b_(cK + p), slot p
#pragma unroll 1
    for (int blk = 0; blk < c; ++blk) {
      pw_block<K, 0>(acc, ring, ap, bq);
      if constexpr (PB < TPI) {
        if ((blk + 1) % PB == 0) chunk_fold<RB, K>(acc, hi);  // another block follows
      }
      ap += K * FIX_TILE;
      bq -= K * FIX_TILE;
    }
    pw_block<K, 1>(acc, ring, ap, bq);
  } else {
#pragma unroll
    for (int p = 0; p < K; ++p) ring[p] = 0;  // never read before written (MODE 2 skips them)
    pw_block<K, 2>(acc, ring, ap, bq);
#pragma unroll 1
    for (int blk = 1; blk < 2 * TPI - c; ++blk) {
      if constexpr (PB < TPI) {
        if (blk % PB == 0) chunk_fold<RB, K>(acc, hi);  // blk blocks done, one more follows
      }
      ap += K * FIX_TILE;
      bq -= K * FIX_TILE;
      pw_block<K, 0>(acc, ring, ap, bq);
    }
  }
  (void)nblocks;
}

template <int RB, int K>
__device__ __forceinline__ u64 pw_ripple(u32 (&x)[K], const u64 (&acc)[K]) {
  constexpr u32 MASK = (1u << RB) - 1u;
  u64 c = 0;
#pragma unroll
  for (int p = 0; p < K; ++p) {
    const u64 t = acc[p] + c;
    x[p] = (u32)t & MASK;
    c = t >> RB;
  }
  return c;
}

template <int RB, int K>
__device__ __forceinline__ void pw_carry_in(u32 (&x)[K], u64 cin) {
  constexpr u32 MASK = (1u << RB) - 1u;
  const u64 t = (u64)x[0] + cin;
  x[0] = (u32)t & MASK;
  x[1] += (u32)(t >> RB);
}

template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(64 * TPI) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_pw_mulmod(
    const LazyDev* __restrict__ md, const u32* __restrict__ Z, int h, int nz, const u32* __restrict__ A,
    const u32* __restrict__ B, u32* __restrict__ out, int count, int stride, u32* __restrict__ head) {
  extern __shared__ uint4 smem4[];
  constexpr int K = LP / TPI, CH = FixedShape<K>::CH;
  static_assert(LP % TPI == 0 && K >= 2, "chunks of at least two positions");
  typedef PwLds<LP, TPI> S;
  u32* lds = reinterpret_cast<u32*>(smem4);
  u32* as = lds;
  u32* bs = lds + LP * FIX_TILE;
  u32* ts = lds;
  u32* bd = lds + S::BD;
  u64* cx = reinterpret_cast<u64*>(lds + S::CX);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
  const int lane = lane_id();
  const int tiles = (count + FIX_TILE - 1) / FIX_TILE;
  // tiles from a queue (head zeroed before the launch): a fast workgroup takes
  // more tiles than a slow one
  __shared__ int next_tile;
  for (;;) {
    if (threadIdx.x == 0) next_tile = (int)atomicAdd(head, 1u);
    __syncthreads();
    const int tile = next_tile;  // the previous tile ended with a barrier
    if (tile >= tiles) break;
    const int base = tile * FIX_TILE;
    {  // 1: positions of a and b
      const int local = base + lane < count ? base + lane : count - 1;
      const size_t so = (size_t)opaque(stride);
      fixed_positions<RB, LP, TPI, L32>(wave, A + local, so, as + lane);
      fixed_positions<RB, LP, TPI, L32>(wave, B + local, so, bs + lane);
    }
    __syncthreads();
    // 2: product chunks wave and TPI + wave
    u32 xl[K], xh[K];
    u64 cl, chi;
    {
      u64 acc[K], hi;
      pw_chunk<RB, LP, TPI, true>(acc, hi, as + lane, bs + lane, wave);
      cl = pw_ripple<RB, K>(xl, acc) + hi;  // below 3 2^(64-RB)
      // the low chunk leaves its 64-bit accumulators here: K words stay live
      // across the high chunk instead of 2K
#pragma unroll
      for (int p = 0; p < K; ++p) asm volatile("" : "+v"(xl[p]));
      asm volatile("" : "+v"(cl));
      __builtin_amdgcn_sched_barrier(0);
      pw_chunk<RB, LP, TPI, false>(acc, hi, as + lane, bs + lane, TPI + wave);
      chi = pw_ripple<RB, K>(xh, acc) + hi;  // the top chunk's carry is 0: a b < 2^(RB 2 LP)
    }
    __syncthreads();  // as, bs read by every wave
    // 3: chunk carries
    cx[wave * FIX_TILE + lane] = cl;
    cx[(TPI + wave) * FIX_TILE + lane] = chi;
    __syncthreads();
    if (wave > 0) pw_carry_in<RB, K>(xl, cx[(wave - 1) * FIX_TILE + lane]);
    pw_carry_in<RB, K>(xh, cx[(TPI + wave - 1) * FIX_TILE + lane]);
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const int t = wave * K + p;
      if (t >= h) ts[(t - h) * FIX_TILE + lane] = xl[p];
      ts[(LP + t - h) * FIX_TILE + lane] = xh[p];
    }
    bd[(2 * wave) * FIX_TILE + lane] = xl[K - 2];
    bd[(2 * wave + 1) * FIX_TILE + lane] = xl[K - 1];
    __syncthreads();
    // 4: V positions [wave K, wave K + K)
    u64 acc[K], vhi = 0;
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const int t = wave * K + p - 2;  // T_t shifted up by two positions
      u32 v;
      if (p >= 2)
        v = xl[p - 2];
      else
        v = wave > 0 ? bd[(2 * (wave - 1) + p) * FIX_TILE + lane] : 0u;
      acc[p] = (t >= 0 && t < h) ? (u64)v : 0ull;
    }
    {
      // rows in batches of 4 (two pairs of the table's layout): the batch's
      // scalar loads and its ts reads are issued together, so one wait covers
      // them all (a scalar load can return out of order).
      const cu32* Zw = (const cu32*)(Z + wave * 2 * CH);
      constexpr int FS = PwSeg<RB, LP, TPI>::FS;
#pragma unroll 1
      for (int r0 = 0; r0 < nz; r0 += 4) {  // nz is a multiple of 4
        u32 zv[4][K], xv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const cu32* zr = Zw + (size_t)((r0 + j) >> 1) * (TPI * 2 * CH) + (j & 1) * CH;
#pragma unroll
          for (int p = 0; p < K; ++p) zv[j][p] = zr[p];
          xv[j] = ts[(r0 + j) * FIX_TILE + lane];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
          for (int p = 0; p < K; ++p) acc[p] = (u64)xv[j] * zv[j][p] + acc[p];
        }
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (FS < 2 * LP) {  // 29-bit shapes: fold every FS rows (wave-uniform)
          if ((r0 + 4) % FS == 0 && r0 + 4 < nz) chunk_fold<RB, K>(acc, vhi);
        }
      }
    }
    // 5: reduction, exact positions, canonical limbs, store (the finish's
    // areas lie above every ts row: no barrier before it)
    fixed_finish<RB, LP, TPI, L32>(wave, md, acc, lds + S::FIN, lane, out + base + lane, (size_t)opaque(stride),
                                   base + lane < count, vhi);
    __syncthreads();  // the next tile's positions overwrite the finish's areas
  }
}

// ---------------------------------------------------------------------------
// Ragged batches: one persistent launch over many sub-batches of mixed width
// ---------------------------------------------------------------------------
struct RaggedSub {
  const LazyDev* md;  // constants of this sub-batch's modulus (configuration of its width)
  const u32* base;    // interleaved [L32][stride]
  const u32* exps;    // interleaved [e_limbs][stride]
  u32* out;           // interleaved [L32][stride]
  int width;          // 32-bit limbs: 64 or 128
  int count, stride, e_limbs, nwin;
  int tile0;          // first queue tile of this sub-batch
};

// the device wall clock the multi-query stamps count, for calibration
__global__ void k_clock(u64* __restrict__ out) {
  if (threadIdx.x == 0) *out = __builtin_amdgcn_s_memrealtime();
}

__global__ void k_fill_u32(u32* __restrict__ p, u32 v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Persistent work-queue loop of one wave: take tiles (one tile = the groups
// of one wave) from `counter` until it passes `total_tiles`; map each tile to
// its sub-batch by binary search over the tile prefix offsets. Every wave
// reaches the exit test after every tile, so the grid drains.
__device__ __forceinline__ int ragged_next(u32* counter, int total_tiles, const RaggedSub* __restrict__ subs,
                                           int nsubs, int* local) {
  u32 t = 0;
  if (lane_id() == 0) t = atomicAdd(counter, 1u);
  t = (u32)__builtin_amdgcn_readfirstlane((int)t);
  if ((int)t >= total_tiles) return -1;
  int lo = 0, hi = nsubs - 1;
  while (lo < hi) {  // last sub with tile0 <= t
    const int mid = (lo + hi + 1) >> 1;
    if (subs[mid].tile0 <= (int)t) lo = mid; else hi = mid - 1;
  }
  *local = (int)t - subs[lo].tile0;
  return lo;
}

template <int RB, int LP, int TPI, int L32>
__device__ __forceinline__ void ragged_tile(const RaggedSub* __restrict__ sbp, int local, u32* area, u32* table,
                                            int w) {
  typedef Lazy<RB, LP, TPI, L32> E;
  const RaggedSub sb = *sbp;
  E M;
  M.init(area, sb.md->N, sb.md->n0);
  const int slot = Grp<TPI>::gid();
  const int item = local * E::G + slot;
  const int it = item < sb.count ? item : sb.count - 1;
  modexp_each_item<RB, LP, TPI, L32>(M, sb.md, sb.base, sb.exps, sb.e_limbs, sb.nwin, w, sb.out, sb.count,
                                     sb.stride, table, E::G, item, it, slot);
}

// Persistent kernel over sub-batches of one width.
template <int RB, int LP, int TPI, int L32>
__global__ __launch_bounds__(BLOCK_THREADS) void k_ragged1(const RaggedSub* __restrict__ subs, int nsubs,
                                                           int total_tiles, u32* __restrict__ counter, int w,
                                                           u32* __restrict__ scratch, size_t wave_words) {
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB, LP, TPI, L32> E;
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * E::AREA;
  u32* table = scratch + ((size_t)blockIdx.x * (BLOCK_THREADS / 64) + threadIdx.x / 64) * wave_words;
  int local = 0;
  for (int s; (s = ragged_next(counter, total_tiles, subs, nsubs, &local)) >= 0;)
    ragged_tile<RB, LP, TPI, L32>(subs + s, local, area, table, w);
}

// Persistent kernel over sub-batches of both widths (64- and 128-limb moduli):
// the width of a tile's sub-batch selects the engine (a wave-uniform branch).
template <int RB1, int LP1, int TPI1, int RB2, int LP2, int TPI2>
__global__ __launch_bounds__(BLOCK_THREADS) void k_ragged2(const RaggedSub* __restrict__ subs, int nsubs,
                                                           int total_tiles, u32* __restrict__ counter, int w,
                                                           u32* __restrict__ scratch, size_t wave_words) {
  extern __shared__ uint4 smem4[];
  typedef Lazy<RB1, LP1, TPI1, 64> E1;
  typedef Lazy<RB2, LP2, TPI2, 128> E2;
  constexpr int AREA = E1::AREA > E2::AREA ? E1::AREA : E2::AREA;
  u32* area = reinterpret_cast<u32*>(smem4) + (threadIdx.x / 64) * AREA;
  u32* table = scratch + ((size_t)blockIdx.x * (BLOCK_THREADS / 64) + threadIdx.x / 64) * wave_words;
  int local = 0;
  for (int s; (s = ragged_next(counter, total_tiles, subs, nsubs, &local)) >= 0;) {
    if (subs[s].width == 64)
      ragged_tile<RB1, LP1, TPI1, 64>(subs + s, local, area, table, w);
    else
      ragged_tile<RB2, LP2, TPI2, 128>(subs + s, local, area, table, w);
  }
}

}  // namespace chx
