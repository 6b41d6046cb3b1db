// Multi-precision Montgomery arithmetic for CDNA4 (gfx950), one bignum per
// lane group.
//
// A bignum of L 32-bit limbs is held by a group of TPI consecutive lanes of a
// wave64, K = L / TPI little-endian limbs per lane in VGPRs: lane rank r owns
// limbs [r*K, r*K + K). G = 64 / TPI groups share a wave. TPI = 1 is "one
// bignum per lane" (no cross-lane traffic); larger TPI trades cross-lane work
// for fewer registers and more lanes per bignum (small batches, wide moduli).
//
// Montgomery product (Montgomery 1985; CIOS ordering of Koc, Acar and Kaliski
// 1996): t = a*b*R^-1 mod N with R = 2^(32L). Operand a is staged in LDS and
// read one limb per row (a group-uniform address, so it is a broadcast read);
// operand b, the modulus N and the accumulator t stay in registers. Each row
// adds a_i*b + m_i*N to t and shifts t down one limb; the shift crosses lanes
// with one DPP row_shl (TPI <= 16) or ds_bpermute (wider groups), and m_i is
// broadcast from group rank 0 (DPP quad_perm for TPI 2/4, ds_bpermute for
// 8..32, readfirstlane for 64). Every lane keeps a small overflow word for its
// top limb; the overflows are folded in after the last row with one
// generate/propagate ballot (a carry-lookahead across the lanes of the group),
// followed by the conditional subtraction of N, so every product is returned
// canonical in [0, N).
//
// Limb products are (u64)x*y + z, which the compiler lowers to v_mad_u64_u32.
// There is no MFMA: this is exact integer arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chx {

typedef uint32_t u32;
typedef uint64_t u64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// ---------------------------------------------------------------------------
// Cross-lane primitives inside a group of TPI lanes
// ---------------------------------------------------------------------------
template <int TPI> struct Grp {
  static constexpr int G = 64 / TPI;
  __device__ __forceinline__ static int rank() { return lane_id() & (TPI - 1); }
  __device__ __forceinline__ static int gid() { return lane_id() / TPI; }

  // value held by rank 0 of the caller's group
  __device__ __forceinline__ static u32 bcast0(u32 v) {
    if constexpr (TPI == 1) {
      return v;
    } else if constexpr (TPI == 2) {
      // every lane has a source in its quad: no "old" value to initialise
      return (u32)__builtin_amdgcn_mov_dpp((int)v, 0xA0, 0xF, 0xF, false);  // quad_perm [0,0,2,2]
    } else if constexpr (TPI == 4) {
      // quad_perm [0,0,0,0]; every source lane exists, so bound_ctrl and the
      // old value 0 never take effect (they let the move fold into a VOP2 user)
      return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x00, 0xF, 0xF, true);
    } else if constexpr (TPI == 8 || TPI == 16) {
      // quad broadcast, then copy quads up the row: row_shr:4 into banks 1 and 3
      // (lanes 4-7, 12-15 of each row of 16); for 16 lanes also row_shr:8 into
      // banks 2 and 3. DPP only, no LDS round trip.
      int x = __builtin_amdgcn_mov_dpp((int)v, 0x00, 0xF, 0xF, false);  // quad_perm [0,0,0,0]
      x = __builtin_amdgcn_update_dpp(x, x, 0x114, 0xF, 0xA, false);     // row_shr:4, banks 1, 3
      if constexpr (TPI == 16) x = __builtin_amdgcn_update_dpp(x, x, 0x118, 0xF, 0xC, false);  // row_shr:8
      return (u32)x;
    } else if constexpr (TPI == 64) {
      return (u32)__builtin_amdgcn_readfirstlane((int)v);
    } else {
      return (u32)__builtin_amdgcn_ds_bpermute((lane_id() & ~(TPI - 1)) << 2, (int)v);
    }
  }

  // bcast0(v) & mask. For 8 and 16 lanes the AND rides on the quad broadcast
  // (v_and_b32_dpp: every source lane of quad_perm [0,0,0,0] exists, so
  // bound_ctrl and the old value never take effect) and the masked quads are
  // then copied up the row: two VALU instructions instead of three.
  __device__ __forceinline__ static u32 bcast0_and(u32 v, u32 mask) {
    if constexpr (TPI == 8 || TPI == 16) {
      int x = __builtin_amdgcn_update_dpp(0, (int)v, 0x00, 0xF, 0xF, true) & (int)mask;
      x = __builtin_amdgcn_update_dpp(x, x, 0x114, 0xF, 0xA, false);  // row_shr:4, banks 1, 3
      if constexpr (TPI == 16) x = __builtin_amdgcn_update_dpp(x, x, 0x118, 0xF, 0xC, false);  // row_shr:8
      return (u32)x;
    } else {
      return bcast0(v) & mask;
    }
  }

  // lane r receives v from rank r+1; the top rank receives 0
  __device__ __forceinline__ static u32 shift_down(u32 v) {
    if constexpr (TPI == 1) {
      return 0u;
    } else if constexpr (TPI <= 16) {
      u32 x = (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x101, 0xF, 0xF, true);  // row_shl:1
      if constexpr (TPI < 16) x = (rank() == TPI - 1) ? 0u : x;
      return x;
    } else {
      u32 x = (u32)__builtin_amdgcn_ds_bpermute((lane_id() + 1) << 2, (int)v);
      return (rank() == TPI - 1) ? 0u : x;
    }
  }

  // lane r receives v from lane r+1 of the wave (the top rank from the next
  // group's rank 0, the wave's last lane 0). Equals shift_down when every
  // group's rank 0 passes 0, without the select that zeroes the top rank.
  __device__ __forceinline__ static u32 shift_down_z0(u32 v) {
    if constexpr (TPI == 1) {
      return 0u;
    } else if constexpr (TPI <= 16) {
      // row_shl:1 within rows of 16 lanes; bound_ctrl gives lane 15 of a row 0
      // (a group never spans rows, so that lane is a top rank)
      return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x101, 0xF, 0xF, true);
    } else {
      return (u32)__builtin_amdgcn_ds_bpermute(((lane_id() + 1) & 63) << 2, (int)v);
    }
  }

  // lane r receives v from rank r-1; rank 0 receives 0
  __device__ __forceinline__ static u32 shift_up(u32 v) {
    if constexpr (TPI == 1) {
      return 0u;
    } else if constexpr (TPI <= 16) {
      u32 x = (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);  // row_shr:1
      if constexpr (TPI < 16) x = (rank() == 0) ? 0u : x;
      return x;
    } else {
      u32 x = (u32)__builtin_amdgcn_ds_bpermute((lane_id() - 1) << 2, (int)v);
      return (rank() == 0) ? 0u : x;
    }
  }

  // value held by the top rank of the caller's group
  __device__ __forceinline__ static u32 bcast_last(u32 v) {
    if constexpr (TPI == 1) {
      return v;
    } else if constexpr (TPI == 2) {
      return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0xF5, 0xF, 0xF, false);  // quad_perm [1,1,3,3]
    } else if constexpr (TPI == 4) {
      return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0xFF, 0xF, 0xF, false);  // quad_perm [3,3,3,3]
    } else {
      return (u32)__builtin_amdgcn_ds_bpermute((lane_id() | (TPI - 1)) << 2, (int)v);
    }
  }

  // bits of a wave ballot that belong to the caller's group (bit r = rank r)
  __device__ __forceinline__ static u64 group_bits(u64 wave_mask) {
    if constexpr (TPI == 64) {
      return wave_mask;
    } else {
      return (wave_mask >> (lane_id() & ~(TPI - 1))) & ((1ull << TPI) - 1ull);
    }
  }

  // Carry-lookahead across the group: gen = this lane produced a carry out of
  // its top limb, prop = a carry entering this lane would leave it. Returns the
  // carry into this lane; *cout receives the carry out of the whole group.
  __device__ __forceinline__ static u32 lookahead(bool gen, bool prop, u32* cout) {
    const u64 g = group_bits(__ballot(gen));
    const u64 p = group_bits(__ballot(prop));
    const u64 gs = g << 1;
    const u64 s = gs + p;
    const u64 cin_mask = s ^ p;
    if constexpr (TPI == 64) {
      *cout = (u32)(((g >> 63) & 1ull) | (s < p ? 1ull : 0ull));
    } else {
      *cout = (u32)((s >> TPI) & 1ull);
    }
    return (u32)((cin_mask >> rank()) & 1ull);
  }
};

// ---------------------------------------------------------------------------
// Multi-limb helpers on a lane-distributed bignum (K limbs in this lane)
// ---------------------------------------------------------------------------

// X += Y (lane-distributed). Returns the carry out of the group (uniform).
template <int K, int TPI>
__device__ __forceinline__ u32 add_dist(u32 (&X)[K], const u32 (&Y)[K]) {
  u32 c = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    u64 s = (u64)X[k] + Y[k] + c;
    X[k] = (u32)s;
    c = (u32)(s >> 32);
  }
  if constexpr (TPI == 1) {
    return c;
  } else {
    u32 all = X[0];
#pragma unroll
    for (int k = 1; k < K; ++k) all &= X[k];
    u32 cout;
    u32 cin = Grp<TPI>::lookahead(c != 0, all == 0xFFFFFFFFu, &cout);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      u64 s = (u64)X[k] + cin;
      X[k] = (u32)s;
      cin = (u32)(s >> 32);
    }
    return cout;
  }
}

// D = X - Y (lane-distributed). Returns the borrow out of the group (uniform).
template <int K, int TPI>
__device__ __forceinline__ u32 sub_dist(u32 (&D)[K], const u32 (&X)[K], const u32 (&Y)[K]) {
  u32 b = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    u64 s = (u64)X[k] - Y[k] - b;
    D[k] = (u32)s;
    b = (u32)(s >> 32) & 1u;
  }
  if constexpr (TPI == 1) {
    return b;
  } else {
    u32 any = D[0];
#pragma unroll
    for (int k = 1; k < K; ++k) any |= D[k];
    u32 bout;
    u32 bin = Grp<TPI>::lookahead(b != 0, any == 0u, &bout);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      u64 s = (u64)D[k] - bin;
      D[k] = (u32)s;
      bin = (u32)(s >> 32) & 1u;
    }
    return bout;
  }
}

// Add a small value v (same for the group) at limb 0 of the group's bignum.
template <int K, int TPI>
__device__ __forceinline__ u32 add_small(u32 (&X)[K], u32 v) {
  u32 Y[K];
#pragma unroll
  for (int k = 0; k < K; ++k) Y[k] = 0;
  Y[0] = (Grp<TPI>::rank() == 0) ? v : 0u;
  return add_dist<K, TPI>(X, Y);
}

template <int K, int TPI>
__device__ __forceinline__ u32 sub_small(u32 (&X)[K], u32 v) {
  u32 Y[K];
#pragma unroll
  for (int k = 0; k < K; ++k) Y[k] = 0;
  Y[0] = (Grp<TPI>::rank() == 0) ? v : 0u;
  return sub_dist<K, TPI>(X, X, Y);
}

// If X >= N (taking the extra top bit `hi` into account), X -= N.
template <int K, int TPI>
__device__ __forceinline__ void cond_sub(u32 (&X)[K], u32 hi, const u32 (&N)[K]) {
  u32 D[K];
  u32 borrow = sub_dist<K, TPI>(D, X, N);
  const bool take = (hi != 0) || (borrow == 0);
#pragma unroll
  for (int k = 0; k < K; ++k) X[k] = take ? D[k] : X[k];
}

// Modular addition X = (X + Y) mod N for X, Y in [0, N).
template <int K, int TPI>
__device__ __forceinline__ void add_mod(u32 (&X)[K], const u32 (&Y)[K], const u32 (&N)[K]) {
  u32 c = add_dist<K, TPI>(X, Y);
  cond_sub<K, TPI>(X, c, N);
}

// Modular subtraction X = (X - Y) mod N for X, Y in [0, N).
template <int K, int TPI>
__device__ __forceinline__ void sub_mod(u32 (&X)[K], const u32 (&Y)[K], const u32 (&N)[K]) {
  u32 D[K];
  const u32 b = sub_dist<K, TPI>(D, X, Y);
  u32 E[K];
#pragma unroll
  for (int k = 0; k < K; ++k) E[k] = D[k];
  (void)add_dist<K, TPI>(E, N);  // computed by every group (no divergence), selected below
#pragma unroll
  for (int k = 0; k < K; ++k) X[k] = b ? E[k] : D[k];
}

// ---------------------------------------------------------------------------
// LDS staging of the streamed operand
// ---------------------------------------------------------------------------
// Layout for one wave: uint4 slots [L/4][G]; limb i of group g lives in
// slot (i/4)*G + g, component i%4. A row reads one uint4 per 4 rows, at a
// group-uniform address: broadcast within the group, consecutive slots across
// groups (conflict-free).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int L, int TPI>
struct Mont {
  static constexpr int K = L / TPI;
  static constexpr int G = 64 / TPI;
  static_assert(L % TPI == 0, "L must be a multiple of TPI");
  static_assert(K % 4 == 0, "K must be a multiple of 4");
  typedef Grp<TPI> grp;

  uint4* lds;   // this wave's staging area: (L/4)*G uint4
  u32 n0inv;    // -N^-1 mod 2^32
  u32 N[K];     // modulus, lane slice

  // write the lane's K limbs of x into the staging area
  __device__ __forceinline__ void stage(const u32 (&x)[K]) {
    const int g = grp::gid();
    const int r = grp::rank();
#pragma unroll
    for (int k4 = 0; k4 < K / 4; ++k4) {
      uint4 v = make_uint4(x[4 * k4], x[4 * k4 + 1], x[4 * k4 + 2], x[4 * k4 + 3]);
      lds[(r * (K / 4) + k4) * G + g] = v;
    }
  }

  // read back limbs (after a wave sync) into a lane-distributed array
  __device__ __forceinline__ void unstage(u32 (&x)[K]) {
    const int g = grp::gid();
    const int r = grp::rank();
#pragma unroll
    for (int k4 = 0; k4 < K / 4; ++k4) {
      uint4 v = lds[(r * (K / 4) + k4) * G + g];
      x[4 * k4] = v.x; x[4 * k4 + 1] = v.y; x[4 * k4 + 2] = v.z; x[4 * k4 + 3] = v.w;
    }
  }

  // t = a*b*R^-1 mod N where a has been staged; b in registers. Canonical out.
  __device__ __forceinline__ void mul_staged(u32 (&t)[K], const u32 (&b)[K]) {
    const int g = grp::gid();
    u32 top = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) t[k] = 0;
    for (int i4 = 0; i4 < L / 4; ++i4) {
      const uint4 a4 = lds[i4 * G + g];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const u32 ai = (s == 0) ? a4.x : (s == 1) ? a4.y : (s == 2) ? a4.z : a4.w;
        // first product of the row decides the Montgomery digit m
        u64 p = (u64)ai * b[0] + t[0];
        const u32 m = grp::bcast0((u32)p * n0inv);
        u64 q = (u64)m * N[0] + (u32)p;
        u64 c1 = p >> 32;
        u64 c2 = q >> 32;
        const u32 low = (u32)q;  // zero in rank 0; goes to rank-1 otherwise
#pragma unroll
        for (int k = 1; k < K; ++k) {
          p = (u64)ai * b[k] + t[k] + c1;
          q = (u64)m * N[k] + (u32)p + c2;
          c1 = p >> 32;
          c2 = q >> 32;
          t[k - 1] = (u32)q;
        }
        u64 s_top = (u64)top + c1 + c2 + grp::shift_down(low);
        t[K - 1] = (u32)s_top;
        top = (u32)(s_top >> 32);
      }
    }
    finish(t, top);
  }

  // fold the per-lane overflow words and reduce to [0, N)
  __device__ __forceinline__ void finish(u32 (&t)[K], u32 top) {
    u32 hi;
    if constexpr (TPI == 1) {
      hi = top;
    } else {
      const u32 cin = grp::shift_up(top);
      u32 Y[K];
#pragma unroll
      for (int k = 0; k < K; ++k) Y[k] = 0;
      Y[0] = cin;
      const u32 cout = add_dist<K, TPI>(t, Y);
      // overflow beyond the top limb of the group = top of the last lane + carry out
      hi = grp::bcast_last(top) + cout;
    }
    cond_sub<K, TPI>(t, hi, N);
  }

  __device__ __forceinline__ void mul(u32 (&t)[K], const u32 (&a)[K], const u32 (&b)[K]) {
    stage(a);
    wave_lds_sync();
    mul_staged(t, b);
    wave_lds_sync();
  }

  __device__ __forceinline__ void sqr(u32 (&x)[K]) {
    u32 t[K];
    stage(x);
    wave_lds_sync();
    mul_staged(t, x);
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] = t[k];
  }

  // x = x * y (mod N), both in Montgomery form
  __device__ __forceinline__ void mul_into(u32 (&x)[K], const u32 (&y)[K]) {
    u32 t[K];
    mul(t, y, x);
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] = t[k];
  }

  // Exact quotient X / N for X = q*N with q < R (right-to-left Hensel
  // division): the Montgomery digits m_i of REDC(X) satisfy
  // M = sum m_i 2^(32i) = -X*N^-1 mod R, so q = R - M (mod R).
  __device__ __forceinline__ void exact_div(u32 (&q)[K], const u32 (&X)[K]) {
    const int g = grp::gid();
    u32* digits = reinterpret_cast<u32*>(lds);
    u32 t[K];
    u32 top = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) t[k] = X[k];
    for (int i = 0; i < L; ++i) {
      const u32 m = grp::bcast0(t[0] * n0inv);
      if (grp::rank() == 0) digits[((i >> 2) * G + g) * 4 + (i & 3)] = m;
      u64 q0 = (u64)m * N[0] + t[0];
      u64 c = q0 >> 32;
      const u32 low = (u32)q0;
#pragma unroll
      for (int k = 1; k < K; ++k) {
        u64 qq = (u64)m * N[k] + t[k] + c;
        c = qq >> 32;
        t[k - 1] = (u32)qq;
      }
      u64 s_top = (u64)top + c + grp::shift_down(low);
      t[K - 1] = (u32)s_top;
      top = (u32)(s_top >> 32);
    }
    wave_lds_sync();
    unstage(q);
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < K; ++k) q[k] = ~q[k];
    (void)add_small<K, TPI>(q, 1u);
  }
};

}  // namespace chx
