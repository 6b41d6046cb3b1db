// Host-side fixed-width bignum helpers: Montgomery constants, conversions to
// Montgomery form and the sliding-window exponent schedule. Values are
// little-endian vectors of 32-bit limbs. Everything here runs once per key or
// per exponent, never per item, so plain shift-and-subtract arithmetic is used.
#pragma once
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace chx {
namespace host {

typedef std::vector<uint32_t> BN;

inline BN bn_from(const uint32_t* p, int n, int width) {
  if (n > width) {
    for (int i = width; i < n; ++i)
      if (p[i]) throw std::invalid_argument("value does not fit in the target width");
    n = width;
  }
  BN r(width, 0u);
  for (int i = 0; i < n; ++i) r[i] = p[i];
  return r;
}

inline int bn_cmp(const BN& a, const BN& b) {
  for (int i = (int)a.size() - 1; i >= 0; --i) {
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  }
  return 0;
}

inline bool bn_is_zero(const BN& a) {
  for (uint32_t w : a)
    if (w) return false;
  return true;
}

inline int bn_bits(const BN& a) {
  for (int i = (int)a.size() - 1; i >= 0; --i) {
    if (a[i]) return 32 * i + 32 - __builtin_clz(a[i]);
  }
  return 0;
}

// x*x truncated to x.size() limbs (callers pass values whose square fits)
inline BN bn_sqr_lo(const BN& x) {
  const size_t W = x.size();
  BN r(W, 0u);
  for (size_t i = 0; i < W; ++i) {
    uint64_t c = 0;
    for (size_t j = 0; i + j < W; ++j) {
      uint64_t t = (uint64_t)x[i] * x[j] + r[i + j] + c;
      r[i + j] = (uint32_t)t;
      c = t >> 32;
    }
  }
  return r;
}

// a -= b, returns borrow
inline uint32_t bn_sub(BN& a, const BN& b) {
  uint64_t br = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    uint64_t d = (uint64_t)a[i] - b[i] - br;
    a[i] = (uint32_t)d;
    br = (d >> 63) & 1;
  }
  return (uint32_t)br;
}

// a = 2a mod n, for a < n (a, n same width; one extra carry bit tracked)
inline void bn_dbl_mod(BN& a, const BN& n) {
  uint32_t carry = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    uint32_t nc = a[i] >> 31;
    a[i] = (a[i] << 1) | carry;
    carry = nc;
  }
  if (carry || bn_cmp(a, n) >= 0) bn_sub(a, n);
}

// v * 2^k mod n (v < n)
inline BN bn_shl_mod(BN v, int k, const BN& n) {
  for (int i = 0; i < k; ++i) bn_dbl_mod(v, n);
  return v;
}

// x mod n for arbitrary x of the same width (binary long division)
inline BN bn_mod(const BN& x, const BN& n) {
  BN r(n.size(), 0u);
  for (int bit = bn_bits(x) - 1; bit >= 0; --bit) {
    bn_dbl_mod(r, n);
    if ((x[bit / 32] >> (bit % 32)) & 1u) {
      // r = r + 1 mod n; r < n and r even after doubling unless it wrapped, so
      // r + 1 <= n; subtract n when equal
      uint64_t c = 1;
      for (size_t i = 0; i < r.size() && c; ++i) {
        uint64_t s = (uint64_t)r[i] + c;
        r[i] = (uint32_t)s;
        c = s >> 32;
      }
      if (bn_cmp(r, n) >= 0) bn_sub(r, n);
    }
  }
  return r;
}

// -n^-1 mod 2^32 for odd n0 (Newton iteration)
inline uint32_t mont_n0inv(uint32_t n0) {
  uint32_t inv = 1;
  for (int i = 0; i < 5; ++i) inv *= 2u - n0 * inv;
  return (uint32_t)(0u - inv);
}

struct MontConsts {
  int L = 0;
  uint32_t n0inv = 0;
  int bits = 0;
  BN N, R1, R2;
};

inline MontConsts mont_consts(const BN& N) {
  if (N.empty() || !(N[0] & 1u)) throw std::invalid_argument("modulus must be odd");
  MontConsts c;
  c.L = (int)N.size();
  c.N = N;
  c.bits = bn_bits(N);
  c.n0inv = mont_n0inv(N[0]);
  BN one(N.size(), 0u);
  one[0] = 1;
  if (bn_cmp(one, N) >= 0) throw std::invalid_argument("modulus must exceed 1");
  c.R1 = bn_shl_mod(one, 32 * c.L, N);
  c.R2 = bn_shl_mod(c.R1, 32 * c.L, N);
  return c;
}

// v * R mod N (v < N required)
inline BN to_mont_host(const BN& v, const MontConsts& c) {
  if (bn_cmp(v, c.N) >= 0) throw std::invalid_argument("value must be below the modulus");
  return bn_shl_mod(v, 32 * c.L, c.N);
}

// Left-to-right sliding-window schedule (Menezes, van Oorschot, Vanstone,
// Handbook of Applied Cryptography, Alg. 14.85) for exponent e with window w.
// Output: sched[0] = index of the first odd window value ((v-1)/2);
// sched[s >= 1] = (squarings << 16) | (index + 1), index + 1 == 0 = no multiply.
// e == 0 yields an empty schedule (the caller returns 1).
inline std::vector<uint32_t> sliding_schedule(const BN& e, int w) {
  std::vector<uint32_t> s;
  int i = bn_bits(e) - 1;
  auto bit = [&](int j) -> uint32_t { return (e[j / 32] >> (j % 32)) & 1u; };
  uint32_t pending_sq = 0;
  bool first = true;
  while (i >= 0) {
    if (!bit(i)) {
      ++pending_sq;
      --i;
      continue;
    }
    int l = i - w + 1;
    if (l < 0) l = 0;
    while (!bit(l)) ++l;
    uint32_t v = 0;
    for (int j = i; j >= l; --j) v = (v << 1) | bit(j);
    const uint32_t idx = (v - 1) / 2;
    const uint32_t len = (uint32_t)(i - l + 1);
    if (first) {
      s.push_back(idx);
      first = false;
    } else {
      pending_sq += len;
      while (pending_sq > 0xFFFFu) {  // split very long runs of zeros
        s.push_back(0xFFFFu << 16);
        pending_sq -= 0xFFFFu;
      }
      s.push_back((pending_sq << 16) | (idx + 1));
    }
    pending_sq = 0;
    i = l - 1;
  }
  if (!first && pending_sq) {
    while (pending_sq > 0xFFFFu) {
      s.push_back(0xFFFFu << 16);
      pending_sq -= 0xFFFFu;
    }
    s.push_back(pending_sq << 16);
  }
  return s;
}

// Sliding-window width for an e-bit exponent: the w that minimises the
// multiplies beyond the e squarings, 2^(w-1) - 1 to build the odd powers plus
// about e / (w + 1) window multiplies (HAC 14.85). Capped at 6, so a per-item
// table holds at most 32 entries.
inline int default_window(int ebits) {
  if (ebits >= 672) return 6;
  if (ebits >= 240) return 5;
  if (ebits >= 80) return 4;
  if (ebits >= 24) return 3;
  return 2;
}

// Value comparison of two vectors of different lengths (high zero limbs ignored).
inline int bn_cmp_val(const BN& a, const BN& b) {
  const size_t n = a.size() > b.size() ? a.size() : b.size();
  for (size_t k = n; k-- > 0;) {
    const uint32_t x = k < a.size() ? a[k] : 0u, y = k < b.size() ? b[k] : 0u;
    if (x != y) return x < y ? -1 : 1;
  }
  return 0;
}

// Full product a * b (a.size() + b.size() limbs), schoolbook.
inline BN bn_mul(const BN& a, const BN& b) {
  BN r(a.size() + b.size(), 0u);
  for (size_t i = 0; i < a.size(); ++i) {
    uint64_t c = 0;
    for (size_t j = 0; j < b.size(); ++j) {
      const uint64_t t = (uint64_t)a[i] * b[j] + r[i + j] + c;
      r[i + j] = (uint32_t)t;
      c = t >> 32;
    }
    r[i + b.size()] = (uint32_t)c;
  }
  return r;
}

// a * b mod m (any widths; result at m's width)
inline BN bn_mulmod(const BN& a, const BN& b, const BN& m) { return bn_mod(bn_mul(a, b), m); }

// Montgomery product a b R^-1 mod N for a, b < N (CIOS, Koc-Acar-Kaliski 1996),
// canonical result.
inline BN mont_mul(const BN& a, const BN& b, const MontConsts& c) {
  const int L = c.L;
  std::vector<uint32_t> t(L + 2, 0u);
  for (int i = 0; i < L; ++i) {
    uint64_t C = 0;
    for (int j = 0; j < L; ++j) {
      const uint64_t s = (uint64_t)a[j] * b[i] + t[j] + C;
      t[j] = (uint32_t)s;
      C = s >> 32;
    }
    uint64_t s = (uint64_t)t[L] + C;
    t[L] = (uint32_t)s;
    t[L + 1] = (uint32_t)(s >> 32);
    const uint32_t m = t[0] * c.n0inv;
    s = (uint64_t)m * c.N[0] + t[0];
    C = s >> 32;
    for (int j = 1; j < L; ++j) {
      s = (uint64_t)m * c.N[j] + t[j] + C;
      t[j - 1] = (uint32_t)s;
      C = s >> 32;
    }
    s = (uint64_t)t[L] + C;
    t[L - 1] = (uint32_t)s;
    t[L] = t[L + 1] + (uint32_t)(s >> 32);
    t[L + 1] = 0;
  }
  BN r(t.begin(), t.begin() + L);
  if (t[L] || bn_cmp(r, c.N) >= 0) bn_sub(r, c.N);
  return r;
}

// b^e mod N (N odd; b of any width), left-to-right binary.
inline BN bn_powmod(const BN& b, const BN& e, const MontConsts& c) {
  const BN x = mont_mul(bn_mod(b, c.N), c.R2, c);
  BN acc = c.R1;
  for (int i = bn_bits(e) - 1; i >= 0; --i) {
    acc = mont_mul(acc, acc, c);
    if ((e[i / 32] >> (i % 32)) & 1u) acc = mont_mul(acc, x, c);
  }
  BN one(c.L, 0u);
  one[0] = 1;
  return mont_mul(acc, one, c);
}

// Miller-Rabin with the first `rounds` primes as bases after trial division
// (HAC 4.24). An integrity check of caller-supplied factors, not a prime
// search: fixed bases catch every accidental composite.
inline bool bn_probable_prime(const BN& n, int rounds) {
  static const uint32_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53};
  const int nbits = bn_bits(n);
  if (nbits <= 32) {  // exact trial division for one-limb values
    const uint32_t v = n[0];
    if (v < 2) return false;
    for (uint32_t d = 2; (uint64_t)d * d <= v; ++d)
      if (v % d == 0) return false;
    return true;
  }
  auto mod_small = [&](uint32_t d) {
    uint64_t r = 0;
    for (size_t k = n.size(); k-- > 0;) r = ((r << 32) | n[k]) % d;
    return (uint32_t)r;
  };
  for (uint32_t d : small)
    if (mod_small(d) == 0) return false;
  BN w(n.begin(), n.begin() + (nbits + 31) / 32);
  const MontConsts c = mont_consts(w);
  BN one(w.size(), 0u), nm1 = w;
  one[0] = 1;
  bn_sub(nm1, one);
  int s = 0;
  while (!((nm1[s / 32] >> (s % 32)) & 1u)) ++s;
  BN d(w.size(), 0u);  // (n - 1) >> s
  for (int bit = s; bit < nbits; ++bit)
    if ((nm1[bit / 32] >> (bit % 32)) & 1u) d[(bit - s) / 32] |= 1u << ((bit - s) % 32);
  for (int k = 0; k < rounds && k < 16; ++k) {
    BN a(w.size(), 0u);
    a[0] = small[k];
    BN x = bn_powmod(a, d, c);
    if (bn_cmp(x, one) == 0 || bn_cmp(x, nm1) == 0) continue;
    bool witness = true;
    for (int i = 1; i < s && witness; ++i) {
      x = bn_mulmod(x, x, w);
      if (bn_cmp(x, nm1) == 0) witness = false;
    }
    if (witness) return false;
  }
  return true;
}

// Consistency of a Paillier secret key given with its factors (g = n + 1):
// p, q distinct odd primes with p q = n; phi inv = 1 (mod n); lcm(p-1, q-1)
// divides phi (so c^phi of every unit c = (1+n)^m r^n is (1+n)^(m phi));
// h_p (-q) = 1 (mod p), h_q (-p) = 1 (mod q) (h_p = L_p(g^(p-1) mod p^2)^-1,
// as g^(p-1) = 1 + (p-1) n mod p^2); q_inv q = 1 (mod p). Under these the
// textbook CRT decryption equals L(c^phi mod n^2) inv mod n on every unit.
// Throws std::invalid_argument naming the first failed condition.
inline void check_paillier_factors(const BN& n, const BN& phi, const BN& inv, const BN& p, const BN& q,
                                   const BN& hp, const BN& hq, const BN& qinv) {
  auto bad = [](const char* what) { throw std::invalid_argument(what); };
  BN one(1, 1u);
  if (bn_bits(p) < 2 || bn_bits(q) < 2 || !(p[0] & 1u) || !(q[0] & 1u)) bad("p and q must be odd and above 2");
  if (bn_cmp_val(p, q) == 0) bad("p and q must differ");
  if (bn_cmp_val(bn_mul(p, q), n) != 0) bad("p * q != n");
  if (!bn_probable_prime(p, 12) || !bn_probable_prime(q, 12)) bad("p and q must be prime");
  if (bn_cmp_val(bn_mulmod(phi, inv, n), one) != 0) bad("phi * inv != 1 (mod n)");
  BN pm1 = p, qm1 = q;  // p and q are odd: clearing bit 0 subtracts 1
  pm1[0] &= ~1u;
  qm1[0] &= ~1u;
  if (!bn_is_zero(bn_mod(phi, pm1)) || !bn_is_zero(bn_mod(phi, qm1))) bad("lcm(p-1, q-1) must divide phi");
  auto neg_mod = [](const BN& x, const BN& m) {  // (-x) mod m
    BN r = bn_mod(x, m);
    if (bn_is_zero(r)) return r;
    BN t = m;
    bn_sub(t, r);
    return t;
  };
  if (bn_cmp_val(bn_mulmod(hp, neg_mod(q, p), p), one) != 0) bad("h_p * (-q) != 1 (mod p)");
  if (bn_cmp_val(bn_mulmod(hq, neg_mod(p, q), q), one) != 0) bad("h_q * (-p) != 1 (mod q)");
  if (bn_cmp_val(bn_mulmod(qinv, q, p), one) != 0) bad("q_inv * q != 1 (mod p)");
}

}  // namespace host
}  // namespace chx
