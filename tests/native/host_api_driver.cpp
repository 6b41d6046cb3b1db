// Line-oriented driver over the host-only C ABI and the host bignum code
// (cuhepy_amd/csrc/host_api.hpp, host_bn.hpp), built with g++ and the
// address and undefined-behaviour sanitizers by tests/test_host_sanitizers.py.
// One command per input line, numbers in hex; one result line per command:
//   S e w                      chx_host_schedule -> "count word..."
//   M n limbs                  chx_host_mont_consts -> "rc n0inv r2"
//   C n phi inv p q h hp hq qi chx_host_paillier_check (h = half limbs) -> "rc message"
//   P b e n                    host::bn_powmod (Montgomery, CIOS) -> "b^e mod n"
//   X a b                      host::bn_mul -> "a*b"
//   D a n                      host::bn_mod -> "a mod n"
//   Q n                        host::bn_probable_prime -> "0|1"
//   E                          error paths of the entry points (null and bad sizes) -> "rc..."
#include <cstdint>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../cuhepy_amd/csrc/host_api.hpp"

using chx::host::BN;

static std::vector<uint32_t> limbs_of(const std::string& hex, int width = 0) {
  std::vector<uint32_t> v;
  for (int end = (int)hex.size(); end > 0; end -= 8) {
    int beg = end - 8 < 0 ? 0 : end - 8;
    v.push_back((uint32_t)std::stoul(hex.substr(beg, end - beg), nullptr, 16));
  }
  if (v.empty()) v.push_back(0);
  while (width > 0 && (int)v.size() < width) v.push_back(0);
  return v;
}

static std::string hex_of(const uint32_t* p, size_t n) {
  std::string s;
  char buf[16];
  bool lead = true;
  for (size_t i = n; i-- > 0;) {
    if (lead) {
      if (p[i] == 0 && i > 0) continue;
      snprintf(buf, sizeof buf, "%x", p[i]);
      lead = false;
    } else {
      snprintf(buf, sizeof buf, "%08x", p[i]);
    }
    s += buf;
  }
  return s.empty() ? "0" : s;
}

static std::string hex_of(const BN& a) { return hex_of(a.data(), a.size()); }

static BN bn(const std::string& hex, int width) {
  std::vector<uint32_t> v = limbs_of(hex, width);
  return chx::host::bn_from(v.data(), (int)v.size(), width);
}

static int width_of(std::initializer_list<const std::string*> xs) {
  size_t w = 1;
  for (auto* x : xs) w = std::max(w, limbs_of(*x).size());
  return (int)w;
}

int main() {
  std::string line;
  while (std::getline(std::cin, line)) {
    std::istringstream in(line);
    std::string op;
    if (!(in >> op)) continue;
    if (op == "S") {
      std::string e;
      int w;
      in >> e >> w;
      std::vector<uint32_t> ev = limbs_of(e);
      int n = chx_host_schedule(ev.data(), (int)ev.size(), w, nullptr, 0);  // size query
      std::vector<uint32_t> out(n > 0 ? n : 1);
      int m = chx_host_schedule(ev.data(), (int)ev.size(), w, out.data(), n);
      std::cout << m;
      for (int i = 0; i < m; ++i) std::cout << ' ' << std::hex << out[i] << std::dec;
      std::cout << '\n';
    } else if (op == "M") {
      std::string n;
      int limbs;
      in >> n >> limbs;
      std::vector<uint32_t> nv = limbs_of(n, limbs);
      nv.resize(limbs);
      uint32_t n0 = 0;
      std::vector<uint32_t> r2(limbs);
      int rc = chx_host_mont_consts(nv.data(), limbs, &n0, r2.data());
      std::cout << rc << ' ' << std::hex << n0 << ' ' << hex_of(r2.data(), r2.size()) << std::dec << '\n';
    } else if (op == "C") {
      std::string n, phi, inv, p, q, hp, hq, qi;
      int h;
      in >> n >> phi >> inv >> p >> q >> h >> hp >> hq >> qi;
      const int w = width_of({&n, &phi, &inv});
      std::vector<uint32_t> nv = limbs_of(n, w), fv = limbs_of(phi, w), iv = limbs_of(inv, w);
      std::vector<uint32_t> pv = limbs_of(p, h), qv = limbs_of(q, h), hpv = limbs_of(hp, h), hqv = limbs_of(hq, h),
                            qiv = limbs_of(qi, h);
      for (auto* v : {&pv, &qv, &hpv, &hqv, &qiv}) v->resize(h);
      int rc = chx_host_paillier_check(nv.data(), w, fv.data(), iv.data(), pv.data(), qv.data(), h, hpv.data(),
                                       hqv.data(), qiv.data());
      char msg[256];
      chx_last_error(msg, sizeof msg);
      std::cout << rc << ' ' << (rc ? msg : "ok") << '\n';
    } else if (op == "P") {
      std::string b, e, n;
      in >> b >> e >> n;
      const int w = width_of({&n});
      chx::host::MontConsts c = chx::host::mont_consts(bn(n, w));
      std::vector<uint32_t> ev = limbs_of(e);
      BN r = chx::host::bn_powmod(bn(b, w), chx::host::bn_from(ev.data(), (int)ev.size(), (int)ev.size()), c);
      std::cout << hex_of(r) << '\n';
    } else if (op == "X") {
      std::string a, b;
      in >> a >> b;
      const int w = width_of({&a, &b});
      std::cout << hex_of(chx::host::bn_mul(bn(a, w), bn(b, w))) << '\n';
    } else if (op == "D") {
      std::string a, n;
      in >> a >> n;
      const int w = width_of({&a, &n});
      std::cout << hex_of(chx::host::bn_mod(bn(a, w), bn(n, w))) << '\n';
    } else if (op == "Q") {
      std::string n;
      in >> n;
      std::cout << (chx::host::bn_probable_prime(bn(n, width_of({&n})), 24) ? 1 : 0) << '\n';
    } else if (op == "E") {
      uint32_t one[2] = {1, 0}, n0 = 0, r2[2];
      std::cout << chx_host_schedule(nullptr, 1, 4, nullptr, 0) << ' '     // null exponent
                << chx_host_schedule(one, 1, 9, nullptr, 0) << ' '         // window out of range
                << chx_host_schedule(one, 1, 4, nullptr, 4) << ' '         // null output with room
                << chx_host_mont_consts(one, 0, &n0, r2) << ' '            // no limbs
                << chx_host_mont_consts(r2, 2, nullptr, r2) << ' '         // null output
                << chx_host_paillier_check(one, 1, one, one, one, one, 0, one, one, one) << ' '  // no half limbs
                << chx_host_paillier_check(one, 1, one, one, nullptr, one, 1, one, one, one) << '\n';
    } else {
      std::cout << "?\n";
    }
  }
  return 0;
}
