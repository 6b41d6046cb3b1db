// Host-only part of the C ABI (include/chx.h): error reporting and the entry
// points that need no device -- sliding-window exponent schedules, Montgomery
// constants and the consistency check of a Paillier key given with its
// factors. Plain C++ over host_bn.hpp: chx.hip includes it, and the sanitizer
// test (tests/native/) builds it with g++ alone.
#pragma once

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/chx.h"
#include "host_bn.hpp"

// message of the calling thread's last failure (chx_last_error)
static thread_local std::string g_last_error;

namespace {

struct ChxError : std::runtime_error {
  int code;
  ChxError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw ChxError(code, buf);
}

// Runs f, translating exceptions into the ABI's error codes and message.
template <class F>
int guarded(F&& f) {
  try {
    f();
    return CHX_OK;
  } catch (const ChxError& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return CHX_EINVAL;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return CHX_ENOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return CHX_EDEVICE;
  }
}

// Host check of a Paillier key given with its factors (host::check_paillier_factors).
void check_factors(const uint32_t* n, int n_limbs, const uint32_t* phi, const uint32_t* inv, const uint32_t* p,
                   const uint32_t* q, int half_limbs, const uint32_t* hp, const uint32_t* hq, const uint32_t* qinv) {
  if (!n || !phi || !inv || !p || !q || !hp || !hq || !qinv || n_limbs <= 0 || n_limbs > 64 || half_limbs <= 0 ||
      half_limbs > 64)
    fail(CHX_EINVAL, "bad key arguments");
  auto v = [](const uint32_t* x, int l) { return chx::host::bn_from(x, l, l); };
  chx::host::check_paillier_factors(v(n, n_limbs), v(phi, n_limbs), v(inv, n_limbs), v(p, half_limbs),
                                    v(q, half_limbs), v(hp, half_limbs), v(hq, half_limbs), v(qinv, half_limbs));
}

}  // namespace

extern "C" {

int chx_last_error(char* buf, size_t size) {
  if (buf && size) {
    size_t n = std::min(size - 1, g_last_error.size());
    memcpy(buf, g_last_error.data(), n);
    buf[n] = 0;
  }
  return (int)g_last_error.size();
}

int chx_host_schedule(const uint32_t* e, int e_limbs, int window, uint32_t* out, int cap) {
  int n = -1;
  int rc = guarded([&] {
    if (!e || e_limbs <= 0 || window < 1 || window > 7 || cap < 0 || (cap > 0 && !out))
      fail(CHX_EINVAL, "bad argument");
    std::vector<uint32_t> s = chx::host::sliding_schedule(chx::host::bn_from(e, e_limbs, e_limbs), window);
    for (int i = 0; i < (int)s.size() && i < cap; ++i) out[i] = s[i];
    n = (int)s.size();
  });
  return rc == CHX_OK ? n : -rc;
}

int chx_host_paillier_check(const uint32_t* n, int n_limbs, const uint32_t* phi, const uint32_t* inv,
                            const uint32_t* p, const uint32_t* q, int half_limbs, const uint32_t* hp,
                            const uint32_t* hq, const uint32_t* qinv) {
  return guarded([&] { check_factors(n, n_limbs, phi, inv, p, q, half_limbs, hp, hq, qinv); });
}

int chx_host_mont_consts(const uint32_t* n, int limbs, uint32_t* n0inv, uint32_t* r2) {
  return guarded([&] {
    if (!n || !n0inv || !r2 || limbs <= 0) fail(CHX_EINVAL, "bad argument");
    chx::host::MontConsts c = chx::host::mont_consts(chx::host::bn_from(n, limbs, limbs));
    *n0inv = c.n0inv;
    for (int i = 0; i < limbs; ++i) r2[i] = c.R2[i];
  });
}

}  // extern "C"
