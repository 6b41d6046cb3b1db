// C-ABI implementation of the gfx950 Paillier library (declared in include/chx.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/chx.h"
#include "host_api.hpp"
#include "host_bn.hpp"
#include "kernels.hpp"
#include "lazy.hpp"

using namespace chx;
using chx::host::BN;

// ---------------------------------------------------------------------------
// errors: ChxError, fail() and guarded() live in host_api.hpp
// ---------------------------------------------------------------------------
namespace {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(CHX_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// Kernels index items with 32-bit ints and limb rows as row * stride + item:
// batches are bounded so every such index fits.
constexpr size_t kMaxStride = (size_t)1 << 28;

// stride of a host-buffer call's device copy (count rounded up to 64)
size_t batch_stride(size_t count) {
  if (count > kMaxStride) fail(CHX_EINVAL, "batch of %zu items exceeds the maximum of %zu", count, kMaxStride);
  return round_up(count, 64);
}

// a caller's interleaved device batch: stride >= count rounded up to 64, both bounded
void check_stride(size_t count, size_t stride) {
  if (count > kMaxStride || stride > kMaxStride)
    fail(CHX_EINVAL, "batch of %zu items / stride %zu exceeds the maximum of %zu", count, stride, kMaxStride);
  if (stride < round_up(count, 64)) fail(CHX_EINVAL, "stride %zu below count rounded to 64", stride);
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t need) {
    if (need <= bytes) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t sz = std::max<size_t>(need, 1 << 20);
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) fail(CHX_ENOMEM, "hipMalloc(%zu): %s", sz, hipGetErrorString(e));
    bytes = sz;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() { return reinterpret_cast<T*>(p); }
};

}  // namespace

// ---------------------------------------------------------------------------
// scratch of the allocation-free entry point (chx_paillier_decrypt_crt_noalloc)
// ---------------------------------------------------------------------------
// One file-scope device array: every GPU the library runs on holds its own copy
// until the process ends. In host code `scratch` names a host-side stand-in, not
// a device address: kernels and copies use the per-device address taken once
// with hipGetSymbolAddress (chx_ctx_create, with the device current). Calls of
// the entry point on one GPU share the array, so calls that overlap on one GPU
// race; nothing in it has to start at zero (each call writes every word it
// reads before reading it).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchBytes / sizeof(uint32_t)];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];             // device address of `scratch` on each GPU
std::string g_scratch_err[kMaxDevices];       // why it could not be taken, if it could not

// With `device` current: take the device address of `scratch` once.
void init_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices || g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return;
  }
  g_scratch_ptr[device] = ptr;
}

uint32_t* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) fail(CHX_EUNSUPPORTED, "device %d has no scratch array", device);
  if (!g_scratch_ptr[device])
    fail(CHX_EDEVICE, "scratch array of device %d unavailable: %s", device,
         g_scratch_err[device].empty() ? "not initialised" : g_scratch_err[device].c_str());
  return static_cast<uint32_t*>(g_scratch_ptr[device]);
}
}  // namespace

enum Slot { S_IN0, S_IN1, S_OUT, S_AOS, S_TAB, S_TMP, S_AUX, S_MID0, S_MID1, S_QUEUE, S_NSLOT };

struct TimedLaunch {
  std::string name;
  hipEvent_t start, stop;
};

struct chx_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  std::mutex mu;
  DevBuf slot[S_NSLOT];
  int tune_tpi[4][3] = {};  // [op][limbs 32/64/128]
  int tune_window[4][3] = {};
  bool timing = false;
  std::vector<TimedLaunch> pending;
  std::map<std::string, std::pair<double, int>> stats;
  uint32_t packed_token = 0;  // per-call value of k_il_digits's flag (never 0)
};

// Lazy-radix constants of one modulus, one block per built lazy configuration
// of its width (the configurations differ in positions per number).
struct LazySet {
  int L32 = 0;
  LazyDev* d[3] = {nullptr, nullptr, nullptr};
};

struct chx_mod {
  chx_ctx* ctx;
  host::MontConsts hc;
  ModDev* d = nullptr;
  LazySet lz;
  uint32_t* d_fixb = nullptr;  // served-multiply constants 2^(RB(i+2)) R' mod N (built on first use)
  uint32_t* d_pwz = nullptr;   // fold table of the pointwise multiply (built on first use)
  int pw_h = -1, pw_nz = 0;    // its first folded position and row count; pw_h -2: not usable
  std::mutex mu;               // guards d_fixb and d_pwz: a modulus may serve several contexts of its device
};

struct ExpSched {
  std::vector<uint32_t> h;
  uint32_t* d = nullptr;
  int window = 0;
  int bits = 0;
};

struct chx_pkey {
  chx_ctx* ctx;
  int W = 0;  // limbs of n^2
  int H = 0;  // W / 2
  bool crt = false;
  host::MontConsts n2, n;
  PaillierDev* d = nullptr;
  LazySet lz_n2, lz_p2, lz_q2;
  BN e_n, e_phi, e_pm1, e_qm1;
  std::map<std::pair<int, int>, ExpSched> scheds;  // (which, window) -> schedule
  BN p2h, q2h;                // p^2, q^2 at H limbs (CRT keys)
  uint32_t* d_crt = nullptr;  // [e_p R | e_q R] mod n^2 (chx_paillier_key_set_idempotents)
};

struct chx_lkey {
  chx_ctx* ctx;
  int W = 0;
  int chunks = 0;
  int noise = 0;
  host::MontConsts n2, n;
  ModDev* d_n2 = nullptr;
  ModDev* d_n = nullptr;
  uint32_t* d_gar = nullptr;   // g_a_inv * R mod n
  uint32_t* d_gtab = nullptr;  // Montgomery-form g table, AoS
  uint32_t* d_ntab = nullptr;  // Montgomery-form noise table, AoS
  // row-group tables (k_lazy_group_table<R>), group[R] for R = 2..5 rows per
  // entry, each built on first use over its first `groups` groups; failed: no
  // memory for it, the next smaller group serves
  struct GroupTab {
    uint32_t* d = nullptr;
    int groups = 0;
    bool failed = false;
  } group[6];
  int table_rows = 0;  // rows the group tables cover, 0 = all (chx_lookup_key_set_table_rows)
  uint32_t* d_n3tab = nullptr;  // noise-triple table (k_lazy_noise_triples), noise^3 entries, with 5-row groups
  bool n3tab_failed = false;
  uint32_t* d_nptab = nullptr;  // noise-pair table (k_lazy_noise_pairs), noise^2 entries
  uint32_t* d_ntab_plain = nullptr;   // the noise table as plain residues (k_lazy_plain)
  uint32_t* d_nptab_plain = nullptr;  // the noise-pair table as plain residues
  LazySet lz_n2;
  BN e_a;
  std::map<int, ExpSched> scheds;
  // half-width decryption (chx_lookup_key_set_crt): c^a mod p^2 and mod q^2
  bool crt = false;
  LazySet lz_p2, lz_q2;
  uint32_t* d_crt = nullptr;  // [e_p R | e_q R] mod n^2, W limbs each
  BN gai;                          // g_a_inv mod n
  LookupCrtDev* d_lcrt = nullptr;  // textbook CRT route (k_lookup_crt_fast), with the factors
};

namespace {

int lidx(int limbs) { return limbs == 32 ? 0 : limbs == 64 ? 1 : limbs == 128 ? 2 : -1; }

// lanes-per-bignum values instantiated for each width
const std::vector<int>& built_tpi(int limbs) {
  static const std::vector<int> t32 = {1, 4};
  static const std::vector<int> t64 = {4, 16};
  static const std::vector<int> t128 = {8, 32};
  static const std::vector<int> none = {};
  return limbs == 32 ? t32 : limbs == 64 ? t64 : limbs == 128 ? t128 : none;
}

// Number of lanes that keeps every SIMD of the chip busy with 2 waves.
// compute units of `device` (cached; 256 if the query fails)
int device_cus(int device) {
  static std::mutex mu;
  static std::map<int, int> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(device);
  if (it != cache.end()) return it->second;
  hipDeviceProp_t prop;
  int v = 256;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0) v = prop.multiProcessorCount;
  cache[device] = v;
  return v;
}

size_t target_lanes(int device) {
  static std::mutex mu;
  static std::map<int, size_t> cache;
  std::lock_guard<std::mutex> lock(mu);  // contexts of different threads share the cache
  auto it = cache.find(device);
  if (it != cache.end()) return it->second;
  hipDeviceProp_t prop;
  size_t v = 256 * 4 * 64 * 2;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    v = (size_t)prop.multiProcessorCount * 4 * 64 * 2;
  cache[device] = v;
  return v;
}

// Default lanes per bignum: the smallest built TPI whose lanes fill the chip,
// else the largest built one (small batches want many lanes per bignum).
int pick_tpi(chx_ctx* ctx, int op, int limbs, size_t count) {
  const int li = lidx(limbs);
  if (li < 0) fail(CHX_EUNSUPPORTED, "width of %d limbs is not built", limbs);
  const auto& opts = built_tpi(limbs);
  const int forced = ctx->tune_tpi[op][li];
  if (forced) {
    if (std::find(opts.begin(), opts.end(), forced) == opts.end())
      fail(CHX_EUNSUPPORTED, "tpi=%d is not built for %d limbs", forced, limbs);
    return forced;
  }
  const size_t want = target_lanes(ctx->device);
  for (int t : opts)
    if (count * (size_t)t >= want) return t;
  return opts.back();
}

int pick_window(chx_ctx* ctx, int op, int limbs, int ebits) {
  const int li = lidx(limbs);
  const int forced = li >= 0 ? ctx->tune_window[op][li] : 0;
  if (forced) return forced;
  return host::default_window(ebits);
}

void record_start(chx_ctx* ctx, const char* name) {
  if (!ctx->timing) return;
  TimedLaunch t;
  t.name = name;
  hip_check(hipEventCreate(&t.start), "hipEventCreate");
  hip_check(hipEventCreate(&t.stop), "hipEventCreate");
  hip_check(hipEventRecord(t.start, ctx->stream), "hipEventRecord");
  ctx->pending.push_back(t);
}

void record_stop(chx_ctx* ctx) {
  if (!ctx->timing || ctx->pending.empty()) return;
  hip_check(hipEventRecord(ctx->pending.back().stop, ctx->stream), "hipEventRecord");
}

void flush_timing(chx_ctx* ctx) {
  for (auto& t : ctx->pending) {
    hip_check(hipEventSynchronize(t.stop), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, t.start, t.stop), "hipEventElapsedTime");
    auto& s = ctx->stats[t.name];
    s.first += ms;
    s.second += 1;
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  ctx->pending.clear();
}

void check_launch(const char* what) { hip_check(hipGetLastError(), what); }

size_t lds_bytes(int L, int tpi) { return (size_t)WAVES_PER_BLOCK * (L / 4) * (64 / tpi) * 16; }

dim3 grid_for(size_t count, int tpi, int rows = 1) {
  const size_t per_block = BLOCK_THREADS / tpi;
  return dim3((unsigned)((count + per_block - 1) / per_block), rows, 1);
}

// ---------------------------------------------------------------------------
// dispatch over (L, TPI)
// ---------------------------------------------------------------------------
#define CHX_DISPATCH(L_, TPI_, CALL)                                                           \
  do {                                                                                         \
    const int l__ = (L_), t__ = (TPI_);                                                        \
    if (l__ == 32 && t__ == 1) { constexpr int L = 32, TPI = 1; CALL; }                        \
    else if (l__ == 32 && t__ == 4) { constexpr int L = 32, TPI = 4; CALL; }                   \
    else if (l__ == 64 && t__ == 4) { constexpr int L = 64, TPI = 4; CALL; }                   \
    else if (l__ == 64 && t__ == 16) { constexpr int L = 64, TPI = 16; CALL; }                 \
    else if (l__ == 128 && t__ == 8) { constexpr int L = 128, TPI = 8; CALL; }                 \
    else if (l__ == 128 && t__ == 32) { constexpr int L = 128, TPI = 32; CALL; }               \
    else fail(CHX_EUNSUPPORTED, "no kernel built for %d limbs with %d lanes per bignum", l__, t__); \
  } while (0)

// CRT halves: H in {32, 64}; lane groups of at least 4 (a p- and a q-half)
#define CHX_DISPATCH_H(H_, TPI_, CALL)                                                         \
  do {                                                                                         \
    const int h__ = (H_), t__ = (TPI_);                                                        \
    if (h__ == 32 && t__ == 4) { constexpr int H = 32, TPI = 4; CALL; }                        \
    else if (h__ == 64 && t__ == 4) { constexpr int H = 64, TPI = 4; CALL; }                   \
    else if (h__ == 64 && t__ == 16) { constexpr int H = 64, TPI = 16; CALL; }                 \
    else fail(CHX_EUNSUPPORTED, "no CRT kernel built for %d-limb halves with %d lanes", h__, t__); \
  } while (0)

// ---------------------------------------------------------------------------
// host <-> device helpers
// ---------------------------------------------------------------------------
void fill_moddev(ModDev& md, const host::MontConsts& c) {
  memset(&md, 0, sizeof md);
  md.L = (u32)c.L;
  md.n0inv = c.n0inv;
  md.bits = (u32)c.bits;
  for (int i = 0; i < c.L; ++i) {
    md.N[i] = c.N[i];
    md.R1[i] = c.R1[i];
    md.R2[i] = c.R2[i];
  }
}

void fill_val(ValDev& v, const BN& x) {
  memset(&v, 0, sizeof v);
  for (size_t i = 0; i < x.size() && i < (size_t)LMAX; ++i) v.v[i] = x[i];
}

template <class T>
T* dev_upload(const T* host, size_t n, hipStream_t s) {
  T* d = nullptr;
  hipError_t e = hipMalloc(&d, sizeof(T) * std::max<size_t>(n, 1));
  if (e != hipSuccess) fail(CHX_ENOMEM, "hipMalloc: %s", hipGetErrorString(e));
  hip_check(hipMemcpyAsync(d, host, sizeof(T) * n, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return d;
}

int pick_width(int bits_needed) {
  if (bits_needed <= 32 * 32) return 32;
  if (bits_needed <= 64 * 32) return 64;
  if (bits_needed <= 128 * 32) return 128;
  fail(CHX_EUNSUPPORTED, "values of %d bits exceed the widest built width (4096 bits)", bits_needed);
}

// validate that every item of an AoS batch is below the modulus
void check_below(const uint32_t* v, size_t count, int limbs, const BN& N, const char* what) {
  for (size_t i = 0; i < count; ++i) {
    const uint32_t* x = v + i * limbs;
    int cmp = 0;
    for (int j = limbs - 1; j >= 0 && cmp == 0; --j) {
      const uint32_t nj = j < (int)N.size() ? N[j] : 0u;
      if (x[j] != nj) cmp = x[j] < nj ? -1 : 1;
    }
    if (cmp >= 0) fail(CHX_EINVAL, "%s[%zu] is not below the modulus", what, i);
  }
}

void aos_to_il(chx_ctx* ctx, const uint32_t* aos, uint32_t* il, size_t count, int L, size_t stride) {
  dim3 g((unsigned)((stride + 31) / 32), (unsigned)((L + 31) / 32));
  hipLaunchKernelGGL(k_aos_to_il, g, dim3(256), 0, ctx->stream, aos, il, (int)count, L, (int)stride);
  check_launch("k_aos_to_il");
}

void il_to_aos(chx_ctx* ctx, const uint32_t* il, uint32_t* aos, size_t count, int L, size_t stride) {
  dim3 g((unsigned)((count + 31) / 32), (unsigned)((L + 31) / 32));
  hipLaunchKernelGGL(k_il_to_aos, g, dim3(256), 0, ctx->stream, il, aos, (int)count, L, (int)stride);
  check_launch("k_il_to_aos");
}

// upload an AoS host batch into an interleaved device slot
uint32_t* upload_il(chx_ctx* ctx, Slot slot, const uint32_t* host, size_t count, int L, size_t stride) {
  const size_t bytes = count * L * sizeof(uint32_t);
  ctx->slot[S_AOS].ensure(bytes);
  ctx->slot[slot].ensure(stride * L * sizeof(uint32_t));
  hip_check(hipMemcpyAsync(ctx->slot[S_AOS].p, host, bytes, hipMemcpyHostToDevice, ctx->stream), "H2D");
  aos_to_il(ctx, ctx->slot[S_AOS].as<uint32_t>(), ctx->slot[slot].as<uint32_t>(), count, L, stride);
  return ctx->slot[slot].as<uint32_t>();
}

void download_il(chx_ctx* ctx, const uint32_t* il, uint32_t* host, size_t count, int L, size_t stride) {
  const size_t bytes = count * L * sizeof(uint32_t);
  ctx->slot[S_AOS].ensure(bytes);
  il_to_aos(ctx, il, ctx->slot[S_AOS].as<uint32_t>(), count, L, stride);
  hip_check(hipMemcpyAsync(host, ctx->slot[S_AOS].p, bytes, hipMemcpyDeviceToHost, ctx->stream), "D2H");
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

// A schedule of the key already on the device, or nullptr (never builds one).
const ExpSched* find_sched_p(chx_pkey* k, int which, int window) {
  auto it = k->scheds.find(std::make_pair(which, window));
  return it == k->scheds.end() ? nullptr : &it->second;
}

const ExpSched& get_sched_p(chx_pkey* k, int which, int window) {
  auto key = std::make_pair(which, window);
  auto it = k->scheds.find(key);
  if (it != k->scheds.end()) return it->second;
  const BN& e = which == 0 ? k->e_n : which == 1 ? k->e_phi : which == 2 ? k->e_pm1 : k->e_qm1;
  ExpSched s;
  s.window = window;
  s.bits = host::bn_bits(e);
  s.h = host::sliding_schedule(e, window);
  if (s.h.empty()) fail(CHX_EINVAL, "zero exponent");
  s.d = dev_upload(s.h.data(), s.h.size(), k->ctx->stream);
  return k->scheds.emplace(key, std::move(s)).first->second;
}

int exp_bits_p(const chx_pkey* k, int which) {
  const BN& e = which == 0 ? k->e_n : which == 1 ? k->e_phi : which == 2 ? k->e_pm1 : k->e_qm1;
  return host::bn_bits(e);
}


// ---------------------------------------------------------------------------
// lazy-radix exponentiation configurations: (radix bits, positions, lanes)
// ---------------------------------------------------------------------------
struct LazyCfg {
  int rb, lp, tpi;
};

const std::vector<LazyCfg>& lazy_cfgs(int L32) {
  // variants whose accumulators fit in VGPRs without spilling (checked with
  // -Rpass-analysis=kernel-resource-usage); index = dispatch slot. Each list
  // runs from the throughput shape to the latency shape for small batches:
  // (28, 40, 8) and (29, 80, 16) give 5 positions per lane (a few more
  // positions than the width needs, so that 8 or 16 lanes divide them), and
  // 4x or 2x the lanes per modulus of the shape before them.
  static const std::vector<LazyCfg> c32 = {{28, 38, 1}, {28, 38, 2}, {28, 40, 8}};
  static const std::vector<LazyCfg> c64 = {{29, 72, 4}, {29, 72, 8}, {29, 80, 16}};
  static const std::vector<LazyCfg> c128 = {{29, 144, 8}, {29, 144, 16}};
  static const std::vector<LazyCfg> none = {};
  return L32 == 32 ? c32 : L32 == 64 ? c64 : L32 == 128 ? c128 : none;
}

// split a value (width L32 limbs) into `lp` positions of `rb` bits
void to_positions(const BN& v, int rb, int lp, uint32_t* out) {
  const uint32_t mask = (1u << rb) - 1u;
  for (int p = 0; p < lp; ++p) {
    const int bit = rb * p, j = bit >> 5, off = bit & 31;
    uint64_t lo = j < (int)v.size() ? v[j] : 0u;
    uint64_t hi = j + 1 < (int)v.size() ? v[j + 1] : 0u;
    out[p] = (uint32_t)(((hi << 32) | lo) >> off) & mask;
  }
}

// Constants of one modulus in the lazy radix (rb, lp), uploaded.
LazyDev* make_lazy_dev(const host::MontConsts& c, int rb, int lp, hipStream_t st) {
  if (rb * lp < 32 * c.L + 2) fail(CHX_EUNSUPPORTED, "lazy configuration too narrow");
  std::unique_ptr<LazyDev> ld(new LazyDev());
  memset(ld.get(), 0, sizeof(LazyDev));
  ld->n0 = c.n0inv & ((1u << rb) - 1u);
  ld->rb = rb;
  ld->lp = lp;
  ld->l32 = c.L;
  to_positions(c.N, rb, lp, ld->N);
  BN one(c.L, 0u);
  one[0] = 1;
  BN r1 = host::bn_shl_mod(one, rb * lp, c.N);  // R' mod N
  BN r2 = host::bn_shl_mod(r1, rb * lp, c.N);   // R'^2 mod N
  BN rw = host::bn_shl_mod(r2, 32 * c.L, c.N);  // 2^(32 L32) R'^2 mod N
  to_positions(r1, rb, lp, ld->R1);
  to_positions(r2, rb, lp, ld->R2);
  to_positions(rw, rb, lp, ld->RW);
  if (rb * lp >= 32 * c.L + rb + 4) {  // Lazy::UNIT: exponentiation modulo N' = kN
    const uint32_t k = ld->n0;            // -N^-1 mod 2^rb, so kN = -1 mod 2^rb
    BN nk(c.L + 1, 0u);
    uint64_t carry = 0;
    for (int i = 0; i < c.L; ++i) {
      const uint64_t t = (uint64_t)c.N[i] * k + carry;
      nk[i] = (uint32_t)t;
      carry = t >> 32;
    }
    nk[c.L] = (uint32_t)carry;
    if (((nk[0] + 1u) & ((1u << rb) - 1u)) != 0u) fail(CHX_EINVAL, "internal: kN is not -1 mod 2^rb");
    to_positions(nk, rb, lp, ld->NK);
  }
  for (int k = 0; k < c.L; ++k) ld->N32[k] = c.N[k];
  return dev_upload(ld.get(), 1, st);
}

LazySet make_lazy_set(const host::MontConsts& c, hipStream_t st) {
  LazySet ls;
  ls.L32 = c.L;
  const auto& cfgs = lazy_cfgs(c.L);
  try {
    for (size_t i = 0; i < cfgs.size() && i < 3; ++i) ls.d[i] = make_lazy_dev(c, cfgs[i].rb, cfgs[i].lp, st);
  } catch (...) {
    for (auto& p : ls.d)
      if (p) (void)hipFree(p), p = nullptr;
    throw;
  }
  return ls;
}

void free_lazy_set(LazySet& ls) {
  for (auto& p : ls.d)
    if (p) (void)hipFree(p), p = nullptr;
}

// pick the lazy configuration for a batch (tuning override, else fill the chip)
// Lookup tables (g table, noise table) are stored once, in the positions of
// configuration 0; only configurations with the same radix and position count
// can read them.
bool lazy_usable(int op, int L32, size_t ci) {
  const auto& cfgs = lazy_cfgs(L32);
  if (op != CHX_OP_LOOKUP) return true;
  return cfgs[ci].rb == cfgs[0].rb && cfgs[ci].lp == cfgs[0].lp;
}

int pick_lazy(chx_ctx* ctx, int L32, size_t count, int op = CHX_OP_MODEXP) {
  const int li = lidx(L32);
  if (li < 0) fail(CHX_EUNSUPPORTED, "width of %d limbs is not built", L32);
  const auto& cfgs = lazy_cfgs(L32);
  const int forced = ctx->tune_tpi[op][li];
  if (forced) {
    for (size_t i = 0; i < cfgs.size(); ++i)
      if (cfgs[i].tpi == forced && lazy_usable(op, L32, i)) return (int)i;
    fail(CHX_EUNSUPPORTED, "tpi=%d is not built for %d-limb op %d", forced, L32, op);
  }
  const size_t want = target_lanes(ctx->device);
  int last = 0;
  for (size_t i = 0; i < cfgs.size(); ++i) {
    if (!lazy_usable(op, L32, i)) continue;
    last = (int)i;
    if (count * (size_t)cfgs[i].tpi >= want) return (int)i;
  }
  return last;
}

template <int RB, int LP, int TPI, int L32>
size_t lazy_lds_bytes() {
  return (size_t)WAVES_PER_BLOCK * Lazy<RB, LP, TPI, L32>::AREA * 4;
}

#define CHX_DISPATCH_LAZY(L32_, IDX_, ...)                                                        \
  do {                                                                                          \
    const int l__ = (L32_), i__ = (IDX_);                                                       \
    if (l__ == 32 && i__ == 0) { constexpr int RB = 28, LP = 38, TPI = 1, L32 = 32; __VA_ARGS__; }      \
    else if (l__ == 32 && i__ == 1) { constexpr int RB = 28, LP = 38, TPI = 2, L32 = 32; __VA_ARGS__; } \
    else if (l__ == 32 && i__ == 2) { constexpr int RB = 28, LP = 40, TPI = 8, L32 = 32; __VA_ARGS__; } \
    else if (l__ == 64 && i__ == 0) { constexpr int RB = 29, LP = 72, TPI = 4, L32 = 64; __VA_ARGS__; }  \
    else if (l__ == 64 && i__ == 1) { constexpr int RB = 29, LP = 72, TPI = 8, L32 = 64; __VA_ARGS__; }  \
    else if (l__ == 64 && i__ == 2) { constexpr int RB = 29, LP = 80, TPI = 16, L32 = 64; __VA_ARGS__; } \
    else if (l__ == 128 && i__ == 0) { constexpr int RB = 29, LP = 144, TPI = 8, L32 = 128; __VA_ARGS__; } \
    else if (l__ == 128 && i__ == 1) { constexpr int RB = 29, LP = 144, TPI = 16, L32 = 128; __VA_ARGS__; } \
    else fail(CHX_EUNSUPPORTED, "no lazy kernel for %d limbs (configuration %d)", l__, i__);   \
  } while (0)

// Window-table words of one job of run_lazy (configuration ci, window w).
size_t lazy_table_words(int L32, int ci, int window, size_t count) {
  return ((size_t)1 << (window - 1)) * lazy_cfgs(L32)[ci].lp * round_up(count, BLOCK_THREADS);
}

// One launch exponentiating one or two batches (CRT halves) of `count` items.
// jobs[j]: input/output interleaved 32-bit buffers, schedule; the modulus
// block of the chosen configuration is taken from sets[j].
// in_hi (optional, per job): upper L32 limbs of double-width inputs.
// tab: the window tables' scratch (njobs * lazy_table_words words), or
// nullptr for the context's slot.
void run_lazy(chx_ctx* ctx, int L32, const LazySet* sets, const ExpSched* const* scheds, const uint32_t* const* in,
              uint32_t* const* out, int njobs, size_t count, size_t stride,
              const uint32_t* const* in_hi = nullptr, uint32_t* tab = nullptr) {
  const int ci = pick_lazy(ctx, L32, count * njobs);
  const LazyCfg cfg = lazy_cfgs(L32)[ci];
  // a double-width input enters the exponentiation below 4N: R' > 16N keeps
  // every product of it below 2N
  if (in_hi && cfg.rb * cfg.lp < 32 * L32 + 4) fail(CHX_EUNSUPPORTED, "lazy configuration too narrow for double-width input");
  int window = scheds[0]->window;
  for (int j = 1; j < njobs; ++j)
    if (scheds[j]->window != window) fail(CHX_EINVAL, "jobs of one launch must share the window width");
  const int tentries = 1 << (window - 1);
  const size_t tstride = round_up(count, BLOCK_THREADS);
  const size_t tab_words = lazy_table_words(L32, ci, window, count);
  if (tab == nullptr) {
    ctx->slot[S_TAB].ensure(tab_words * njobs * sizeof(uint32_t));
    tab = ctx->slot[S_TAB].as<uint32_t>();
  }
  LazyJob jb[2];
  for (int j = 0; j < 2; ++j) {
    const int s = j < njobs ? j : 0;
    jb[j].md = sets[s].d[ci];
    jb[j].sched = scheds[s]->d;
    jb[j].nsched = (int)scheds[s]->h.size();
    jb[j].in = in[s];
    jb[j].in_hi = in_hi ? in_hi[s] : nullptr;
    jb[j].out = out[s];
    jb[j].table = tab + (size_t)s * tab_words;
  }
  record_start(ctx, "lazy_modexp");
  CHX_DISPATCH_LAZY(L32, ci, {
    hipLaunchKernelGGL((k_lazy_modexp<RB, LP, TPI, L32>), grid_for(count, TPI, njobs), dim3(BLOCK_THREADS),
                       (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, jb[0], jb[1], (int)count, (int)stride, tentries,
                       (int)tstride);
  });
  record_stop(ctx);
  check_launch("k_lazy_modexp");
}

// --------------------------------------------------------------------------
// launchers (device-resident, interleaved)
// --------------------------------------------------------------------------
void launch_mulmod(chx_ctx* ctx, const ModDev* md, int L, const uint32_t* a, const uint32_t* b, uint32_t* out,
                   size_t count, size_t stride) {
  const int tpi = pick_tpi(ctx, CHX_OP_MULMOD, L, count);
  record_start(ctx, "mulmod");
  CHX_DISPATCH(L, tpi, {
    hipLaunchKernelGGL((k_mulmod<L, TPI>), grid_for(count, TPI), dim3(BLOCK_THREADS), lds_bytes(L, TPI),
                       ctx->stream, md, a, b, out, (int)count, (int)stride);
  });
  record_stop(ctx);
  check_launch("k_mulmod");
}

// Pointwise multiply of a modulus handle: the lazy engine for batches of
// kLazyMulmodMin or more (two lazy products per item), the canonical engine
// below that or when the context forces its lane count.
constexpr size_t kLazyMulmodMin = 8192;

// workgroups of k_fixed_mulmod: enough to fill every CU at the occupancy its
// LDS and registers allow (the kernel strides over the tiles)
extern "C++" {
template <class F>
static unsigned fixed_blocks(chx_ctx* ctx, F kernel, int threads, size_t lds, size_t tiles) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  const size_t want = (size_t)device_cus(ctx->device) * per_cu;
  return (unsigned)std::max<size_t>(1, std::min(tiles, want));
}
}

// Head of a work queue in the context's memory, zeroed on its stream (a fill
// kernel) for the launch that follows: launches on one stream take turns.
u32* queue_head(chx_ctx* ctx) {
  ctx->slot[S_QUEUE].ensure(256);
  u32* h = ctx->slot[S_QUEUE].as<u32>();
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, h, 0u, 1);
  check_launch("k_fill_u32");
  return h;
}

// The one-reduction pointwise multiply (k_pw_mulmod) and the served multiply
// run at the positions of the exponentiation engine (lazy configuration 0:
// 29 bits at 2048- and 4096-bit moduli), whose modulus constants (lz.d[0])
// they use, with 18 or 19 positions per chunk.
#define CHX_DISPATCH_PW(L32_, ...)                                                                 \
  do {                                                                                            \
    const int l__ = (L32_);                                                                       \
    if (l__ == 32) { constexpr int RB = 28, LP = 38, TPI = 2, L32 = 32; __VA_ARGS__; }             \
    else if (l__ == 64) { constexpr int RB = 29, LP = 72, TPI = 4, L32 = 64; __VA_ARGS__; }        \
    else if (l__ == 128) { constexpr int RB = 29, LP = 144, TPI = 8, L32 = 128; __VA_ARGS__; }     \
    else fail(CHX_EUNSUPPORTED, "no pointwise kernel for %d limbs", l__);                         \
  } while (0)

static LazyCfg pw_shape(int L32) {
  return L32 == 32 ? LazyCfg{28, 38, 2} : L32 == 64 ? LazyCfg{29, 72, 4} : LazyCfg{29, 144, 8};
}
// Shape of the served multiply (k_fixed_operand, k_fixed_mulmod; query
// operands are prepared in its positions): 72 and 144 positions of 29 bits at
// 2048- and 4096-bit moduli; its row loop folds its accumulators every
// FixedSeg rows (one and two folds per entry).
static LazyCfg serve_cfg(int L32) {
  return L32 == 32 ? LazyCfg{28, 38, 1} : L32 == 64 ? LazyCfg{29, 72, 4} : LazyCfg{29, 144, 8};
}

#define CHX_DISPATCH_SERVE(L32_, ...)                                                            \
  do {                                                                                         \
    const int l__ = (L32_);                                                                    \
    if (l__ == 32) { constexpr int RB = 28, LP = 38, TPI = 1, L32 = 32; __VA_ARGS__; }           \
    else if (l__ == 64) { constexpr int RB = 29, LP = 72, TPI = 4, L32 = 64; __VA_ARGS__; }      \
    else if (l__ == 128) { constexpr int RB = 29, LP = 144, TPI = 8, L32 = 128; __VA_ARGS__; }   \
    else fail(CHX_EUNSUPPORTED, "no served multiply for %d limbs", l__);                         \
  } while (0)

// The served and pointwise multiplies use the modulus constants of lazy
// configuration 0 (lz.d[0]); their positions must be the same.
static const LazyDev* shared_constants(const chx_mod* m, const LazyCfg& c) {
  const LazyCfg z = lazy_cfgs(m->hc.L)[0];
  if (z.rb != c.rb || z.lp != c.lp || !m->lz.d[0]) fail(CHX_EINVAL, "internal: served positions differ from the engine's");
  return m->lz.d[0];
}

// B_i = 2^(RB (i + 2)) R' mod N as positions (LP rows), uploaded once per
// modulus, on first use.
static void serve_constants(chx_ctx* ctx, chx_mod* m) {
  if (lidx(m->hc.L) < 0) fail(CHX_EUNSUPPORTED, "no served multiply for %d limbs", m->hc.L);
  std::lock_guard<std::mutex> lock(m->mu);
  if (m->d_fixb) return;
  const LazyCfg c = serve_cfg(m->hc.L);
  (void)shared_constants(m, c);
  std::vector<uint32_t> h((size_t)c.lp * c.lp);
  BN one(m->hc.L, 0u);
  one[0] = 1;
  for (int i = 0; i < c.lp; ++i) {
    BN b = host::bn_shl_mod(one, c.rb * (i + 2 + c.lp), m->hc.N);
    to_positions(b, c.rb, c.lp, h.data() + (size_t)i * c.lp);
  }
  m->d_fixb = dev_upload(h.data(), h.size(), ctx->stream);
}


// Fold table of k_pw_mulmod: rows r < nz hold Z = 2^(RB (h + r + 2)) mod N as
// canonical positions; the two rows 2g, 2g + 1 of chunk w (positions
// [w K, w K + K)) are 2 CH consecutive words at ((g TPI + w) 2 CH) (CH = K
// rounded up to 4, zero padding). h = 2 LP - nz, with 2^(RB h) <= N / 2
// (h <= (bits(N) - 2) / RB) and nz a multiple of 4; usable when nz <= 2 LP -
// 12 TPI - 2 (the kernel's LDS rows), i.e. for moduli down to about half the
// width.
// Returns false when the modulus cannot take the kernel.
bool pw_constants(chx_ctx* ctx, chx_mod* m) {
  std::lock_guard<std::mutex> lock(m->mu);
  if (m->pw_h == -2) return false;
  if (m->d_pwz) return true;
  const int L = m->hc.L;
  if (lidx(L) < 0) {
    m->pw_h = -2;
    return false;
  }
  const LazyCfg c = pw_shape(L);
  (void)shared_constants(m, c);
  const int k = c.lp / c.tpi, ch = (k + 3) / 4 * 4;
  const int h0 = (host::bn_bits(m->hc.N) - 2) / c.rb;
  const int nz = (2 * c.lp - h0 + 3) / 4 * 4;
  const int h = 2 * c.lp - nz;
  if (h0 < 0 || h < 0 || nz > 2 * c.lp - 12 * c.tpi - 2) {
    m->pw_h = -2;
    return false;
  }
  std::vector<uint32_t> tab((size_t)nz * c.tpi * ch, 0u), p(c.lp);
  BN one(L, 0u);
  one[0] = 1;
  BN z = host::bn_shl_mod(one, c.rb * (h + 2), m->hc.N);
  for (int r = 0; r < nz; ++r) {
    to_positions(z, c.rb, c.lp, p.data());
    for (int q = 0; q < c.lp; ++q) tab[((size_t)(r / 2) * c.tpi + q / k) * 2 * ch + (r % 2) * ch + q % k] = p[q];
    z = host::bn_shl_mod(z, c.rb, m->hc.N);
  }
  m->d_pwz = dev_upload(tab.data(), tab.size(), ctx->stream);
  m->pw_h = h;
  m->pw_nz = nz;
  return true;
}

void launch_mulmod_mod(chx_ctx* ctx, const chx_mod* mc, const uint32_t* a, const uint32_t* b, uint32_t* out,
                       size_t count, size_t stride) {
  chx_mod* m = const_cast<chx_mod*>(mc);
  const int L = m->hc.L;
  const int li = lidx(L);
  if (li < 0 || count < kLazyMulmodMin || ctx->tune_tpi[CHX_OP_MULMOD][li] != 0) {
    launch_mulmod(ctx, m->d, L, a, b, out, count, stride);
    return;
  }
  // window 2 of CHX_OP_MULMOD selects the two-product form below
  if (ctx->tune_window[CHX_OP_MULMOD][li] != 2 && pw_constants(ctx, m)) {
    const size_t tiles = (count + FIX_TILE - 1) / FIX_TILE;
    u32* head = queue_head(ctx);
    record_start(ctx, "pw_mulmod");
    CHX_DISPATCH_PW(L, {
      auto kern = k_pw_mulmod<RB, LP, TPI, L32>;
      const size_t lds = (size_t)PwLds<LP, TPI>::WORDS * 4;
      static std::once_flag lds_attr;
      std::call_once(lds_attr, [&] {
        hip_check(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                  "hipFuncSetAttribute");
      });
      hipLaunchKernelGGL(kern, dim3(fixed_blocks(ctx, kern, 64 * TPI, lds, tiles)), dim3(64 * TPI), lds,
                         ctx->stream, m->lz.d[0], m->d_pwz, m->pw_h, m->pw_nz, a, b, out, (int)count,
                         (int)stride, head);
    });
    record_stop(ctx);
    check_launch("k_pw_mulmod");
    return;
  }
  const int ci = pick_lazy(ctx, L, count);
  record_start(ctx, "mulmod");
  CHX_DISPATCH_LAZY(L, ci, {
    hipLaunchKernelGGL((k_lazy_mulmod<RB, LP, TPI, L32>), grid_for(count, TPI), dim3(BLOCK_THREADS),
                       (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, m->lz.d[ci], a, b, out, (int)count,
                       (int)stride);
  });
  record_stop(ctx);
  check_launch("k_lazy_mulmod");
}

void launch_modexp(chx_ctx* ctx, const chx_mod* m, const uint32_t* base, uint32_t* out, size_t count,
                   size_t stride, const ExpSched& s) {
  const ExpSched* sp = &s;
  run_lazy(ctx, m->hc.L, &m->lz, &sp, &base, &out, 1, count, stride);
}

// per-item exponents (fixed window); exps interleaved, e_limbs limbs per item
void launch_modexp_each(chx_ctx* ctx, const LazySet& ls, const uint32_t* base, const uint32_t* exps, int e_limbs,
                        int ebits, uint32_t* out, size_t count, size_t stride) {
  const int L = ls.L32;
  const int ci = pick_lazy(ctx, L, count, CHX_OP_RAGGED);
  const LazyCfg cfg = lazy_cfgs(L)[ci];
  const int li = lidx(L);
  int w = ctx->tune_window[CHX_OP_RAGGED][li];
  if (!w) w = ebits >= 512 ? 5 : ebits >= 64 ? 4 : 2;
  const int nwin = std::max(1, (ebits + w - 1) / w);
  const size_t tstride = round_up(count, BLOCK_THREADS);
  ctx->slot[S_TAB].ensure(((size_t)1 << w) * cfg.lp * tstride * sizeof(uint32_t));
  record_start(ctx, "lazy_modexp_each");
  CHX_DISPATCH_LAZY(L, ci, {
    hipLaunchKernelGGL((k_lazy_modexp_each<RB, LP, TPI, L32>), grid_for(count, TPI), dim3(BLOCK_THREADS),
                       (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, ls.d[ci], base, exps, e_limbs, nwin, w, out,
                       (int)count, (int)stride, ctx->slot[S_TAB].as<uint32_t>(), (int)tstride);
  });
  record_stop(ctx);
  check_launch("k_lazy_modexp_each");
}

// m = L(c^e mod n^2) * cR mod n (+ optional Hamming popcount): lazy
// exponentiation into a scratch buffer, then the canonical finish.
void launch_decrypt(chx_ctx* ctx, const LazySet& lz, const ModDev* nm, const uint32_t* cR, int L,
                    const uint32_t* c, uint32_t* out, size_t count, size_t stride, const ExpSched& s,
                    const uint32_t* mask, uint32_t* pop) {
  ctx->slot[S_MID0].ensure(stride * L * sizeof(uint32_t));
  uint32_t* u = ctx->slot[S_MID0].as<uint32_t>();
  const ExpSched* sp = &s;
  run_lazy(ctx, L, &lz, &sp, &c, &u, 1, count, stride);
  const int tpi = pick_tpi(ctx, CHX_OP_MULMOD, L, count);
  record_start(ctx, "decrypt_finish");
  CHX_DISPATCH(L, tpi, {
    hipLaunchKernelGGL((k_decrypt_finish<L, TPI>), grid_for(count, TPI), dim3(BLOCK_THREADS), lds_bytes(L, TPI),
                       ctx->stream, nm, cR, u, out, (int)count, (int)stride, mask, pop);
  });
  record_stop(ctx);
  check_launch("k_decrypt_finish");
}

// m = L(u) * cR mod n for u = c^e mod n^2 computed at half width: c^e mod
// p^2 and mod q^2 in one launch (both halves read the whole ciphertext and
// reduce it on entry), then k_lookup_crt_finish recombines u exactly with the
// idempotents eR = [e_p R | e_q R] mod n^2 and finishes as k_decrypt_finish,
// so the result equals launch_decrypt's for every c in [0, n^2).
// fast (optional, a lookup key's textbook CRT constants): k_lookup_crt_fast
// for every item, then the exact finish only for the items it lists.
void launch_exact_crt(chx_ctx* ctx, const LazySet& p2, const LazySet& q2, const ExpSched& s, const ModDev* n2m,
                      const uint32_t* eR, const ModDev* nm, const uint32_t* cR, int L, const uint32_t* c,
                      uint32_t* out, size_t count, size_t stride, const uint32_t* mask, uint32_t* pop,
                      const LookupCrtDev* fast = nullptr) {
  const int H = L / 2;
  const size_t half_words = (size_t)H * stride;
  ctx->slot[S_MID0].ensure(2 * half_words * sizeof(uint32_t));
  uint32_t* u = ctx->slot[S_MID0].as<uint32_t>();
  const LazySet sets[2] = {p2, q2};
  const ExpSched* scheds[2] = {&s, &s};
  const uint32_t* ins[2] = {c, c};
  const uint32_t* his[2] = {c + half_words, c + half_words};
  uint32_t* outs[2] = {u, u + half_words};
  run_lazy(ctx, H, sets, scheds, ins, outs, 2, count, stride, his);
  const int tpi = pick_tpi(ctx, CHX_OP_MULMOD, L, count);
  uint32_t* list = nullptr;
  if (fast) {
    ctx->slot[S_AUX].ensure((count + 1) * sizeof(uint32_t));
    list = ctx->slot[S_AUX].as<uint32_t>();
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, list, 0u, 1);  // a fill kernel, not a memset
    check_launch("k_fill_u32");
    // the finish splits each lane group into a p- and a q-half: at least 2 lanes
    const int ftpi = std::max(pick_tpi(ctx, CHX_OP_MULMOD, H, count), 4);
    record_start(ctx, "decrypt_finish");
    CHX_DISPATCH_H(H, ftpi, {
      hipLaunchKernelGGL((k_lookup_crt_fast<H, TPI>), grid_for(count, TPI), dim3(BLOCK_THREADS), lds_bytes(H, TPI),
                         ctx->stream, fast, u, out, (int)count, (int)stride, list, mask, pop);
    });
    record_stop(ctx);
    check_launch("k_lookup_crt_fast");
  }
  record_start(ctx, "decrypt_finish");  // with `fast`: the listed items only
  CHX_DISPATCH(L, tpi, {
    hipLaunchKernelGGL((k_lookup_crt_finish<L, TPI>), grid_for(count, TPI), dim3(BLOCK_THREADS), lds_bytes(L, TPI),
                       ctx->stream, n2m, eR, nm, cR, u, out, (int)count, (int)stride, mask, pop, list);
  });
  record_stop(ctx);
  check_launch("k_lookup_crt_finish");
}

void launch_encrypt(chx_pkey* k, const uint32_t* m, const uint32_t* r, uint32_t* c, size_t count,
                    size_t stride) {
  chx_ctx* ctx = k->ctx;
  const int L = k->W;
  if (k->d_crt) {  // r^n mod p^2 and mod q^2 at half width, recombined in the finish
    const int H = L / 2;
    const ExpSched& s = get_sched_p(k, 0, pick_window(ctx, CHX_OP_MODEXP, H, exp_bits_p(k, 0)));
    const size_t half_words = (size_t)H * stride;
    ctx->slot[S_MID0].ensure(2 * half_words * sizeof(uint32_t));
    uint32_t* u = ctx->slot[S_MID0].as<uint32_t>();
    const LazySet sets[2] = {k->lz_p2, k->lz_q2};
    const ExpSched* scheds[2] = {&s, &s};
    const uint32_t* ins[2] = {r, r};
    const uint32_t* his[2] = {r + half_words, r + half_words};
    uint32_t* outs[2] = {u, u + half_words};
    run_lazy(ctx, H, sets, scheds, ins, outs, 2, count, stride, his);
    const int tpi = pick_tpi(ctx, CHX_OP_MULMOD, L, count);
    record_start(ctx, "encrypt_finish");
    CHX_DISPATCH(L, tpi, {
      hipLaunchKernelGGL((k_encrypt_crt_finish<L, TPI>), grid_for(count, TPI), dim3(BLOCK_THREADS),
                         lds_bytes(L, TPI), ctx->stream, k->d, k->d_crt, m, u, c, (int)count, (int)stride);
    });
    record_stop(ctx);
    check_launch("k_encrypt_crt_finish");
    return;
  }
  const ExpSched& s = get_sched_p(k, 0, pick_window(ctx, CHX_OP_MODEXP, L, exp_bits_p(k, 0)));
  ctx->slot[S_MID0].ensure(stride * L * sizeof(uint32_t));
  uint32_t* y = ctx->slot[S_MID0].as<uint32_t>();
  const ExpSched* sp = &s;
  run_lazy(ctx, L, &k->lz_n2, &sp, &r, &y, 1, count, stride);  // y = r^n mod n^2
  const int tpi = pick_tpi(ctx, CHX_OP_MULMOD, L, count);
  record_start(ctx, "encrypt_finish");
  CHX_DISPATCH(L, tpi, {
    hipLaunchKernelGGL((k_encrypt_finish<L, TPI>), grid_for(count, TPI), dim3(BLOCK_THREADS), lds_bytes(L, TPI),
                       ctx->stream, k->d, m, y, c, (int)count, (int)stride);
  });
  record_stop(ctx);
  check_launch("k_encrypt_finish");
}

// Scratch handed to launch_crt by the allocation-free entry point: `words`
// 32-bit words at `p` (device memory).
struct FixedScratch {
  uint32_t* p;
  size_t words;
};

// With fixed != nullptr nothing is allocated: the exponentiation uses the
// default window and the schedules built with the key, and the residues and
// window tables are carved from `fixed` (refused if it is too small).
void launch_crt(chx_pkey* k, const uint32_t* c, uint32_t* m, size_t count, size_t stride,
                uint32_t* bad = nullptr, const uint32_t* mask = nullptr, uint32_t* pop = nullptr,
                const FixedScratch* fixed = nullptr) {
  chx_ctx* ctx = k->ctx;
  if (!k->crt) fail(CHX_EINVAL, "key was created without CRT factors");
  const int H = k->H;
  const int ebits = std::max(exp_bits_p(k, 2), exp_bits_p(k, 3));
  const int w = fixed ? host::default_window(ebits) : pick_window(ctx, CHX_OP_MODEXP, H, ebits);
  const ExpSched* spp = fixed ? find_sched_p(k, 2, w) : &get_sched_p(k, 2, w);
  const ExpSched* sqp = fixed ? find_sched_p(k, 3, w) : &get_sched_p(k, 3, w);
  if (!spp || !sqp) fail(CHX_EINVAL, "the key's CRT schedules were not built at creation");
  const ExpSched& sp = *spp;
  const ExpSched& sq = *sqp;
  const size_t half_words = (size_t)H * stride;
  uint32_t* u = nullptr;
  uint32_t* tab = nullptr;
  if (fixed) {
    const size_t u_words = round_up(2 * half_words, 64);  // 256-byte aligned tables after it
    const size_t t_words = 2 * lazy_table_words(H, pick_lazy(ctx, H, count * 2), w, count);
    if (u_words + t_words > fixed->words)
      fail(CHX_EINVAL,
           "%zu ciphertexts (stride %zu) need %zu bytes of scratch, above the %zu-byte array of the "
           "allocation-free entry point",
           count, stride, (u_words + t_words) * sizeof(uint32_t), fixed->words * sizeof(uint32_t));
    u = fixed->p;
    tab = fixed->p + u_words;
  } else {
    ctx->slot[S_MID1].ensure(2 * half_words * sizeof(uint32_t));
    u = ctx->slot[S_MID1].as<uint32_t>();
  }
  // the finish splits each lane group into a p- and a q-half: at least 2 lanes
  const int tpi = std::max(pick_tpi(ctx, CHX_OP_MULMOD, H, count), 4);
  // both halves read the whole ciphertext (rows [0, H) and [H, 2H)) and reduce
  // it modulo p^2 and q^2 on entry
  const LazySet sets[2] = {k->lz_p2, k->lz_q2};
  const ExpSched* scheds[2] = {&sp, &sq};
  const uint32_t* ins[2] = {c, c};
  const uint32_t* his[2] = {c + half_words, c + half_words};
  uint32_t* outs[2] = {u, u + half_words};
  run_lazy(ctx, H, sets, scheds, ins, outs, 2, count, stride, his, tab);
  record_start(ctx, "crt_finish");
  CHX_DISPATCH_H(H, tpi, {
    hipLaunchKernelGGL((k_crt_finish<H, TPI>), grid_for(count, TPI), dim3(BLOCK_THREADS), lds_bytes(H, TPI),
                       ctx->stream, k->d, u, m, (int)count, (int)stride, bad, mask, pop);
  });
  record_stop(ctx);
  check_launch("k_crt_finish");
}

const ExpSched& get_sched_l(chx_lkey* k, int window) {
  auto it = k->scheds.find(window);
  if (it != k->scheds.end()) return it->second;
  ExpSched s;
  s.window = window;
  s.bits = host::bn_bits(k->e_a);
  s.h = host::sliding_schedule(k->e_a, window);
  if (s.h.empty()) fail(CHX_EINVAL, "zero exponent");
  s.d = dev_upload(s.h.data(), s.h.size(), k->ctx->stream);
  return k->scheds.emplace(window, std::move(s)).first->second;
}

uint32_t* lazy_group_table(chx_ctx* ctx, const LazySet& ls, const uint32_t* gtab, size_t groups, int R);
uint32_t* lazy_noise_triples(chx_ctx* ctx, const LazySet& ls, const uint32_t* nptab, const uint32_t* ntab, int nn);

// any_packed (optional) is asked before a row-group table would be built:
// false means that no item has all its digits in the packing alphabet, so no
// wave can use a row group and none is built.
void launch_lookup_encrypt(chx_lkey* k, const uint8_t* digits, const uint8_t* nidx, int nmult, uint32_t* c,
                           size_t count, size_t stride, const std::function<bool()>& any_packed = nullptr) {
  chx_ctx* ctx = k->ctx;
  if (!k->d_gtab || !k->d_ntab) fail(CHX_EINVAL, "lookup key was created without tables (decrypt-only)");
  const int L = k->W;
  const int ci = pick_lazy(ctx, L, count, CHX_OP_LOOKUP);
  // rows per table product: the context's lookup window (0 = quads). The
  // group tables (16^R entries per R rows: 18.9 MB, 200 MB, 2.42 GB and 30.8
  // GB for 256 rows at a 4096-bit n^2; the 128 rows a 512-bit packed
  // embedding uses take half, chx_lookup_key_set_table_rows) are built on first use; without the memory for
  // one, the next smaller group serves. Window 5 also builds the noise-triple
  // table (16.8 M entries, 9.7 GB at a 4096-bit n^2 and 256 noise entries).
  const int window = ctx->tune_window[CHX_OP_LOOKUP][lidx(L)];
  int rows = window == 0 ? 4 : window;
  int packed = -1;  // any_packed's answer, asked once
  auto ensure = [&](int R) {
    chx_lkey::GroupTab& t = k->group[R];
    if (t.d || t.failed || k->chunks < R) return;
    if (any_packed && packed < 0) packed = any_packed() ? 1 : 0;
    if (packed == 0) return;
    size_t groups = (size_t)(k->chunks / R);
    if (k->table_rows > 0) groups = std::min(groups, (size_t)((k->table_rows + R - 1) / R));
    try {
      t.d = lazy_group_table(ctx, k->lz_n2, k->d_gtab, groups, R);
      t.groups = (int)groups;
    } catch (const ChxError& e) {
      if (e.code != CHX_ENOMEM) throw;
      (void)hipGetLastError();  // the failed allocation is not this launch's error
      t.failed = true;
    }
  };
  for (; rows >= 2; --rows) {
    ensure(rows);
    if (k->group[rows].d) break;
  }
  const uint32_t* ptab = rows >= 2 ? k->group[rows].d : nullptr;
  const int pgroups = rows >= 2 ? k->group[rows].groups : 0;
  if (window == 5 && k->noise > 0 && nmult > 3 && !k->d_n3tab && !k->n3tab_failed) {
    try {
      k->d_n3tab = lazy_noise_triples(ctx, k->lz_n2, k->d_nptab, k->d_ntab, k->noise);
    } catch (const ChxError& e) {
      if (e.code != CHX_ENOMEM) throw;
      (void)hipGetLastError();
      k->n3tab_failed = true;
    }
  }
  const uint32_t* n3tab = window == 5 ? k->d_n3tab : nullptr;
  record_start(ctx, "lookup_encrypt");
  CHX_DISPATCH_LAZY(L, ci, {
    hipLaunchKernelGGL((k_lazy_lookup_encrypt<RB, LP, TPI, L32>), grid_for(count, TPI), dim3(BLOCK_THREADS),
                       (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, k->lz_n2.d[ci], k->d_gtab, ptab, rows,
                       pgroups, k->d_ntab, k->d_nptab, n3tab, k->d_ntab_plain, k->d_nptab_plain, digits, k->chunks,
                       nidx, nmult, k->noise, c, (int)count, (int)stride);
  });
  record_stop(ctx);
  check_launch("k_lazy_lookup_encrypt");
}

// convert an AoS table of plain residues (entries x L32 limbs, host) into
// lazy Montgomery positions on the device (entries x LP); returns the buffer
uint32_t* lazy_table_upload(chx_ctx* ctx, const LazySet& ls, const uint32_t* host, size_t entries) {
  const int L = ls.L32;
  const LazyCfg cfg = lazy_cfgs(L)[0];
  const int ci = 0;
  uint32_t* src = nullptr;
  uint32_t* dst = nullptr;
  hip_check(hipMalloc(&src, entries * L * sizeof(uint32_t)), "hipMalloc");
  hipError_t e = hipMalloc(&dst, entries * cfg.lp * sizeof(uint32_t));
  if (e != hipSuccess) {
    (void)hipFree(src);
    fail(CHX_ENOMEM, "hipMalloc: %s", hipGetErrorString(e));
  }
  hip_check(hipMemcpyAsync(src, host, entries * L * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream), "H2D");
  CHX_DISPATCH_LAZY(L, ci, {
    hipLaunchKernelGGL((k_lazy_table<RB, LP, TPI, L32>), grid_for(entries, TPI), dim3(BLOCK_THREADS),
                       (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, ls.d[ci], src, dst, (int)entries);
  });
  check_launch("k_lazy_table");
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  (void)hipFree(src);
  return dst;
}

// noise-pair table from the device noise table (nn entries, positions of
// configuration 0); returns the buffer (nn^2 entries)
uint32_t* lazy_noise_pairs(chx_ctx* ctx, const LazySet& ls, const uint32_t* ntab, int nn) {
  const int L = ls.L32;
  const LazyCfg cfg = lazy_cfgs(L)[0];
  const int ci = 0;
  const size_t entries = (size_t)nn * nn;
  uint32_t* dst = nullptr;
  hipError_t e = hipMalloc(&dst, entries * cfg.lp * sizeof(uint32_t));
  if (e != hipSuccess) fail(CHX_ENOMEM, "hipMalloc: %s", hipGetErrorString(e));
  CHX_DISPATCH_LAZY(L, ci, {
    hipLaunchKernelGGL((k_lazy_noise_pairs<RB, LP, TPI, L32>), grid_for(entries, TPI), dim3(BLOCK_THREADS),
                       (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, ls.d[ci], ntab, nn, dst);
  });
  hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    (void)hipFree(dst);
    fail(CHX_EDEVICE, "k_lazy_noise_pairs: %s", hipGetErrorString(le));
  }
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return dst;
}

// noise-triple table from the device noise-pair and noise tables (nn^2 and
// nn entries, positions of configuration 0); returns the buffer (nn^3 entries)
uint32_t* lazy_noise_triples(chx_ctx* ctx, const LazySet& ls, const uint32_t* nptab, const uint32_t* ntab, int nn) {
  const int L = ls.L32;
  const LazyCfg cfg = lazy_cfgs(L)[0];
  const int ci = 0;
  const size_t entries = (size_t)nn * nn * nn;
  if (!nptab || !ntab) fail(CHX_EINVAL, "noise triples need the noise and noise-pair tables");
  if (entries > (size_t)INT32_MAX / 2) fail(CHX_EINVAL, "too many noise-triple entries");
  uint32_t* dst = nullptr;
  hipError_t e = hipMalloc(&dst, entries * cfg.lp * sizeof(uint32_t));
  if (e != hipSuccess) fail(CHX_ENOMEM, "hipMalloc: %s", hipGetErrorString(e));
  CHX_DISPATCH_LAZY(L, ci, {
    hipLaunchKernelGGL((k_lazy_noise_triples<RB, LP, TPI, L32>), grid_for(entries, TPI), dim3(BLOCK_THREADS),
                       (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, ls.d[ci], nptab, ntab, nn, dst);
  });
  hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    (void)hipFree(dst);
    fail(CHX_EDEVICE, "k_lazy_noise_triples: %s", hipGetErrorString(le));
  }
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return dst;
}

// plain residues (MontMul(e, 1)) of a device table in lazy Montgomery form
// (positions of configuration 0); returns the buffer (entries x LP)
uint32_t* lazy_plain_copy(chx_ctx* ctx, const LazySet& ls, const uint32_t* in, size_t entries) {
  const int L = ls.L32;
  const LazyCfg cfg = lazy_cfgs(L)[0];
  const int ci = 0;
  uint32_t* dst = nullptr;
  hipError_t e = hipMalloc(&dst, entries * cfg.lp * sizeof(uint32_t));
  if (e != hipSuccess) fail(CHX_ENOMEM, "hipMalloc: %s", hipGetErrorString(e));
  CHX_DISPATCH_LAZY(L, ci, {
    hipLaunchKernelGGL((k_lazy_plain<RB, LP, TPI, L32>), grid_for(entries, TPI), dim3(BLOCK_THREADS),
                       (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, ls.d[ci], in, dst, (int)entries);
  });
  hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    (void)hipFree(dst);
    fail(CHX_EDEVICE, "k_lazy_plain: %s", hipGetErrorString(le));
  }
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return dst;
}

// row-group table of `groups` groups of R = 2 to 5 rows from the device g
// table (both in the positions of configuration 0); returns the buffer
// (groups x 16^R entries)
uint32_t* lazy_group_table(chx_ctx* ctx, const LazySet& ls, const uint32_t* gtab, size_t groups, int R) {
  const int L = ls.L32;
  const LazyCfg cfg = lazy_cfgs(L)[0];
  const int ci = 0;
  if (R < 2 || R > 5) fail(CHX_EINVAL, "row groups of %d rows are not built", R);
  const size_t entries = groups * ((size_t)1 << (4 * R));
  if (entries > (size_t)INT32_MAX / 2) fail(CHX_EINVAL, "too many table rows");
  uint32_t* dst = nullptr;
  hipError_t e = hipMalloc(&dst, entries * cfg.lp * sizeof(uint32_t));
  if (e != hipSuccess) fail(CHX_ENOMEM, "hipMalloc: %s", hipGetErrorString(e));
  CHX_DISPATCH_LAZY(L, ci, {
    if (R == 2)
      hipLaunchKernelGGL((k_lazy_group_table<RB, LP, TPI, L32, 2>), grid_for(entries, TPI), dim3(BLOCK_THREADS),
                         (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, ls.d[ci], gtab, dst, (int)entries);
    else if (R == 5)
      hipLaunchKernelGGL((k_lazy_group_table<RB, LP, TPI, L32, 5>), grid_for(entries, TPI), dim3(BLOCK_THREADS),
                         (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, ls.d[ci], gtab, dst, (int)entries);
    else if (R == 4)
      hipLaunchKernelGGL((k_lazy_group_table<RB, LP, TPI, L32, 4>), grid_for(entries, TPI), dim3(BLOCK_THREADS),
                         (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, ls.d[ci], gtab, dst, (int)entries);
    else
      hipLaunchKernelGGL((k_lazy_group_table<RB, LP, TPI, L32, 3>), grid_for(entries, TPI), dim3(BLOCK_THREADS),
                         (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, ls.d[ci], gtab, dst, (int)entries);
  });
  hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    (void)hipFree(dst);
    fail(CHX_EDEVICE, "k_lazy_group_table: %s", hipGetErrorString(le));
  }
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return dst;
}

template <class F>
int with_ctx(chx_ctx* ctx, F&& f) {
  if (!ctx) {
    g_last_error = "null context";
    return CHX_EINVAL;
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  return guarded([&] {
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    f();
  });
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* chx_version(void) { return "cuhepy_amd-chx 0.1 (gfx950)"; }

int chx_ctx_create(int device, chx_ctx** out) {
  return guarded([&] {
    if (!out) fail(CHX_EINVAL, "null out pointer");
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) fail(CHX_EINVAL, "device %d not present (%d visible)", device, n);
    hip_check(hipSetDevice(device), "hipSetDevice");
    init_scratch(device);  // the allocation-free entry point's array, once per device
    chx_ctx* c = new chx_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      fail(CHX_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    c->stream = c->own;
    *out = c;
  });
}

int chx_ctx_destroy(chx_ctx* ctx) {
  if (!ctx) return CHX_OK;
  int rc = guarded([&] {
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    (void)hipStreamSynchronize(ctx->stream);
    flush_timing(ctx);
    for (auto& b : ctx->slot) b.release();
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
  });
  delete ctx;
  return rc;
}

int chx_ctx_set_stream(chx_ctx* ctx, void* stream) {
  return with_ctx(ctx, [&] { ctx->stream = stream ? (hipStream_t)stream : ctx->own; });
}

void* chx_ctx_get_stream(chx_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int chx_ctx_sync(chx_ctx* ctx) {
  return with_ctx(ctx, [&] { hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize"); });
}

int chx_ctx_set_tuning(chx_ctx* ctx, int op, int limbs, int tpi, int window) {
  return with_ctx(ctx, [&] {
    const int li = lidx(limbs);
    if (op < 0 || op > 3 || li < 0) fail(CHX_EINVAL, "bad op %d or limbs %d", op, limbs);
    if (tpi) {
      bool ok = false;
      if (op == CHX_OP_MODEXP || op == CHX_OP_LOOKUP || op == CHX_OP_RAGGED) {
        const auto& cfgs = lazy_cfgs(limbs);
        for (size_t i = 0; i < cfgs.size(); ++i) ok = ok || (cfgs[i].tpi == tpi && lazy_usable(op, limbs, i));
      } else {
        const auto& opts = built_tpi(limbs);
        ok = std::find(opts.begin(), opts.end(), tpi) != opts.end();
      }
      if (!ok) fail(CHX_EUNSUPPORTED, "tpi=%d is not built for op %d at %d limbs", tpi, op, limbs);
    }
    if (window < 0 || window > 7) fail(CHX_EINVAL, "window %d out of range [1, 7]", window);
    if (op == CHX_OP_LOOKUP && window > 5) fail(CHX_EINVAL, "lookup rows per product %d out of range [1, 5]", window);
    if (op == CHX_OP_MULMOD && window > 2) fail(CHX_EINVAL, "mulmod form %d out of range [1, 2]", window);
    ctx->tune_tpi[op][li] = tpi;
    ctx->tune_window[op][li] = window;
  });
}

int chx_supported_tpi(int limbs, int* tpis, int cap) {
  const auto& v = built_tpi(limbs);
  for (int i = 0; i < (int)v.size() && i < cap; ++i) tpis[i] = v[i];
  return (int)v.size();
}

int chx_supported_tpi_op(int op, int limbs, int* tpis, int cap) {
  if (op == CHX_OP_MULMOD) return chx_supported_tpi(limbs, tpis, cap);
  const auto& v = lazy_cfgs(limbs);
  int k = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    if (!lazy_usable(op, limbs, i)) continue;
    if (k < cap) tpis[k] = v[i].tpi;
    ++k;
  }
  return k;
}

size_t chx_ctx_scratch_bytes(chx_ctx* ctx) {
  if (!ctx) return 0;
  size_t s = 0;
  for (auto& b : ctx->slot) s += b.bytes;
  return s;
}

int chx_mod_create(chx_ctx* ctx, const uint32_t* n, int limbs, chx_mod** out) {
  return with_ctx(ctx, [&] {
    if (!n || !out) fail(CHX_EINVAL, "null argument");
    if (lidx(limbs) < 0) fail(CHX_EUNSUPPORTED, "limbs must be 32, 64 or 128 (got %d)", limbs);
    chx_mod* m = new chx_mod();
    m->ctx = ctx;
    try {
      m->hc = host::mont_consts(host::bn_from(n, limbs, limbs));
      ModDev md;
      fill_moddev(md, m->hc);
      m->d = dev_upload(&md, 1, ctx->stream);
      m->lz = make_lazy_set(m->hc, ctx->stream);
    } catch (...) {
      (void)chx_mod_destroy(m);  // frees what was already allocated on the device
      throw;
    }
    *out = m;
  });
}

int chx_mod_destroy(chx_mod* m) {
  if (!m) return CHX_OK;
  int rc = guarded([&] {
    hip_check(hipSetDevice(m->ctx->device), "hipSetDevice");
    if (m->d) (void)hipFree(m->d);
    if (m->d_fixb) (void)hipFree(m->d_fixb);
    if (m->d_pwz) (void)hipFree(m->d_pwz);
    free_lazy_set(m->lz);
  });
  delete m;
  return rc;
}

int chx_mod_limbs(const chx_mod* m) { return m ? m->hc.L : 0; }

int chx_mulmod(chx_ctx* ctx, const chx_mod* m, const uint32_t* a, const uint32_t* b, uint32_t* out,
               size_t count) {
  return with_ctx(ctx, [&] {
    if (!m || !a || !b || !out) fail(CHX_EINVAL, "null argument");
    if (m->ctx->device != ctx->device) fail(CHX_EINVAL, "modulus lives on another device");
    if (count == 0) return;
    const int L = m->hc.L;
    const size_t stride = batch_stride(count);
    uint32_t* da = upload_il(ctx, S_IN0, a, count, L, stride);
    uint32_t* db = upload_il(ctx, S_IN1, b, count, L, stride);
    ctx->slot[S_OUT].ensure(stride * L * 4);
    launch_mulmod_mod(ctx, m, da, db, ctx->slot[S_OUT].as<uint32_t>(), count, stride);
    download_il(ctx, ctx->slot[S_OUT].as<uint32_t>(), out, count, L, stride);
  });
}

int chx_mulmod_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* a, const uint32_t* b, uint32_t* out,
                   size_t count, size_t stride) {
  return with_ctx(ctx, [&] {
    if (!m || !a || !b || !out) fail(CHX_EINVAL, "null argument");
    if (m->ctx->device != ctx->device) fail(CHX_EINVAL, "modulus lives on another device");
    if (count == 0) return;
    check_stride(count, stride);
    launch_mulmod_mod(ctx, m, a, b, out, count, stride);
  });
}

// Query operand of the served multiply: the table Y_i = y 2^(RB (i + 2)) mod N,
// LP rows of TPI chunks of CH words (k_fixed_operand, k_fixed_mulmod).
int chx_mulmod_operand_words(const chx_mod* m) {
  if (!m) return 0;
  const int L = m->hc.L;
  if (L != 32 && L != 64 && L != 128) return 0;
  const LazyCfg c = serve_cfg(L);
  const int k = c.lp / c.tpi;
  return c.lp * c.tpi * ((k + 3) / 4 * 4);
}

static void launch_operands(chx_ctx* ctx, chx_mod* m, const uint32_t* y, uint32_t* ops, size_t count) {
  const int L = m->hc.L;
  serve_constants(ctx, m);
  const LazyCfg c = serve_cfg(L);
  if (count * (size_t)c.lp > (size_t)INT32_MAX / 64) fail(CHX_EINVAL, "too many operands (%zu)", count);
  CHX_DISPATCH_SERVE(L, {
    hipLaunchKernelGGL((k_fixed_operand<RB, LP, TPI, L32>), grid_for(count * LP, TPI), dim3(BLOCK_THREADS),
                       (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, m->lz.d[0], m->d_fixb, y, ops,
                       (int)count);
  });
  check_launch("k_fixed_operand");
}

int chx_mulmod_operands_dev(chx_ctx* ctx, const chx_mod* mc, const uint32_t* y, uint32_t* ops, size_t count) {
  chx_mod* m = const_cast<chx_mod*>(mc);
  return with_ctx(ctx, [&] {
    if (!m || !y || !ops) fail(CHX_EINVAL, "null argument");
    if (m->ctx->device != ctx->device) fail(CHX_EINVAL, "modulus lives on another device");
    if (count == 0) return;
    launch_operands(ctx, m, y, ops, count);
  });
}

int chx_mulmod_operands(chx_ctx* ctx, const chx_mod* mc, const uint32_t* y, uint32_t* ops, size_t count) {
  chx_mod* m = const_cast<chx_mod*>(mc);
  return with_ctx(ctx, [&] {
    if (!m || !y || !ops) fail(CHX_EINVAL, "null argument");
    if (m->ctx->device != ctx->device) fail(CHX_EINVAL, "modulus lives on another device");
    if (count == 0) return;
    const size_t bytes = count * (size_t)m->hc.L * sizeof(uint32_t);
    ctx->slot[S_AUX].ensure(bytes);
    // pageable source: the copy has read it when the call returns; the kernel
    // follows it on the context's stream (no host-device synchronisation)
    hip_check(hipMemcpyAsync(ctx->slot[S_AUX].p, y, bytes, hipMemcpyHostToDevice, ctx->stream), "H2D");
    launch_operands(ctx, m, ctx->slot[S_AUX].as<uint32_t>(), ops, count);
  });
}


// One multi-query launch (k_fixed_mulmod); done/stamps optional (both or neither).
static void multi_launch(chx_ctx* ctx, const chx_mod* m, const uint32_t* x, size_t stride, size_t begin,
                         size_t count, int nq, const uint32_t* const* ops, uint32_t* const* outs,
                         size_t out_stride, uint32_t* done, uint64_t* stamps) {
  if (!m || !x || !ops || !outs) fail(CHX_EINVAL, "null argument");
  if (m->ctx->device != ctx->device) fail(CHX_EINVAL, "modulus lives on another device");
  if (nq < 1 || nq > MQ_MAX) fail(CHX_EINVAL, "nq=%d out of range [1, %d]", nq, MQ_MAX);
  if (count == 0) return;
  if (begin + count > stride) fail(CHX_EINVAL, "range [%zu, %zu) exceeds the index stride %zu", begin,
                                   begin + count, stride);
  if (begin + count > out_stride) fail(CHX_EINVAL, "range exceeds the output stride %zu", out_stride);
  if (stride > kMaxStride || out_stride > kMaxStride) fail(CHX_EINVAL, "stride exceeds the maximum of %zu", kMaxStride);
  MultiQuery mq;
  memset(&mq, 0, sizeof(mq));
  mq.nq = nq;
  for (int q = 0; q < nq; ++q) {
    if (!ops[q] || !outs[q]) fail(CHX_EINVAL, "query %d: null pointer", q);
    mq.op[q] = ops[q];
    mq.out[q] = outs[q];
  }
  const int L = m->hc.L;
  const size_t tiles = (count + FIX_TILE - 1) / FIX_TILE;
  if (done || stamps) {
    if (!done || !stamps) fail(CHX_EINVAL, "done and stamps go together");
    mq.done = done;
    mq.stamp = reinterpret_cast<u64*>(stamps);
    // counters, the unit queue's head and the stamps from zero (fill kernels,
    // no memset node if the caller captures)
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, done, 0u, nq + 1);
    check_launch("k_fill_u32");
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, reinterpret_cast<u32*>(stamps), 0u, 2 * nq);
    check_launch("k_fill_u32");
    mq.head = done + nq;
  } else {
    mq.head = queue_head(ctx);
  }
  record_start(ctx, "mulmod_multi");
  CHX_DISPATCH_SERVE(L, {
    auto kern = k_fixed_mulmod<RB, LP, TPI, L32>;
    const size_t lds = (size_t)FixedLds<LP, TPI>::WORDS * 4;
    static std::once_flag lds_attr;
    std::call_once(lds_attr, [&] {
      hip_check(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                "hipFuncSetAttribute");
    });
    hipLaunchKernelGGL(kern, dim3(fixed_blocks(ctx, kern, 64 * TPI, lds, tiles * nq)), dim3(64 * TPI), lds,
                       ctx->stream, m->lz.d[0], x, (int)begin, (int)count, (int)stride, mq,
                       (int)out_stride);
  });
  record_stop(ctx);
  check_launch("k_fixed_mulmod");
}

int chx_mulmod_multi_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* x, size_t stride, size_t begin,
                         size_t count, int nq, const uint32_t* const* ops, uint32_t* const* outs,
                         size_t out_stride) {
  return with_ctx(ctx, [&] {
    multi_launch(ctx, m, x, stride, begin, count, nq, ops, outs, out_stride, nullptr, nullptr);
  });
}

int chx_mulmod_multi_stamped_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* x, size_t stride, size_t begin,
                                 size_t count, int nq, const uint32_t* const* ops, uint32_t* const* outs,
                                 size_t out_stride, uint32_t* done, uint64_t* stamps) {
  return with_ctx(ctx, [&] {
    if (!done || !stamps) fail(CHX_EINVAL, "null argument");
    multi_launch(ctx, m, x, stride, begin, count, nq, ops, outs, out_stride, done, stamps);
  });
}

// The stamps' clock rate, measured: two reads of the device clock (k_clock)
// about 20 ms apart against the host's steady clock (the rate the runtime
// reports as the wall clock rate need not be this counter's).
int chx_ctx_wall_clock_hz(chx_ctx* ctx, double* hz) {
  return with_ctx(ctx, [&] {
    if (!hz) fail(CHX_EINVAL, "null argument");
    ctx->slot[S_AUX].ensure(16);
    u64* d = ctx->slot[S_AUX].as<u64>();
    auto read = [&](u64* host) {
      hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, ctx->stream, d);
      check_launch("k_clock");
      hip_check(hipMemcpyAsync(host, d, 8, hipMemcpyDeviceToHost, ctx->stream), "D2H");
      hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    };
    u64 a = 0, b = 0;
    read(&a);  // warm
    const auto t0 = std::chrono::steady_clock::now();
    read(&a);
    const auto t1 = std::chrono::steady_clock::now();
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    const auto t2 = std::chrono::steady_clock::now();
    read(&b);
    const auto t3 = std::chrono::steady_clock::now();
    // each read lies between its bracketing host times: take the midpoints
    const double host = std::chrono::duration<double>((t2 - t0) + (t3 - t1)).count() / 2;
    if (b <= a || host <= 0) fail(CHX_EDEVICE, "device clock did not advance");
    *hz = (double)(b - a) / host;
  });
}

int chx_mulmod_bcast_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* x, const uint32_t* y_host,
                         uint32_t* out, size_t count, size_t stride) {
  return with_ctx(ctx, [&] {
    if (!m || !x || !y_host || !out) fail(CHX_EINVAL, "null argument");
    if (m->ctx->device != ctx->device) fail(CHX_EINVAL, "modulus lives on another device");
    if (count == 0) return;
    const int L = m->hc.L;
    check_stride(count, stride);
    BN yR = host::to_mont_host(host::bn_mod(host::bn_from(y_host, L, L), m->hc.N), m->hc);
    ctx->slot[S_AUX].ensure(L * 4);
    hip_check(hipMemcpyAsync(ctx->slot[S_AUX].p, yR.data(), L * 4, hipMemcpyHostToDevice, ctx->stream), "H2D");
    const int tpi = pick_tpi(ctx, CHX_OP_MULMOD, L, count);
    record_start(ctx, "mulmod_bcast");
    CHX_DISPATCH(L, tpi, {
      hipLaunchKernelGGL((k_mulmod_bcast<L, TPI>), grid_for(count, TPI), dim3(BLOCK_THREADS), lds_bytes(L, TPI),
                         ctx->stream, m->d, x, ctx->slot[S_AUX].as<uint32_t>(), out, (int)count, (int)stride);
    });
    record_stop(ctx);
    check_launch("k_mulmod_bcast");
  });
}

static ExpSched make_sched(chx_ctx* ctx, const uint32_t* e, int e_limbs, int L, int op) {
  BN eb = host::bn_from(e, e_limbs, e_limbs);
  ExpSched s;
  s.bits = host::bn_bits(eb);
  s.window = pick_window(ctx, op, L, s.bits);
  s.h = host::sliding_schedule(eb, s.window);
  return s;
}

int chx_modexp(chx_ctx* ctx, const chx_mod* m, const uint32_t* base, size_t count, const uint32_t* e,
               int e_limbs, uint32_t* out) {
  return with_ctx(ctx, [&] {
    if (!m || !base || !out || (!e && e_limbs)) fail(CHX_EINVAL, "null argument");
    if (m->ctx->device != ctx->device) fail(CHX_EINVAL, "modulus lives on another device");
    if (count == 0) return;
    const int L = m->hc.L;
    check_below(base, count, L, m->hc.N, "base");
    ExpSched s = make_sched(ctx, e, e_limbs, L, CHX_OP_MODEXP);
    if (s.h.empty()) {  // e == 0: result 1 (N > 1)
      std::vector<uint32_t> one((size_t)count * L, 0u);
      for (size_t i = 0; i < count; ++i) one[i * L] = 1u;
      memcpy(out, one.data(), one.size() * 4);
      return;
    }
    const size_t stride = batch_stride(count);
    uint32_t* db = upload_il(ctx, S_IN0, base, count, L, stride);
    ctx->slot[S_AUX].ensure(s.h.size() * 4);
    hip_check(hipMemcpyAsync(ctx->slot[S_AUX].p, s.h.data(), s.h.size() * 4, hipMemcpyHostToDevice, ctx->stream),
              "H2D");
    s.d = ctx->slot[S_AUX].as<uint32_t>();
    ctx->slot[S_OUT].ensure(stride * L * 4);
    launch_modexp(ctx, m, db, ctx->slot[S_OUT].as<uint32_t>(), count, stride, s);
    download_il(ctx, ctx->slot[S_OUT].as<uint32_t>(), out, count, L, stride);
    s.d = nullptr;
  });
}

int chx_modexp_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* base, uint32_t* out, size_t count,
                   size_t stride, const uint32_t* e, int e_limbs) {
  return with_ctx(ctx, [&] {
    if (!m || !base || !out || !e) fail(CHX_EINVAL, "null argument");
    if (m->ctx->device != ctx->device) fail(CHX_EINVAL, "modulus lives on another device");
    if (count == 0) return;
    check_stride(count, stride);
    const int L = m->hc.L;
    ExpSched s = make_sched(ctx, e, e_limbs, L, CHX_OP_MODEXP);
    if (s.h.empty()) fail(CHX_EINVAL, "zero exponent is handled by the host-buffer entry point only");
    ctx->slot[S_AUX].ensure(s.h.size() * 4);
    hip_check(hipMemcpyAsync(ctx->slot[S_AUX].p, s.h.data(), s.h.size() * 4, hipMemcpyHostToDevice, ctx->stream),
              "H2D");
    s.d = ctx->slot[S_AUX].as<uint32_t>();
    launch_modexp(ctx, m, base, out, count, stride, s);
    s.d = nullptr;
  });
}

int chx_modexp_each(chx_ctx* ctx, const chx_mod* m, const uint32_t* base, size_t count, const uint32_t* exps,
                    int e_limbs, uint32_t* out) {
  return with_ctx(ctx, [&] {
    if (!m || !base || !out || !exps || e_limbs <= 0) fail(CHX_EINVAL, "null argument");
    if (m->ctx->device != ctx->device) fail(CHX_EINVAL, "modulus lives on another device");
    if (count == 0) return;
    const int L = m->hc.L;
    check_below(base, count, L, m->hc.N, "base");
    int ebits = 0;
    for (size_t i = 0; i < count * (size_t)e_limbs; ++i)
      if (exps[i]) ebits = std::max(ebits, 32 * (int)(i % e_limbs) + 32 - __builtin_clz(exps[i]));
    const size_t stride = batch_stride(count);
    uint32_t* db = upload_il(ctx, S_IN0, base, count, L, stride);
    uint32_t* de = upload_il(ctx, S_IN1, exps, count, e_limbs, stride);
    ctx->slot[S_OUT].ensure(stride * L * 4);
    launch_modexp_each(ctx, m->lz, db, de, e_limbs, std::max(ebits, 1), ctx->slot[S_OUT].as<uint32_t>(), count,
                       stride);
    download_il(ctx, ctx->slot[S_OUT].as<uint32_t>(), out, count, L, stride);
  });
}

int chx_modexp_each_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* base, const uint32_t* exps, int e_limbs,
                        int e_bits, uint32_t* out, size_t count, size_t stride) {
  return with_ctx(ctx, [&] {
    if (!m || !base || !out || !exps || e_limbs <= 0 || e_bits <= 0 || e_bits > 32 * e_limbs)
      fail(CHX_EINVAL, "bad argument");
    if (m->ctx->device != ctx->device) fail(CHX_EINVAL, "modulus lives on another device");
    if (count == 0) return;
    check_stride(count, stride);
    launch_modexp_each(ctx, m->lz, base, exps, e_limbs, e_bits, out, count, stride);
  });
}

// ---------------------------------------------------------------------------
// ragged batches
// ---------------------------------------------------------------------------
}  // extern "C"

struct chx_ragged {
  chx_ctx* ctx = nullptr;
  int variant = 0, window = 5, nsubs = 0, total_tiles = 0, blocks = 0;
  std::vector<chx_ragged_sub> subs;
  std::vector<RaggedSub> hsubs;  // plain form (one entry per sub-batch, input order)
  // persistent form: [0] all sub-batches (mixed kernel) or [0] 64-limb and
  // [1] 128-limb sub-batches (one persistent kernel per width)
  int nq = 0;
  int q_nsubs[2] = {0, 0}, q_tiles[2] = {0, 0};
  RaggedSub* d_q[2] = {nullptr, nullptr};
  uint32_t* d_counter = nullptr;
  uint32_t* d_scratch = nullptr;
  uint32_t* d_plain_tab = nullptr;
  size_t wave_words = 0, plain_tab_words = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// variant 0: per-width persistent kernels (4 / 8 lanes), variant 1: per-width
// (8 / 16 lanes), variant 2: one mixed persistent kernel (8 / 16 lanes),
// variant 3: one mixed persistent kernel (4 / 8 lanes). The mixed queues hold
// the costliest tiles first (longest-processing-time order), so the cheap ones
// fill the tail.
const int kRaggedCfg[4] = {0, 1, 1, 0};  // lazy configuration index used by each variant

void ragged_enqueue(chx_ragged* rg) {
  chx_ctx* ctx = rg->ctx;
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, rg->d_counter, 0u, 2);
  auto mixed = [&](auto e1, auto e2, auto kernel) {
    typedef decltype(e1) E1;
    typedef decltype(e2) E2;
    constexpr int AREA = E1::AREA > E2::AREA ? E1::AREA : E2::AREA;
    hipLaunchKernelGGL(kernel, dim3(rg->blocks), dim3(BLOCK_THREADS), (size_t)WAVES_PER_BLOCK * AREA * 4,
                       ctx->stream, rg->d_q[0], rg->q_nsubs[0], rg->q_tiles[0], rg->d_counter, rg->window,
                       rg->d_scratch, rg->wave_words);
  };
  if (rg->variant == 2) {
    mixed(Lazy<29, 72, 8, 64>(), Lazy<29, 144, 16, 128>(), k_ragged2<29, 72, 8, 29, 144, 16>);
    return;
  }
  if (rg->variant == 3) {
    mixed(Lazy<29, 72, 4, 64>(), Lazy<29, 144, 8, 128>(), k_ragged2<29, 72, 4, 29, 144, 8>);
    return;
  }
  for (int q = 0; q < 2; ++q) {
    if (!rg->q_nsubs[q]) continue;
    const int L = q == 0 ? 64 : 128;
    const int ci = kRaggedCfg[rg->variant];
    CHX_DISPATCH_LAZY(L, ci, {
      typedef Lazy<RB, LP, TPI, L32> E;
      hipLaunchKernelGGL((k_ragged1<RB, LP, TPI, L32>), dim3(rg->blocks), dim3(BLOCK_THREADS),
                         (size_t)WAVES_PER_BLOCK * E::AREA * 4, ctx->stream, rg->d_q[q], rg->q_nsubs[q],
                         rg->q_tiles[q], rg->d_counter + q, rg->window, rg->d_scratch, rg->wave_words);
    });
  }
}

}  // namespace

extern "C" {

int chx_ragged_create(chx_ctx* ctx, const chx_ragged_sub* subs, int nsubs, int variant, int window,
                      int waves_per_cu, chx_ragged** out) {
  return with_ctx(ctx, [&] {
    if (!subs || nsubs <= 0 || !out) fail(CHX_EINVAL, "bad argument");
    if (variant < 0 || variant > 3) fail(CHX_EUNSUPPORTED, "ragged variant %d not built", variant);
    if (window == 0) window = 5;
    if (window < 2 || window > 6) fail(CHX_EINVAL, "window %d out of range [2, 6]", window);
    if (waves_per_cu == 0) waves_per_cu = 8;
    if (waves_per_cu < 1 || waves_per_cu > 16) fail(CHX_EINVAL, "waves_per_cu out of range");
    std::unique_ptr<chx_ragged> rg(new chx_ragged());
    rg->ctx = ctx;
    rg->variant = variant;
    rg->window = window;
    rg->nsubs = nsubs;
    rg->subs.assign(subs, subs + nsubs);
    const int ci = kRaggedCfg[variant];
    const bool is_mixed = variant >= 2;
    std::vector<RaggedSub> q[2];
    int tiles = 0;
    int lp_max_g = 0;
    // queue order: input order per width; costliest tiles first when mixed
    std::vector<int> order(nsubs);
    for (int s = 0; s < nsubs; ++s) order[s] = s;
    if (is_mixed) {
      auto tile_cost = [&](int s) {
        const int L = subs[s].mod ? subs[s].mod->hc.L : 0;
        const auto& cf = lazy_cfgs(L == 128 ? 128 : 64)[ci];
        return (double)subs[s].e_bits * cf.lp * cf.lp * (64 / cf.tpi);
      };
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tile_cost(a) > tile_cost(b); });
    }
    std::vector<RaggedSub> by_input(nsubs);
    for (int s : order) {
      const chx_ragged_sub& u = subs[s];
      if (!u.mod || !u.base || !u.exps || !u.out) fail(CHX_EINVAL, "sub-batch %d: null pointer", s);
      const int L = u.mod->hc.L;
      if (L != 64 && L != 128) fail(CHX_EUNSUPPORTED, "sub-batch %d: width %d limbs (64 or 128 supported)", s, L);
      if (u.count == 0 || u.stride < round_up(u.count, 64) || u.stride > kMaxStride)
        fail(CHX_EINVAL, "sub-batch %d: bad count/stride", s);
      if (u.e_limbs == 0 || u.e_bits == 0 || u.e_bits > 32 * u.e_limbs)
        fail(CHX_EINVAL, "sub-batch %d: bad exponent size", s);
      const LazyCfg cfg = lazy_cfgs(L)[ci];
      const int G = 64 / cfg.tpi;
      RaggedSub r;
      r.md = u.mod->lz.d[ci];
      r.base = u.base;
      r.exps = u.exps;
      r.out = u.out;
      r.width = L;
      r.count = (int)u.count;
      r.stride = (int)u.stride;
      r.e_limbs = (int)u.e_limbs;
      r.nwin = ((int)u.e_bits + window - 1) / window;
      const int qi = is_mixed ? 0 : (L == 64 ? 0 : 1);
      r.tile0 = rg->q_tiles[qi];
      rg->q_tiles[qi] += ((int)u.count + G - 1) / G;
      tiles += ((int)u.count + G - 1) / G;
      lp_max_g = std::max(lp_max_g, cfg.lp * G);
      by_input[s] = r;
      q[qi].push_back(r);
    }
    rg->hsubs = by_input;
    for (int i = 0; i < 2; ++i) {
      rg->q_nsubs[i] = (int)q[i].size();
      if (!q[i].empty()) rg->d_q[i] = dev_upload(q[i].data(), q[i].size(), ctx->stream);
    }
    rg->total_tiles = tiles;
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, ctx->device), "hipGetDeviceProperties");
    const int waves = prop.multiProcessorCount * waves_per_cu;
    rg->blocks = std::max(1, waves / WAVES_PER_BLOCK);
    rg->wave_words = ((size_t)1 << window) * lp_max_g;
    hip_check(hipMalloc(&rg->d_counter, 64), "hipMalloc");
    hip_check(hipMalloc(&rg->d_scratch, (size_t)rg->blocks * WAVES_PER_BLOCK * rg->wave_words * 4), "hipMalloc");
    *out = rg.release();
  });
}

int chx_ragged_run(chx_ragged* rg, int use_graph) {
  if (!rg) {
    g_last_error = "null plan";
    return CHX_EINVAL;
  }
  chx_ctx* ctx = rg->ctx;
  return with_ctx(ctx, [&] {
    if (!use_graph) {
      record_start(ctx, "ragged");
      ragged_enqueue(rg);
      record_stop(ctx);
      check_launch("k_ragged1/k_ragged2");
      return;
    }
    if (!rg->exec) {
      hip_check(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      ragged_enqueue(rg);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(ctx->stream, &g);
      hip_check(e, "hipStreamEndCapture");
      rg->graph = g;
      hip_check(hipGraphInstantiate(&rg->exec, rg->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    record_start(ctx, "ragged");
    hip_check(hipGraphLaunch(rg->exec, ctx->stream), "hipGraphLaunch");
    record_stop(ctx);
  });
}

int chx_ragged_run_plain(chx_ragged* rg) {
  if (!rg) {
    g_last_error = "null plan";
    return CHX_EINVAL;
  }
  chx_ctx* ctx = rg->ctx;
  return with_ctx(ctx, [&] {
    size_t need = 0;
    for (const auto& u : rg->subs) {
      const LazyCfg cfg = lazy_cfgs(u.mod->hc.L)[kRaggedCfg[rg->variant]];
      need = std::max(need, ((size_t)1 << rg->window) * cfg.lp * round_up(u.count, BLOCK_THREADS));
    }
    if (need > rg->plain_tab_words) {
      if (rg->d_plain_tab) (void)hipFree(rg->d_plain_tab);
      rg->d_plain_tab = nullptr;
      hip_check(hipMalloc(&rg->d_plain_tab, need * 4), "hipMalloc");
      rg->plain_tab_words = need;
    }
    record_start(ctx, "ragged_plain");
    for (size_t s = 0; s < rg->subs.size(); ++s) {
      const chx_ragged_sub& u = rg->subs[s];
      const int L = u.mod->hc.L;
      const int ci = kRaggedCfg[rg->variant];
      const size_t tstride = round_up(u.count, BLOCK_THREADS);
      CHX_DISPATCH_LAZY(L, ci, {
        hipLaunchKernelGGL((k_lazy_modexp_each<RB, LP, TPI, L32>), grid_for(u.count, TPI), dim3(BLOCK_THREADS),
                           (lazy_lds_bytes<RB, LP, TPI, L32>()), ctx->stream, rg->hsubs[s].md, u.base, u.exps,
                           (int)u.e_limbs, rg->hsubs[s].nwin, rg->window, u.out, (int)u.count, (int)u.stride,
                           rg->d_plain_tab, (int)tstride);
      });
    }
    record_stop(ctx);
    check_launch("k_lazy_modexp_each");
  });
}

int chx_ragged_destroy(chx_ragged* rg) {
  if (!rg) return CHX_OK;
  int rc = guarded([&] {
    hip_check(hipSetDevice(rg->ctx->device), "hipSetDevice");
    (void)hipStreamSynchronize(rg->ctx->stream);
    if (rg->exec) (void)hipGraphExecDestroy(rg->exec);
    if (rg->graph) (void)hipGraphDestroy(rg->graph);
    for (void* p : {(void*)rg->d_q[0], (void*)rg->d_q[1], (void*)rg->d_counter, (void*)rg->d_scratch,
                    (void*)rg->d_plain_tab})
      if (p) (void)hipFree(p);
  });
  delete rg;
  return rc;
}

// ---------------------------------------------------------------------------
// Paillier
// ---------------------------------------------------------------------------

int chx_paillier_key_create(chx_ctx* ctx, const uint32_t* n, int n_limbs, const uint32_t* phi,
                            const uint32_t* inv, const uint32_t* p, const uint32_t* q, int half_limbs,
                            const uint32_t* hp, const uint32_t* hq, const uint32_t* qinv, chx_pkey** out) {
  return with_ctx(ctx, [&] {
    if (!n || !phi || !inv || !out || n_limbs <= 0 || n_limbs > 64) fail(CHX_EINVAL, "bad key arguments");
    BN nb = host::bn_from(n, n_limbs, n_limbs);
    const int nbits = host::bn_bits(nb);
    if (nbits < 8) fail(CHX_EINVAL, "modulus too small");
    const int W = pick_width(2 * nbits);
    std::unique_ptr<chx_pkey> k(new chx_pkey());
    k->ctx = ctx;
    k->W = W;
    k->H = W / 2;
    // n^2
    BN nW = host::bn_from(n, n_limbs, W);
    BN n2 = host::bn_sqr_lo(nW);
    k->n2 = host::mont_consts(n2);
    k->n = host::mont_consts(nW);
    BN phiW = host::bn_from(phi, n_limbs, W);
    BN invW = host::bn_from(inv, n_limbs, W);
    if (host::bn_cmp(invW, nW) >= 0) fail(CHX_EINVAL, "inv must be below n");
    k->e_n = nW;
    k->e_phi = phiW;
    std::unique_ptr<PaillierDev> pd(new PaillierDev());
    memset(pd.get(), 0, sizeof(PaillierDev));
    fill_moddev(pd->n2, k->n2);
    fill_moddev(pd->n, k->n);
    // n * R mod n^2: n < n^2, shift n by 32W bits mod n^2
    fill_val(pd->nR, host::to_mont_host(nW, k->n2));
    fill_val(pd->muR, host::to_mont_host(invW, k->n));
    try {  // a failure below frees what was already allocated on the device
      if (p && q && hp && hq && qinv && half_limbs > 0) {
        const int H = W / 2;
        if (half_limbs > H) fail(CHX_EINVAL, "factors wider than half the width");
        // the textbook CRT pass equals the reference formula only for a
        // consistent key: refuse any other
        check_factors(n, n_limbs, phi, inv, p, q, half_limbs, hp, hq, qinv);
        if (H != 32 && H != 64) fail(CHX_EUNSUPPORTED, "CRT needs 32- or 64-limb halves (width %d)", W);
        BN pb = host::bn_from(p, half_limbs, H), qb = host::bn_from(q, half_limbs, H);
        // p^2 and q^2 must fit in H limbs, so p and q in H/2 (the CRT finish
        // works modulo p and q at that width)
        if (host::bn_bits(pb) > 16 * H || host::bn_bits(qb) > 16 * H)
          fail(CHX_EINVAL, "p and q must be below 2^%d for %d-limb CRT halves", 16 * H, H);
        // p^2, q^2 at H limbs
        auto sq = [](const BN& x) { return host::bn_sqr_lo(x); };
        host::MontConsts cp2 = host::mont_consts(sq(pb)), cq2 = host::mont_consts(sq(qb));
        BN nH = host::bn_from(n, n_limbs, H);
        host::MontConsts cn = host::mont_consts(nH);
        fill_moddev(pd->p2, cp2);
        fill_moddev(pd->q2, cq2);
        k->p2h = cp2.N;
        k->q2h = cq2.N;
        k->lz_p2 = make_lazy_set(cp2, ctx->stream);
        k->lz_q2 = make_lazy_set(cq2, ctx->stream);
        fill_moddev(pd->nh, cn);
        const int Hh = H / 2;
        BN ph = host::bn_from(pb.data(), Hh, Hh), qh = host::bn_from(qb.data(), Hh, Hh);
        host::MontConsts cph = host::mont_consts(ph), cqh = host::mont_consts(qh);
        fill_moddev(pd->ph, cph);
        fill_moddev(pd->qh, cqh);
        auto below = [&](const uint32_t* v, const BN& m, const char* what) {
          BN x = host::bn_from(v, half_limbs, H);
          for (int i = Hh; i < H; ++i)
            if (x[i]) fail(CHX_EINVAL, "%s must be below its modulus", what);
          x.resize(Hh);
          if (host::bn_cmp(x, m) >= 0) fail(CHX_EINVAL, "%s must be below its modulus", what);
          return x;
        };
        fill_val(pd->hpR, host::to_mont_host(below(hp, ph, "h_p"), cph));
        fill_val(pd->hqR, host::to_mont_host(below(hq, qh, "h_q"), cqh));
        fill_val(pd->qinvR, host::to_mont_host(below(qinv, ph, "q^-1 mod p"), cph));
        fill_val(pd->qRn, host::to_mont_host(qb, cn));
        BN one(H, 0u);
        one[0] = 1;
        k->e_pm1 = pb;
        host::bn_sub(k->e_pm1, one);
        k->e_qm1 = qb;
        host::bn_sub(k->e_qm1, one);
        k->crt = true;
        // the default-window CRT schedules now, so that the allocation-free
        // entry point (chx_paillier_decrypt_crt_noalloc) finds them on the device
        const int w = host::default_window(std::max(exp_bits_p(k.get(), 2), exp_bits_p(k.get(), 3)));
        (void)get_sched_p(k.get(), 2, w);
        (void)get_sched_p(k.get(), 3, w);
      }
      k->lz_n2 = make_lazy_set(k->n2, ctx->stream);
      k->d = dev_upload(pd.get(), 1, ctx->stream);
    } catch (...) {
      (void)chx_paillier_key_destroy(k.release());
      throw;
    }
    *out = k.release();
  });
}

int chx_paillier_key_destroy(chx_pkey* k) {
  if (!k) return CHX_OK;
  int rc = guarded([&] {
    hip_check(hipSetDevice(k->ctx->device), "hipSetDevice");
    if (k->d) (void)hipFree(k->d);
    if (k->d_crt) (void)hipFree(k->d_crt);
    free_lazy_set(k->lz_n2);
    free_lazy_set(k->lz_p2);
    free_lazy_set(k->lz_q2);
    for (auto& s : k->scheds)
      if (s.second.d) (void)hipFree(s.second.d);
  });
  delete k;
  return rc;
}

int chx_paillier_width(const chx_pkey* k) { return k ? k->W : 0; }

// Checked CRT idempotents -> [e_p R | e_q R] mod n^2 in device memory
static uint32_t* upload_idempotents(chx_ctx* ctx, const host::MontConsts& n2, const BN& p2, const BN& q2,
                                    const uint32_t* ep, const uint32_t* eq) {
  const int W = n2.L, H = (int)p2.size();
  BN e1 = host::bn_from(ep, W, W), e2 = host::bn_from(eq, W, W);
  if (host::bn_cmp(e1, n2.N) >= 0 || host::bn_cmp(e2, n2.N) >= 0) fail(CHX_EINVAL, "CRT idempotents must be below n^2");
  BN one(H, 0u), zero(H, 0u);
  one[0] = 1;
  if (host::bn_cmp(host::bn_mod(e1, p2), one) != 0 || host::bn_cmp(host::bn_mod(e1, q2), zero) != 0 ||
      host::bn_cmp(host::bn_mod(e2, q2), one) != 0 || host::bn_cmp(host::bn_mod(e2, p2), zero) != 0)
    fail(CHX_EINVAL, "e_p must be 1 mod p^2 and 0 mod q^2, e_q the converse");
  std::vector<uint32_t> er(2 * (size_t)W);
  BN t1 = host::to_mont_host(e1, n2), t2 = host::to_mont_host(e2, n2);
  std::copy(t1.begin(), t1.end(), er.begin());
  std::copy(t2.begin(), t2.end(), er.begin() + W);
  return dev_upload(er.data(), 2 * W, ctx->stream);
}

int chx_paillier_key_set_idempotents(chx_ctx* ctx, chx_pkey* k, const uint32_t* ep, const uint32_t* eq) {
  return with_ctx(ctx, [&] {
    if (!k || !ep || !eq) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (!k->crt) fail(CHX_EINVAL, "key was created without CRT factors");
    if (k->d_crt) fail(CHX_EINVAL, "key already has CRT idempotents");
    k->d_crt = upload_idempotents(ctx, k->n2, k->p2h, k->q2h, ep, eq);
  });
}

int chx_paillier_key_has_idempotents(const chx_pkey* k) { return k && k->d_crt ? 1 : 0; }

int chx_paillier_encrypt(chx_ctx* ctx, const chx_pkey* kc, const uint32_t* m, const uint32_t* r, uint32_t* c,
                         size_t count) {
  chx_pkey* k = const_cast<chx_pkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !m || !r || !c) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (count == 0) return;
    const int L = k->W;
    check_below(m, count, L, k->n.N, "m");
    check_below(r, count, L, k->n.N, "r");
    const size_t stride = batch_stride(count);
    uint32_t* dm = upload_il(ctx, S_IN0, m, count, L, stride);
    uint32_t* dr = upload_il(ctx, S_IN1, r, count, L, stride);
    ctx->slot[S_OUT].ensure(stride * L * 4);
    launch_encrypt(k, dm, dr, ctx->slot[S_OUT].as<uint32_t>(), count, stride);
    download_il(ctx, ctx->slot[S_OUT].as<uint32_t>(), c, count, L, stride);
  });
}

int chx_paillier_encrypt_dev(chx_ctx* ctx, const chx_pkey* kc, const uint32_t* m, const uint32_t* r,
                             uint32_t* c, size_t count, size_t stride) {
  chx_pkey* k = const_cast<chx_pkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !m || !r || !c) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (count == 0) return;
    check_stride(count, stride);
    launch_encrypt(k, m, r, c, count, stride);
  });
}

static void paillier_decrypt_impl(chx_pkey* k, const uint32_t* dc, uint32_t* dm, size_t count, size_t stride,
                                  const uint32_t* mask, uint32_t* pop) {
  chx_ctx* ctx = k->ctx;
  const int L = k->W;
  PaillierDev* d = k->d;
  if (k->d_crt) {  // c^phi at half width, recombined exactly
    const ExpSched& s = get_sched_p(k, 1, pick_window(ctx, CHX_OP_MODEXP, L / 2, exp_bits_p(k, 1)));
    launch_exact_crt(ctx, k->lz_p2, k->lz_q2, s, &d->n2, k->d_crt, &d->n, d->muR.v, L, dc, dm, count, stride, mask,
                     pop);
    return;
  }
  const ExpSched& s = get_sched_p(k, 1, pick_window(ctx, CHX_OP_MODEXP, L, exp_bits_p(k, 1)));
  launch_decrypt(ctx, k->lz_n2, &d->n, d->muR.v, L, dc, dm, count, stride, s, mask, pop);
}

// Host entry points of a key with CRT factors: the textbook CRT decryption
// (launch_crt, two half-width exponents of |p| and |q| bits) equals
// scheme.py:86-104 for every unit c, since every unit of Z_(n^2)^* is an
// encryption (1 + n)^m r^n (the key passed chx_host_paillier_check). Its
// finish lists every non-unit (p | c or q | c); only those items are
// decrypted again by paillier_decrypt_impl, whose results equal the formula
// for every c, and their outputs replaced. When more than a quarter of the
// batch is listed the whole batch takes the exact form instead (one pass).
// Returns false (nothing launched) for a key without CRT factors; else the
// listed item indices in `redo` (empty when the batch was fully decrypted).
static bool decrypt_units_first(chx_pkey* k, const uint32_t* dc, uint32_t* dm, size_t count, size_t stride,
                                const uint32_t* mask, uint32_t* pop, std::vector<uint32_t>& redo) {
  redo.clear();
  if (!k->crt) return false;
  chx_ctx* ctx = k->ctx;
  ctx->slot[S_AUX].ensure((count + 1) * sizeof(uint32_t));
  uint32_t* bad = ctx->slot[S_AUX].as<uint32_t>();
  hip_check(hipMemsetAsync(bad, 0, sizeof(uint32_t), ctx->stream), "hipMemsetAsync");
  launch_crt(k, dc, dm, count, stride, bad, mask, pop);  // dm: rows [0, H)
  uint32_t nbad = 0;
  hip_check(hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, ctx->stream), "D2H");
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if (nbad > count) fail(CHX_EDEVICE, "non-unit list overflow");
  if ((size_t)nbad * 4 > count) {  // many non-units: one exact pass over the batch
    paillier_decrypt_impl(k, dc, dm, count, stride, mask, pop);
    return true;
  }
  if (dm)  // m < n: the upper half of the W-limb rows is zero
    hip_check(hipMemsetAsync(dm + (size_t)k->H * stride, 0, (size_t)k->H * stride * sizeof(uint32_t), ctx->stream),
              "hipMemsetAsync");
  if (nbad) {
    redo.resize(nbad);
    hip_check(hipMemcpyAsync(redo.data(), bad + 1, (size_t)nbad * 4, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  }
  return true;
}

// Exact decryption of the listed items of a host batch c (W limbs each), after
// the main pass's results are on the host: m rows (W limbs) or odd-bit counts
// (pop, with the device mask) of those items are replaced.
static void redo_exact(chx_pkey* k, const uint32_t* c, const std::vector<uint32_t>& redo, uint32_t* m,
                       const uint32_t* dmask, uint32_t* pop) {
  if (redo.empty()) return;
  chx_ctx* ctx = k->ctx;
  const int L = k->W;
  const size_t nb = redo.size(), st = batch_stride(nb);
  std::vector<uint32_t> sub(nb * L);
  for (size_t i = 0; i < nb; ++i) memcpy(&sub[i * L], c + (size_t)redo[i] * L, L * sizeof(uint32_t));
  uint32_t* dc = upload_il(ctx, S_IN0, sub.data(), nb, L, st);
  if (m) {
    ctx->slot[S_OUT].ensure(st * L * 4);
    paillier_decrypt_impl(k, dc, ctx->slot[S_OUT].as<uint32_t>(), nb, st, nullptr, nullptr);
    download_il(ctx, ctx->slot[S_OUT].as<uint32_t>(), sub.data(), nb, L, st);
    for (size_t i = 0; i < nb; ++i) memcpy(m + (size_t)redo[i] * L, &sub[i * L], L * sizeof(uint32_t));
  } else {
    ctx->slot[S_OUT].ensure(nb * 4);
    paillier_decrypt_impl(k, dc, nullptr, nb, st, dmask, ctx->slot[S_OUT].as<uint32_t>());
    hip_check(hipMemcpyAsync(sub.data(), ctx->slot[S_OUT].p, nb * 4, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    for (size_t i = 0; i < nb; ++i) pop[redo[i]] = sub[i];
  }
}

int chx_paillier_decrypt(chx_ctx* ctx, const chx_pkey* kc, const uint32_t* c, uint32_t* m, size_t count) {
  chx_pkey* k = const_cast<chx_pkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !c || !m) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (count == 0) return;
    const int L = k->W;
    check_below(c, count, L, k->n2.N, "c");
    const size_t stride = batch_stride(count);
    uint32_t* dc = upload_il(ctx, S_IN0, c, count, L, stride);
    ctx->slot[S_OUT].ensure(stride * L * 4);
    std::vector<uint32_t> redo;
    if (!decrypt_units_first(k, dc, ctx->slot[S_OUT].as<uint32_t>(), count, stride, nullptr, nullptr, redo))
      paillier_decrypt_impl(k, dc, ctx->slot[S_OUT].as<uint32_t>(), count, stride, nullptr, nullptr);
    download_il(ctx, ctx->slot[S_OUT].as<uint32_t>(), m, count, L, stride);
    redo_exact(k, c, redo, m, nullptr, nullptr);
  });
}

int chx_paillier_decrypt_dev(chx_ctx* ctx, const chx_pkey* kc, const uint32_t* c, uint32_t* m, size_t count,
                             size_t stride) {
  chx_pkey* k = const_cast<chx_pkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !c || !m) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (count == 0) return;
    check_stride(count, stride);
    paillier_decrypt_impl(k, c, m, count, stride, nullptr, nullptr);
  });
}

int chx_paillier_decrypt_crt(chx_ctx* ctx, const chx_pkey* kc, const uint32_t* c, uint32_t* m, size_t count) {
  chx_pkey* k = const_cast<chx_pkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !c || !m) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (count == 0) return;
    const int L = k->W;
    check_below(c, count, L, k->n2.N, "c");
    const size_t stride = batch_stride(count);
    uint32_t* dc = upload_il(ctx, S_IN0, c, count, L, stride);
    ctx->slot[S_OUT].ensure(stride * k->H * 4);
    launch_crt(k, dc, ctx->slot[S_OUT].as<uint32_t>(), count, stride);
    download_il(ctx, ctx->slot[S_OUT].as<uint32_t>(), m, count, k->H, stride);
  });
}

int chx_paillier_decrypt_crt_dev(chx_ctx* ctx, const chx_pkey* kc, const uint32_t* c, uint32_t* m,
                                 size_t count, size_t stride) {
  chx_pkey* k = const_cast<chx_pkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !c || !m) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (count == 0) return;
    check_stride(count, stride);
    launch_crt(k, c, m, count, stride);
  });
}

int chx_paillier_decrypt_crt_noalloc(chx_ctx* ctx, const chx_pkey* kc, const uint32_t* c, uint32_t* m,
                                     size_t count, size_t stride) {
  chx_pkey* k = const_cast<chx_pkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !c || !m) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (!k->crt) fail(CHX_EINVAL, "key was created without CRT factors");
    if (count == 0) return;
    check_stride(count, stride);
    const FixedScratch fs{scratch_ptr(ctx->device), kNoAllocScratchBytes / sizeof(uint32_t)};
    // HIP-event timing would create events: off for this call
    struct TimingOff {
      chx_ctx* c;
      bool was;
      ~TimingOff() { c->timing = was; }
    } off{ctx, ctx->timing};
    ctx->timing = false;
    launch_crt(k, c, m, count, stride, nullptr, nullptr, nullptr, &fs);
  });
}

int chx_paillier_decode_hamming(chx_ctx* ctx, const chx_pkey* kc, const uint32_t* c, const uint32_t* mask,
                                uint32_t* counts, size_t count) {
  chx_pkey* k = const_cast<chx_pkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !c || !mask || !counts) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (count == 0) return;
    const int L = k->W;
    check_below(c, count, L, k->n2.N, "c");
    const size_t stride = batch_stride(count);
    uint32_t* dc = upload_il(ctx, S_IN0, c, count, L, stride);
    ctx->slot[S_IN1].ensure(L * 4);
    hip_check(hipMemcpyAsync(ctx->slot[S_IN1].p, mask, L * 4, hipMemcpyHostToDevice, ctx->stream), "H2D");
    ctx->slot[S_OUT].ensure(count * 4);
    std::vector<uint32_t> redo;
    if (!decrypt_units_first(k, dc, nullptr, count, stride, ctx->slot[S_IN1].as<uint32_t>(),
                             ctx->slot[S_OUT].as<uint32_t>(), redo))
      paillier_decrypt_impl(k, dc, nullptr, count, stride, ctx->slot[S_IN1].as<uint32_t>(),
                            ctx->slot[S_OUT].as<uint32_t>());
    hip_check(hipMemcpyAsync(counts, ctx->slot[S_OUT].p, count * 4, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    redo_exact(k, c, redo, nullptr, ctx->slot[S_IN1].as<uint32_t>(), counts);
  });
}

// ---------------------------------------------------------------------------
// lookup
// ---------------------------------------------------------------------------
int chx_lookup_key_create(chx_ctx* ctx, const uint32_t* n, int n_limbs, const uint32_t* a, int a_limbs,
                          const uint32_t* g_a_inv, const uint32_t* g_table, int chunks,
                          const uint32_t* noise_table, int noise_entries, chx_lkey** out) {
  return with_ctx(ctx, [&] {
    if (!n || !a || !g_a_inv || !out) fail(CHX_EINVAL, "null argument");
    const bool tables = chunks > 0;
    if (n_limbs <= 0 || n_limbs > 64 || chunks < 0 || (tables && (!g_table || !noise_table)) ||
        (tables && (noise_entries <= 0 || noise_entries > 256)))
      fail(CHX_EINVAL, "bad sizes");
    BN nb = host::bn_from(n, n_limbs, n_limbs);
    const int W = pick_width(2 * host::bn_bits(nb));
    std::unique_ptr<chx_lkey> k(new chx_lkey());
    k->ctx = ctx;
    k->W = W;
    k->chunks = chunks;
    k->noise = noise_entries;
    BN nW = host::bn_from(n, n_limbs, W);
    BN n2 = host::bn_sqr_lo(nW);
    k->n2 = host::mont_consts(n2);
    k->n = host::mont_consts(nW);
    k->e_a = host::bn_from(a, a_limbs, std::max(a_limbs, 1));
    try {  // a failure below frees what was already allocated on the device
      ModDev md;
      fill_moddev(md, k->n2);
      k->d_n2 = dev_upload(&md, 1, ctx->stream);
      fill_moddev(md, k->n);
      k->d_n = dev_upload(&md, 1, ctx->stream);
      k->lz_n2 = make_lazy_set(k->n2, ctx->stream);
      k->gai = host::bn_mod(host::bn_from(g_a_inv, n_limbs, W), nW);
      BN gar = host::to_mont_host(host::bn_from(g_a_inv, n_limbs, W), k->n);
      k->d_gar = dev_upload(gar.data(), W, ctx->stream);
      if (tables) {
        // tables to the lazy Montgomery domain on the device (once per key)
        k->d_gtab = lazy_table_upload(ctx, k->lz_n2, g_table, (size_t)chunks * 256);
        k->d_ntab = lazy_table_upload(ctx, k->lz_n2, noise_table, (size_t)noise_entries);
        k->d_nptab = lazy_noise_pairs(ctx, k->lz_n2, k->d_ntab, noise_entries);
        k->d_ntab_plain = lazy_plain_copy(ctx, k->lz_n2, k->d_ntab, (size_t)noise_entries);
        k->d_nptab_plain = lazy_plain_copy(ctx, k->lz_n2, k->d_nptab, (size_t)noise_entries * noise_entries);
      }
    } catch (...) {
      (void)chx_lookup_key_destroy(k.release());
      throw;
    }
    *out = k.release();
  });
}

int chx_lookup_key_destroy(chx_lkey* k) {
  if (!k) return CHX_OK;
  int rc = guarded([&] {
    hip_check(hipSetDevice(k->ctx->device), "hipSetDevice");
    for (void* p : {(void*)k->d_n2, (void*)k->d_n, (void*)k->d_gar, (void*)k->d_gtab, (void*)k->d_ntab,
                    (void*)k->d_n3tab, (void*)k->d_nptab, (void*)k->d_ntab_plain, (void*)k->d_nptab_plain})
      if (p) (void)hipFree(p);
    for (auto& t : k->group)
      if (t.d) (void)hipFree(t.d);
    if (k->d_crt) (void)hipFree(k->d_crt);
    if (k->d_lcrt) (void)hipFree(k->d_lcrt);
    free_lazy_set(k->lz_n2);
    free_lazy_set(k->lz_p2);
    free_lazy_set(k->lz_q2);
    for (auto& s : k->scheds)
      if (s.second.d) (void)hipFree(s.second.d);
  });
  delete k;
  return rc;
}

// Constants of the textbook CRT route of a lookup key (k_lookup_crt_fast),
// uploaded; nullptr (the route stays off, every item takes the exact
// recombination) when p or q does not fit the quarter width or has no
// inverse modulo the other.
LookupCrtDev* make_lookup_crt(const BN& nW, const BN& pb, const BN& qb, const BN& gai, int H, hipStream_t st) {
  const int Hh = H / 2;
  for (int i = Hh; i < H; ++i)
    if (pb[i] || qb[i]) return nullptr;
  BN ph = host::bn_from(pb.data(), Hh, Hh), qh = host::bn_from(qb.data(), Hh, Hh);
  host::MontConsts cph = host::mont_consts(ph), cqh = host::mont_consts(qh);
  BN nH = host::bn_from(nW.data(), H, H);
  host::MontConsts cn = host::mont_consts(nH);
  // inverses by Fermat (p, q prime), each checked
  BN two(Hh, 0u);
  two[0] = 2;
  BN pm2 = ph, qm2 = qh;
  host::bn_sub(pm2, two);
  host::bn_sub(qm2, two);
  BN qinv = host::bn_powmod(host::bn_mod(qh, ph), pm2, cph);  // q^-1 mod p
  BN pinv = host::bn_powmod(host::bn_mod(ph, qh), qm2, cqh);  // p^-1 mod q
  BN one(Hh, 0u);
  one[0] = 1;
  if (host::bn_cmp(host::bn_mulmod(qinv, host::bn_mod(qh, ph), ph), one) != 0 ||
      host::bn_cmp(host::bn_mulmod(pinv, host::bn_mod(ph, qh), qh), one) != 0)
    return nullptr;
  LookupCrtDev d;
  memset(&d, 0, sizeof d);
  fill_moddev(d.nh, cn);
  fill_moddev(d.ph, cph);
  fill_moddev(d.qh, cqh);
  const BN hp = host::bn_mulmod(qinv, host::bn_mod(gai, ph), ph);
  const BN hq = host::bn_mulmod(pinv, host::bn_mod(gai, qh), qh);
  fill_val(d.hpR, host::to_mont_host(hp, cph));
  fill_val(d.hqR, host::to_mont_host(hq, cqh));
  fill_val(d.qinvR, host::to_mont_host(qinv, cph));
  fill_val(d.qRn, host::to_mont_host(host::bn_from(qb.data(), H, H), cn));
  fill_val(d.pR3, host::to_mont_host(cph.R2, cph));  // R^2 R = R^3 mod p
  fill_val(d.qR3, host::to_mont_host(cqh.R2, cqh));
  return dev_upload(&d, 1, st);
}

int chx_lookup_key_set_crt(chx_ctx* ctx, chx_lkey* k, const uint32_t* p, const uint32_t* q, int half_limbs,
                           const uint32_t* ep, const uint32_t* eq) {
  return with_ctx(ctx, [&] {
    if (!k || !p || !q || !ep || !eq) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (k->crt) fail(CHX_EINVAL, "key already has CRT factors");
    const int W = k->W, H = W / 2;
    if (H != 32 && H != 64) fail(CHX_EUNSUPPORTED, "CRT needs 32- or 64-limb halves (width %d)", W);
    if (half_limbs <= 0 || half_limbs > H) fail(CHX_EINVAL, "factors wider than half the width");
    BN pb = host::bn_from(p, half_limbs, H), qb = host::bn_from(q, half_limbs, H);
    // the halves exponentiate modulo p^2 and q^2 at H limbs
    if (host::bn_bits(pb) > 16 * H || host::bn_bits(qb) > 16 * H)
      fail(CHX_EINVAL, "p and q must be below 2^%d for %d-limb CRT halves", 16 * H, H);
    if (host::bn_bits(pb) < 2 || host::bn_bits(qb) < 2 || !(pb[0] & 1u) || !(qb[0] & 1u))
      fail(CHX_EINVAL, "p and q must be odd and above 1");
    BN pq(W, 0u);  // p * q at W limbs
    for (int i = 0; i < H; ++i) {
      uint64_t c = 0;
      for (int j = 0; j < H; ++j) {
        const uint64_t t = (uint64_t)pb[i] * qb[j] + pq[i + j] + c;
        pq[i + j] = (uint32_t)t;
        c = t >> 32;
      }
      pq[i + H] = (uint32_t)c;
    }
    if (host::bn_cmp(pq, k->n.N) != 0) fail(CHX_EINVAL, "p * q != n");
    BN p2 = host::bn_sqr_lo(pb), q2 = host::bn_sqr_lo(qb);
    uint32_t* er = upload_idempotents(ctx, k->n2, p2, q2, ep, eq);
    k->d_crt = er;
    try {  // a failure below leaves the key without CRT factors
      k->lz_p2 = make_lazy_set(host::mont_consts(p2), ctx->stream);
      k->lz_q2 = make_lazy_set(host::mont_consts(q2), ctx->stream);
      k->d_lcrt = make_lookup_crt(k->n.N, pb, qb, k->gai, H, ctx->stream);
    } catch (...) {
      free_lazy_set(k->lz_p2);
      free_lazy_set(k->lz_q2);
      if (k->d_crt) (void)hipFree(k->d_crt);
      k->d_crt = nullptr;
      throw;
    }
    k->crt = true;
  });
}

int chx_lookup_key_has_crt(const chx_lkey* k) { return k && k->crt ? 1 : 0; }

int chx_lookup_key_set_table_rows(chx_ctx* ctx, chx_lkey* k, int rows) {
  return with_ctx(ctx, [&] {
    if (!k) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (rows < 0) fail(CHX_EINVAL, "rows %d out of range", rows);
    if (rows == k->table_rows) return;
    // tables built for the old coverage go; the next encryption that asks
    // for one builds it for the new one
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    for (auto& t : k->group) {
      if (t.d) (void)hipFree(t.d);
      t = chx_lkey::GroupTab();
    }
    k->table_rows = rows;
  });
}

int chx_lookup_width(const chx_lkey* k) { return k ? k->W : 0; }

int chx_lookup_encrypt(chx_ctx* ctx, const chx_lkey* kc, const uint8_t* digits, const uint8_t* noise_idx,
                       int nmult, uint32_t* c, size_t count) {
  chx_lkey* k = const_cast<chx_lkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !digits || !c || (nmult > 0 && !noise_idx)) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (count == 0) return;
    if (nmult < 0 || nmult > 64) fail(CHX_EINVAL, "nmult out of range");
    for (size_t i = 0; i < count * (size_t)nmult; ++i)
      if (noise_idx[i] >= k->noise) fail(CHX_EINVAL, "noise index %u out of range", noise_idx[i]);
    const int L = k->W;
    const size_t stride = batch_stride(count);
    const size_t dbytes = count * (size_t)k->chunks, nbytes = count * (size_t)nmult;
    // arbitrary plaintexts (e.g. decrypted ones being re-encrypted) have no
    // item whose digits all lie in the packing alphabet: then the row-group
    // tables would go unused, and single rows serve without building them
    bool any_packed = false;
    for (size_t i = 0; i < count && !any_packed; ++i) {
      const uint8_t* d = digits + i * (size_t)k->chunks;
      int j = 0;
      while (j < k->chunks && (d[j] & PACKED_ZERO_BITS) == 0) ++j;
      any_packed = j == k->chunks;
    }
    ctx->slot[S_IN0].ensure(dbytes + 16);
    ctx->slot[S_IN1].ensure(nbytes + 16);
    hip_check(hipMemcpyAsync(ctx->slot[S_IN0].p, digits, dbytes, hipMemcpyHostToDevice, ctx->stream), "H2D");
    if (nbytes)
      hip_check(hipMemcpyAsync(ctx->slot[S_IN1].p, noise_idx, nbytes, hipMemcpyHostToDevice, ctx->stream), "H2D");
    ctx->slot[S_OUT].ensure(stride * L * 4);
    launch_lookup_encrypt(k, ctx->slot[S_IN0].as<uint8_t>(), ctx->slot[S_IN1].as<uint8_t>(), nmult,
                          ctx->slot[S_OUT].as<uint32_t>(), count, stride, [&] { return any_packed; });
    download_il(ctx, ctx->slot[S_OUT].as<uint32_t>(), c, count, L, stride);
  });
}

int chx_lookup_encrypt_dev(chx_ctx* ctx, const chx_lkey* kc, const uint8_t* digits, const uint8_t* noise_idx,
                           int nmult, uint32_t* c, size_t count, size_t stride) {
  chx_lkey* k = const_cast<chx_lkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !digits || !c || (nmult > 0 && !noise_idx)) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (count == 0) return;
    check_stride(count, stride);
    launch_lookup_encrypt(k, digits, noise_idx, nmult, c, count, stride);
  });
}

static void lookup_decrypt_impl(chx_lkey* k, const uint32_t* dc, uint32_t* dm, size_t count, size_t stride,
                                const uint32_t* mask, uint32_t* pop) {
  chx_ctx* ctx = k->ctx;
  const int L = k->W;
  if (k->crt) {
    const ExpSched& s = get_sched_l(k, pick_window(ctx, CHX_OP_MODEXP, L / 2, host::bn_bits(k->e_a)));
    launch_exact_crt(ctx, k->lz_p2, k->lz_q2, s, k->d_n2, k->d_crt, k->d_n, k->d_gar, L, dc, dm, count, stride, mask,
                     pop, k->d_lcrt);
    return;
  }
  const ExpSched& s = get_sched_l(k, pick_window(ctx, CHX_OP_MODEXP, L, host::bn_bits(k->e_a)));
  launch_decrypt(ctx, k->lz_n2, k->d_n, k->d_gar, L, dc, dm, count, stride, s, mask, pop);
}

int chx_lookup_decrypt(chx_ctx* ctx, const chx_lkey* kc, const uint32_t* c, uint32_t* m, size_t count) {
  chx_lkey* k = const_cast<chx_lkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !c || !m) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (count == 0) return;
    const int L = k->W;
    check_below(c, count, L, k->n2.N, "c");
    const size_t stride = batch_stride(count);
    uint32_t* dc = upload_il(ctx, S_IN0, c, count, L, stride);
    ctx->slot[S_OUT].ensure(stride * L * 4);
    lookup_decrypt_impl(k, dc, ctx->slot[S_OUT].as<uint32_t>(), count, stride, nullptr, nullptr);
    download_il(ctx, ctx->slot[S_OUT].as<uint32_t>(), m, count, L, stride);
  });
}

int chx_lookup_decode_hamming(chx_ctx* ctx, const chx_lkey* kc, const uint32_t* c, const uint32_t* mask,
                              uint32_t* counts, size_t count) {
  chx_lkey* k = const_cast<chx_lkey*>(kc);
  return with_ctx(ctx, [&] {
    if (!k || !c || !mask || !counts) fail(CHX_EINVAL, "null argument");
    if (k->ctx != ctx) fail(CHX_EINVAL, "key belongs to another context");
    if (count == 0) return;
    const int L = k->W;
    check_below(c, count, L, k->n2.N, "c");
    const size_t stride = batch_stride(count);
    uint32_t* dc = upload_il(ctx, S_IN0, c, count, L, stride);
    ctx->slot[S_IN1].ensure(L * 4);
    hip_check(hipMemcpyAsync(ctx->slot[S_IN1].p, mask, L * 4, hipMemcpyHostToDevice, ctx->stream), "H2D");
    ctx->slot[S_OUT].ensure(count * 4);
    lookup_decrypt_impl(k, dc, nullptr, count, stride, ctx->slot[S_IN1].as<uint32_t>(),
                        ctx->slot[S_OUT].as<uint32_t>());
    hip_check(hipMemcpyAsync(counts, ctx->slot[S_OUT].p, count * 4, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  });
}

// Decrypt dc under `from` into a scratch buffer, take the plaintexts' base-2^8
// digits on the device and encrypt them under `to` into out (same stride).
static void lookup_reencrypt_impl(chx_ctx* ctx, chx_lkey* from, chx_lkey* to, const uint32_t* dc,
                                  const uint8_t* nidx, int nmult, uint32_t* out, size_t count, size_t stride) {
  if (from->ctx != ctx || to->ctx != ctx) fail(CHX_EINVAL, "both keys must belong to this context");
  if (!to->d_gtab || !to->d_ntab) fail(CHX_EINVAL, "target key was created without tables (decrypt-only)");
  // m < n_from: every digit of a plaintext has a row of the target's g table
  const int mbits = host::bn_bits(from->n.N);
  if (mbits > 8 * to->chunks)
    fail(CHX_EINVAL, "plaintexts of a %d-bit n have more base-256 digits than the target's %d message chunks", mbits,
         to->chunks);
  const int La = from->W;  // m < n fits the low La/2 limbs (32 La >= bits(n^2) >= 2 bits(n) - 1)
  ctx->slot[S_MID1].ensure(stride * La * sizeof(uint32_t));
  uint32_t* dm = ctx->slot[S_MID1].as<uint32_t>();
  lookup_decrypt_impl(from, dc, dm, count, stride, nullptr, nullptr);
  ctx->slot[S_TMP].ensure(count * (size_t)to->chunks + 16);
  ctx->slot[S_AUX].ensure(16);
  uint32_t* flag = ctx->slot[S_AUX].as<uint32_t>();
  uint32_t token = ++ctx->packed_token;
  if (token == 0) token = ctx->packed_token = 1;
  record_start(ctx, "il_digits");
  hipLaunchKernelGGL(k_il_digits, grid_for(count, 1), dim3(BLOCK_THREADS), 0, ctx->stream, dm, La / 2,
                     (int)stride, (int)count, ctx->slot[S_TMP].as<uint8_t>(), to->chunks, flag, token);
  record_stop(ctx);
  check_launch("k_il_digits");
  // read back only when a row-group table would be built: a stale flag can
  // only cause a build, never a different result
  auto any_packed = [&] {
    uint32_t h = 0;
    hip_check(hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return h == token;
  };
  launch_lookup_encrypt(to, ctx->slot[S_TMP].as<uint8_t>(), nidx, nmult, out, count, stride, any_packed);
}

int chx_lookup_reencrypt(chx_ctx* ctx, const chx_lkey* fc, const chx_lkey* tc, const uint32_t* c,
                         const uint8_t* noise_idx, int nmult, uint32_t* out, size_t count) {
  chx_lkey* from = const_cast<chx_lkey*>(fc);
  chx_lkey* to = const_cast<chx_lkey*>(tc);
  return with_ctx(ctx, [&] {
    if (!from || !to || !c || !out || (nmult > 0 && !noise_idx)) fail(CHX_EINVAL, "null argument");
    if (count == 0) return;
    if (nmult < 0 || nmult > 64) fail(CHX_EINVAL, "nmult out of range");
    for (size_t i = 0; i < count * (size_t)nmult; ++i)
      if (noise_idx[i] >= to->noise) fail(CHX_EINVAL, "noise index %u out of range", noise_idx[i]);
    const int La = from->W, Lb = to->W;
    check_below(c, count, La, from->n2.N, "c");
    const size_t stride = batch_stride(count);
    uint32_t* dc = upload_il(ctx, S_IN0, c, count, La, stride);
    const size_t nbytes = count * (size_t)nmult;
    ctx->slot[S_IN1].ensure(nbytes + 16);
    if (nbytes)
      hip_check(hipMemcpyAsync(ctx->slot[S_IN1].p, noise_idx, nbytes, hipMemcpyHostToDevice, ctx->stream), "H2D");
    ctx->slot[S_OUT].ensure(stride * Lb * 4);
    lookup_reencrypt_impl(ctx, from, to, dc, ctx->slot[S_IN1].as<uint8_t>(), nmult, ctx->slot[S_OUT].as<uint32_t>(),
                          count, stride);
    download_il(ctx, ctx->slot[S_OUT].as<uint32_t>(), out, count, Lb, stride);
  });
}

int chx_lookup_reencrypt_dev(chx_ctx* ctx, const chx_lkey* fc, const chx_lkey* tc, const uint32_t* c,
                             const uint8_t* noise_idx, int nmult, uint32_t* out, size_t count, size_t stride) {
  chx_lkey* from = const_cast<chx_lkey*>(fc);
  chx_lkey* to = const_cast<chx_lkey*>(tc);
  return with_ctx(ctx, [&] {
    if (!from || !to || !c || !out || (nmult > 0 && !noise_idx)) fail(CHX_EINVAL, "null argument");
    if (count == 0) return;
    if (nmult < 0 || nmult > 64) fail(CHX_EINVAL, "nmult out of range");
    check_stride(count, stride);
    lookup_reencrypt_impl(ctx, from, to, c, noise_idx, nmult, out, count, stride);
  });
}

// ---------------------------------------------------------------------------
// layout + measurement
// ---------------------------------------------------------------------------
int chx_aos_to_il_dev(chx_ctx* ctx, const uint32_t* aos, uint32_t* il, size_t count, int limbs, size_t stride) {
  return with_ctx(ctx, [&] {
    if (!aos || !il || limbs <= 0 || limbs > 4096) fail(CHX_EINVAL, "bad argument");
    if (stride < count || stride > kMaxStride) fail(CHX_EINVAL, "stride %zu below count or above the maximum", stride);
    if (count) aos_to_il(ctx, aos, il, count, limbs, stride);
  });
}

int chx_il_to_aos_dev(chx_ctx* ctx, const uint32_t* il, uint32_t* aos, size_t count, int limbs, size_t stride) {
  return with_ctx(ctx, [&] {
    if (!aos || !il || limbs <= 0 || limbs > 4096) fail(CHX_EINVAL, "bad argument");
    if (stride < count || stride > kMaxStride) fail(CHX_EINVAL, "stride %zu below count or above the maximum", stride);
    if (count) il_to_aos(ctx, il, aos, count, limbs, stride);
  });
}

int chx_probe_rate(chx_ctx* ctx, int kind, int blocks, int iters, double* seconds, double* ops) {
  return with_ctx(ctx, [&] {
    if (blocks <= 0 || iters <= 0 || !seconds || !ops || kind < 0 || kind > 7) fail(CHX_EINVAL, "bad argument");
    ctx->slot[S_AUX].ensure((size_t)blocks * 256 * 8);
    u64* sink = ctx->slot[S_AUX].as<u64>();
    auto launch = [&](int it) {
      switch (kind) {
        case 0: hipLaunchKernelGGL(k_probe<0>, dim3(blocks), dim3(256), 0, ctx->stream, sink, it, 7u); break;
        case 1: hipLaunchKernelGGL(k_probe<1>, dim3(blocks), dim3(256), 0, ctx->stream, sink, it, 7u); break;
        case 2: hipLaunchKernelGGL(k_probe<2>, dim3(blocks), dim3(256), 0, ctx->stream, sink, it, 7u); break;
        case 3: hipLaunchKernelGGL(k_probe<3>, dim3(blocks), dim3(256), 0, ctx->stream, sink, it, 7u); break;
        case 4: hipLaunchKernelGGL(k_probe<4>, dim3(blocks), dim3(256), 0, ctx->stream, sink, it, 7u); break;
        case 5: hipLaunchKernelGGL(k_probe<5>, dim3(blocks), dim3(256), 0, ctx->stream, sink, it, 7u); break;
        case 6: hipLaunchKernelGGL(k_probe<6>, dim3(blocks), dim3(256), 0, ctx->stream, sink, it, 7u); break;
        default: hipLaunchKernelGGL(k_probe<7>, dim3(blocks), dim3(256), 0, ctx->stream, sink, it, 7u); break;
      }
    };
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "hipEventCreate");
    hip_check(hipEventCreate(&b), "hipEventCreate");
    launch(8);
    hip_check(hipEventRecord(a, ctx->stream), "hipEventRecord");
    launch(iters);
    hip_check(hipEventRecord(b, ctx->stream), "hipEventRecord");
    check_launch("k_probe");
    hip_check(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *seconds = ms * 1e-3;
    *ops = (double)blocks * 256.0 * (double)iters * 16.0;
  });
}

int chx_probe_copy(chx_ctx* ctx, const uint32_t* src, uint32_t* dst, size_t count, size_t stride, int limbs,
                   int tpi) {
  return with_ctx(ctx, [&] {
    if (!src || !dst) fail(CHX_EINVAL, "null argument");
    if (count == 0) return;
    check_stride(count, stride);
    bool done = false;
#define CHX_COPY_CASE(L_, T_)                                                                                  \
    if (limbs == L_ && tpi == T_) {                                                                            \
      hipLaunchKernelGGL((k_copy_groups<T_, L_>), grid_for(count, T_), dim3(BLOCK_THREADS), 0, ctx->stream, src, \
                         dst, (int)count, (int)stride);                                                        \
      done = true;                                                                                             \
    }
    CHX_COPY_CASE(128, 1)  // one lane per item: the row pattern of the fixed and pointwise kernels' loads
    CHX_COPY_CASE(64, 4)
    CHX_COPY_CASE(64, 8)
    CHX_COPY_CASE(128, 8)
    CHX_COPY_CASE(128, 16)
#undef CHX_COPY_CASE
    if (!done) fail(CHX_EUNSUPPORTED, "copy probe built for (64, 4|8) and (128, 8|16) only");
    check_launch("k_copy_groups");
  });
}

int chx_probe_mad(chx_ctx* ctx, int blocks, int iters, double* seconds, double* macs) {
  return chx_probe_rate(ctx, 0, blocks, iters, seconds, macs);
}

int chx_ctx_set_timing(chx_ctx* ctx, int enable) {
  return with_ctx(ctx, [&] { ctx->timing = enable != 0; });
}

int chx_ctx_kernel_time(chx_ctx* ctx, const char* name, double* total_ms, int* launches) {
  return with_ctx(ctx, [&] {
    if (!name || !total_ms || !launches) fail(CHX_EINVAL, "null argument");
    flush_timing(ctx);
    auto it = ctx->stats.find(name);
    *total_ms = it == ctx->stats.end() ? 0.0 : it->second.first;
    *launches = it == ctx->stats.end() ? 0 : it->second.second;
  });
}

int chx_ctx_reset_timing(chx_ctx* ctx) {
  return with_ctx(ctx, [&] {
    flush_timing(ctx);
    ctx->stats.clear();
  });
}

}  // extern "C"
