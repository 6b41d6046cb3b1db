/*
 * chx.h -- C ABI of the MI355X (gfx950) Paillier batched-modexp library.
 *
 * This is the drop-in boundary for the reference's two pybind11 extensions,
 *   cuhepy.paillier._gpu_ext.paillier_GPU_client        (class PaillierGPUClient,
 *       src/cuhepy/paillier/_gpu_ext/paillier_GPU_client.cu:1136-1165)
 *   cuhepy.paillier._lookup_gpu_ext.paillier_GPU_lookup_client
 *       (class PaillierGPULookupClient, .../paillier_GPU_lookup_client.cu:2673-2721)
 * The Python classes of the same names in cuhepy_amd.paillier are rebuilt on
 * top of these functions through ctypes; each function below names the
 * reference method or kernel whose work it performs.
 *
 * Conventions
 *  - Bignums are little-endian arrays of 32-bit limbs. Host-buffer entry points
 *    take "AoS" batches: item i occupies limbs [i*limbs, (i+1)*limbs), the
 *    layout of the reference's cgbn_mem_t arrays (cu:310-313).
 *  - *_dev entry points take device pointers in the limb-interleaved layout:
 *    limb j of item i at [j*stride + i], stride >= count rounded up to 64.
 *    They enqueue on the context's stream and return without synchronising.
 *  - Every function returns CHX_OK (0) or an error code; chx_last_error()
 *    returns the message of the calling thread's last failure. Inputs are
 *    validated (values >= modulus are rejected, not reduced or truncated).
 *  - Thread safety: calls on one context are serialised by a per-context lock;
 *    different contexts run concurrently. ctypes releases the GIL around every
 *    call.
 */
#ifndef CHX_H
#define CHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chx_ctx chx_ctx;   /* device context: stream, scratch, tuning */
typedef struct chx_mod chx_mod;   /* one odd modulus + Montgomery constants */
typedef struct chx_pkey chx_pkey; /* Paillier key (n, n^2, decrypt and CRT constants) */
typedef struct chx_lkey chx_lkey; /* Paillier-lookup key (tables in HBM) */

enum {
  CHX_OK = 0,
  CHX_EINVAL = 1,      /* bad argument or value out of range */
  CHX_EDEVICE = 2,     /* HIP runtime error */
  CHX_ENOMEM = 3,      /* device allocation failed */
  CHX_EUNSUPPORTED = 4 /* width / lane-group combination not built */
};

/* operation ids for chx_ctx_set_tuning / chx_ctx_get_tuning */
enum {
  CHX_OP_MODEXP = 0, /* fixed-exponent modexp, encrypt, decrypt, CRT */
  CHX_OP_MULMOD = 1, /* pointwise / broadcast mulmod */
  CHX_OP_LOOKUP = 2, /* lookup-table encryption */
  CHX_OP_RAGGED = 3  /* per-item exponent modexp */
};

const char* chx_version(void);
/* copy the calling thread's last error message; returns its full length */
int chx_last_error(char* buf, size_t size);

/* ---- context (no reference counterpart: the reference uses the default stream
 *      and cudaMalloc/cudaFree on every call, cu:917-961) ---- */
int chx_ctx_create(int device, chx_ctx** out);
int chx_ctx_destroy(chx_ctx* ctx);
/* use an external hipStream_t (e.g. torch's); NULL restores the owned stream */
int chx_ctx_set_stream(chx_ctx* ctx, void* stream);
void* chx_ctx_get_stream(chx_ctx* ctx);
int chx_ctx_sync(chx_ctx* ctx);
/* lanes per bignum (tpi) and window width for `op` at `limbs`; 0 = automatic.
 * For CHX_OP_LOOKUP the window is the number of 8-bit table rows one product
 * covers: 1 = single rows, 2 = row pairs, 3 = row triples, 4 = row quads,
 * 5 = row quints with three noise factors per product (16^R-entry tables per
 * R rows, built on first use; the next smaller group when the memory for one
 * cannot be allocated; quints take 604 MB per 5 rows and 9.7 GB of noise
 * triples at a 4096-bit n^2, see chx_lookup_key_set_table_rows); groups are
 * used for waves whose digits are all in the bit-packing alphabet; 0 = 4.
 * For CHX_OP_MULMOD
 * the window selects the form of the pointwise multiply of batches of 8,192
 * or more: 0 or 1 = one product and one fold (k_pw_mulmod), 2 = two
 * Montgomery products (k_lazy_mulmod); a forced tpi selects the canonical
 * engine at any size. Results are identical. */
int chx_ctx_set_tuning(chx_ctx* ctx, int op, int limbs, int tpi, int window);
/* lanes-per-bignum values built for `limbs` (returns count, fills up to cap);
 * op CHX_OP_MODEXP lists the exponentiation engine's values, other ops the
 * canonical engine's */
int chx_supported_tpi(int limbs, int* tpis, int cap);
int chx_supported_tpi_op(int op, int limbs, int* tpis, int cap);
/* bytes of device scratch currently held by the context */
size_t chx_ctx_scratch_bytes(chx_ctx* ctx);

/* ---- moduli ---- */
/* limbs in {32, 64, 128}; n odd, n < 2^(32*limbs) */
int chx_mod_create(chx_ctx* ctx, const uint32_t* n, int limbs, chx_mod** out);
int chx_mod_destroy(chx_mod* m);
int chx_mod_limbs(const chx_mod* m);

/* out[i] = a[i]*b[i] mod n for any a, b below 2^(32*limbs) -- encode_hamming_server /
 * gpu_pointwise_mulmod_hex
 * (cu:672-732, modmul_pointwise_kernel cu:430-454; lcu:1505-1561) */
int chx_mulmod(chx_ctx* ctx, const chx_mod* m, const uint32_t* a, const uint32_t* b, uint32_t* out,
               size_t count);
int chx_mulmod_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* a, const uint32_t* b, uint32_t* out,
                   size_t count, size_t stride);
/* ---- served Hamming evaluation (many queries against one resident index) ----
 * encode_hamming_server multiplies a query ciphertext into each index ciphertext
 * mod n^2 (paillier_GPU_client.cu:986-991 -> gpu_pointwise_mulmod_hex cu:672-732;
 * paillier_GPU_lookup_client.cu:2312-2329 -> gpu_pointwise_mulmod_bigint lcu:1506).
 * A query is first turned into an operand of chx_mulmod_operand_words(m) words
 * (y: device, count x limbs, any value below 2^(32 limbs); ops: device, count x
 * words). chx_mulmod_multi_dev then computes, in one launch,
 *   outs[q][i] = x[i] * y_q mod n   for i in [begin, begin + count), q < nq <= 32
 * with x interleaved [limbs][stride] and every outs[q] interleaved
 * [limbs][out_stride], indexed by the global entry i. ops and outs are host
 * arrays of nq device pointers. Results are canonical (below n). The operand is
 * the query's table y * 2^(RB (i + 2)) mod n for every position i of the
 * multiply's radix (144 positions of 29 bits at a 4096-bit n: 92 KB), so the
 * multiply needs no reduction rows. */
int chx_mulmod_operand_words(const chx_mod* m);
int chx_mulmod_operands_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* y, uint32_t* ops, size_t count);
/* the same from host limbs (count x limbs): stream-ordered, no synchronisation */
int chx_mulmod_operands(chx_ctx* ctx, const chx_mod* m, const uint32_t* y, uint32_t* ops, size_t count);
int chx_mulmod_multi_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* x, size_t stride, size_t begin,
                         size_t count, int nq, const uint32_t* const* ops, uint32_t* const* outs,
                         size_t out_stride);
/* chx_mulmod_multi_dev that also reports when each query is complete: done
 * (device, nq + 1 words, zeroed by the call on the stream) counts each
 * query's finished tiles and holds the head of a queue of (query, tile)
 * units, and stamps[q] (device, nq u64, zeroed by the call) receives the
 * device clock (chx_ctx_wall_clock_hz) when query q's last tile finished
 * (the largest of its tiles' finishing clocks). The workgroups
 * take the units from the queue in order, query by query, so the queries
 * complete one after another, each overlapping only the start of the next
 * (no reference counterpart: the reference serves one query per call). */
int chx_mulmod_multi_stamped_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* x, size_t stride, size_t begin,
                                 size_t count, int nq, const uint32_t* const* ops, uint32_t* const* outs,
                                 size_t out_stride, uint32_t* done, uint64_t* stamps);
/* frequency of the device clock the stamps above count (Hz), measured
 * against the host's steady clock over about 20 ms */
int chx_ctx_wall_clock_hz(chx_ctx* ctx, double* hz);
/* out[i] = x[i]*y mod n with y shared (one query, canonical engine) */
int chx_mulmod_bcast_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* x, const uint32_t* y_host,
                         uint32_t* out, size_t count, size_t stride);

/* out[i] = base[i]^e mod n, one exponent for the batch -- pow_mod_mem (cu:332-360),
 * pow_mod_small_exp (lcu:728-775); gmpy2.powmod in scheme.py:81-83, 101-104 */
int chx_modexp(chx_ctx* ctx, const chx_mod* m, const uint32_t* base, size_t count, const uint32_t* e,
               int e_limbs, uint32_t* out);
int chx_modexp_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* base, uint32_t* out, size_t count,
                   size_t stride, const uint32_t* e, int e_limbs);

/* out[i] = base[i]^e[i] mod n with a per-item exponent (fixed window); exps:
 * count x e_limbs (AoS). No reference counterpart (config 4's ragged modexp;
 * the lookup noise table g_n^r_k, lookup.py:113-117, lcu:592-626).
 * _dev: exps interleaved; e_bits = bit length of the widest exponent. */
int chx_modexp_each(chx_ctx* ctx, const chx_mod* m, const uint32_t* base, size_t count, const uint32_t* exps,
                    int e_limbs, uint32_t* out);
int chx_modexp_each_dev(chx_ctx* ctx, const chx_mod* m, const uint32_t* base, const uint32_t* exps, int e_limbs,
                        int e_bits, uint32_t* out, size_t count, size_t stride);

/* ---- ragged batches (config 4; no reference counterpart) ----
 * Many sub-batches of mixed width (64 or 128 limbs), each with its own modulus
 * and per-item exponents, in ONE persistent launch: waves pull tiles from an
 * atomic work queue. The launch (queue reset by a fill kernel + the persistent
 * kernel) is captured once into a hipGraph and replayed. Device buffers are
 * interleaved; they must stay valid while the plan exists. */
typedef struct chx_ragged chx_ragged;
typedef struct {
  const chx_mod* mod;   /* modulus of the sub-batch: 64 or 128 limbs */
  const uint32_t* base; /* device [limbs][stride] */
  const uint32_t* exps; /* device [e_limbs][stride] */
  uint32_t* out;        /* device [limbs][stride] */
  uint32_t count, stride, e_limbs, e_bits;
} chx_ragged_sub;
/* variant 0: one persistent kernel per width (4 lanes per 2048-bit, 8 per
 * 4096-bit modulus), both in the one graph; variant 1: the same with 8 and 16
 * lanes; variant 2: a single persistent kernel over both widths (8 and 16
 * lanes); variant 3: a single persistent kernel over both widths (4 and 8
 * lanes). The single-kernel variants queue the costliest tiles first.
 * window 2..6 (0 = 5); waves_per_cu: persistent waves per CU (0 = 8). */
int chx_ragged_create(chx_ctx* ctx, const chx_ragged_sub* subs, int nsubs, int variant, int window,
                      int waves_per_cu, chx_ragged** out);
/* persistent launch; use_graph != 0 replays the captured hipGraph */
int chx_ragged_run(chx_ragged* rg, int use_graph);
/* the plain form: one non-persistent launch per sub-batch, no graph */
int chx_ragged_run_plain(chx_ragged* rg);
int chx_ragged_destroy(chx_ragged* rg);

/* ---- Paillier (g = n + 1) ----
 * n_limbs: limbs of n (<= 64); the working width is the smallest of 32/64/128
 * limbs that holds n^2. inv = phi^-1 mod n. When p != NULL, CRT decryption is
 * enabled: p, q (half_limbs each, p*q = n), hp = L_p(g^(p-1) mod p^2)^-1 mod p,
 * hq likewise, qinv = q^-1 mod p; such a key must pass chx_host_paillier_check
 * (CHX_EINVAL otherwise). Replaces PaillierCPU::key_gen output handling
 * (cu:183-216) and load_stringified_keys (cu:826-851). */
int chx_paillier_key_create(chx_ctx* ctx, const uint32_t* n, int n_limbs, const uint32_t* phi,
                            const uint32_t* inv, const uint32_t* p, const uint32_t* q, int half_limbs,
                            const uint32_t* hp, const uint32_t* hq, const uint32_t* qinv, chx_pkey** out);
int chx_paillier_key_destroy(chx_pkey* k);
int chx_paillier_width(const chx_pkey* k); /* limbs of n^2 */
/* Exact half-width decryption for a key created with p and q: give it the CRT
 * idempotents ep = q^2 (q^-2 mod p^2), eq = p^2 (p^-2 mod q^2) (W limbs each,
 * W = chx_paillier_width). From then on chx_paillier_decrypt(_dev) and
 * chx_paillier_decode_hamming compute c^phi mod p^2 and mod q^2 at half width
 * and recombine c^phi mod n^2 before L, so every c in [0, n^2) gives the result
 * of Paillier.decrypt (scheme.py:86-104) at half the multiply-adds of the
 * full-width form (the reference's decrypt_kernel, cu:549-587, works at full
 * width); chx_paillier_encrypt(_dev) likewise computes r^n mod p^2 and mod q^2
 * and recombines r^n mod n^2 (scheme.py:63-83, encrypt_kernel cu:458-546). Unlike chx_paillier_decrypt_crt, which follows the textbook CRT
 * formula and agrees with it on valid ciphertexts only, this form agrees on
 * non-units too. Refused (CHX_EINVAL) without CRT factors or with wrong
 * idempotents; the key then keeps the full-width form. */
int chx_paillier_key_set_idempotents(chx_ctx* ctx, chx_pkey* k, const uint32_t* ep, const uint32_t* eq);
/* 1 when chx_paillier_key_set_idempotents has succeeded on k, else 0 */
int chx_paillier_key_has_idempotents(const chx_pkey* k);

/* c[i] = (1 + m[i]*n) * r[i]^n mod n^2; m < n, 0 < r < n, gcd(r, n) = 1 is the
 * caller's draw -- encrypt_kernel (cu:458-546) / Paillier.encrypt (scheme.py:63-83).
 * m, r, c: `width` limbs per item. */
int chx_paillier_encrypt(chx_ctx* ctx, const chx_pkey* k, const uint32_t* m, const uint32_t* r,
                         uint32_t* c, size_t count);
int chx_paillier_encrypt_dev(chx_ctx* ctx, const chx_pkey* k, const uint32_t* m, const uint32_t* r,
                             uint32_t* c, size_t count, size_t stride);
/* m[i] = L(c[i]^phi mod n^2) * inv mod n -- decrypt_only_kernel (cu:549-587) /
 * Paillier.decrypt (scheme.py:86-104). c < n^2 required. For a key created
 * with p and q, this and chx_paillier_decode_hamming first run the CRT
 * decryption (equal to the formula for every unit c) and decrypt the batch
 * again in the exact form only when it holds a non-unit (p | c or q | c);
 * the results are the formula's for every c either way. */
int chx_paillier_decrypt(chx_ctx* ctx, const chx_pkey* k, const uint32_t* c, uint32_t* m, size_t count);
int chx_paillier_decrypt_dev(chx_ctx* ctx, const chx_pkey* k, const uint32_t* c, uint32_t* m, size_t count,
                             size_t stride);
/* CRT decryption (no reference counterpart; same result as chx_paillier_decrypt
 * for c coprime to n). m has width/2 limbs per item. */
int chx_paillier_decrypt_crt(chx_ctx* ctx, const chx_pkey* k, const uint32_t* c, uint32_t* m, size_t count);
int chx_paillier_decrypt_crt_dev(chx_ctx* ctx, const chx_pkey* k, const uint32_t* c, uint32_t* m,
                                 size_t count, size_t stride);
/* Allocation-free form of chx_paillier_decrypt_crt_dev (same arguments, same
 * results). No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). It calls no hipMalloc or hipFree: its
 * scratch (the CRT residues and per-item window tables, about 20 KB per
 * ciphertext at n 2048-bit, 5 KB at n 1024-bit) comes from one 64 MiB device
 * array per GPU, held until the process ends, and a batch needing more is
 * refused with CHX_EINVAL. It uses the default window (context tuning of the
 * window is ignored) and the schedules built with the key. Calls on one GPU
 * share the array: calls that overlap on one GPU (several contexts or
 * streams) race. HIP-event timing is off for the call. */
int chx_paillier_decrypt_crt_noalloc(chx_ctx* ctx, const chx_pkey* k, const uint32_t* c, uint32_t* m,
                                     size_t count, size_t stride);
/* fused decrypt + odd-bit popcount (decode_hamming_kernel cu:592-668, never
 * launched there; decode_hamming_client cu:995-1038): counts[i] =
 * popcount(m[i] & mask), mask = `width` limbs. */
int chx_paillier_decode_hamming(chx_ctx* ctx, const chx_pkey* k, const uint32_t* c, const uint32_t* mask,
                                uint32_t* counts, size_t count);

/* ---- Paillier-lookup ----
 * n_limbs: limbs of n; width = limbs of n^2 as above. a: decrypt exponent
 * (a_limbs), g_a_inv < n. Tables are plain residues mod n^2, `width` limbs per
 * entry: g_table[chunks][256] (lookup.py:104-110, lcu:550-588) and
 * noise_table[256] (lookup.py:113-117); they are converted to Montgomery form
 * on the device and kept resident in HBM (ensure_device_tables_, lcu:2625-2652).
 * The key also builds, on the device, a noise-pair table with plain-residue
 * copies (noise^2 entries; 40 MB each at a 4096-bit n^2); the row-group
 * tables (pairs, triples, quads, quints) are built by the first encryption that
 * uses them (see chx_ctx_set_tuning), quints with a noise-triple table.
 * chunks = 0 (tables NULL) creates a decrypt-only key. */
int chx_lookup_key_create(chx_ctx* ctx, const uint32_t* n, int n_limbs, const uint32_t* a, int a_limbs,
                          const uint32_t* g_a_inv, const uint32_t* g_table, int chunks,
                          const uint32_t* noise_table, int noise_entries, chx_lkey** out);
int chx_lookup_key_destroy(chx_lkey* k);
int chx_lookup_width(const chx_lkey* k);
/* Half-width decryption for a lookup key: attach n's factors p, q (half_limbs
 * limbs each, p q = n, p^2 and q^2 below 2^(16 W)) and the CRT idempotents
 * ep = q^2 (q^-2 mod p^2), eq = p^2 (p^-2 mod q^2) (W limbs each, W =
 * chx_lookup_width). From then on chx_lookup_decrypt, _decode_hamming and
 * _reencrypt compute c^a mod p^2 and mod q^2 and recombine c^a mod n^2 before
 * the floor division L, so every c in [0, n^2) gives the result of
 * PaillierLookup.decrypt (lookup.py:257-278), which exponentiates at full
 * width; the reference has no CRT form. Items whose c^a is 1 mod n (every
 * ciphertext the scheme makes) take the textbook CRT finish at half and
 * quarter width instead of the recombination; the result is the same. The
 * secret key's phi yields p and q (p + q = n - phi + 1). Needs a 2048- or 4096-bit width; refused with
 * CHX_EUNSUPPORTED otherwise, and the key keeps the full-width form. */
int chx_lookup_key_set_crt(chx_ctx* ctx, chx_lkey* k, const uint32_t* p, const uint32_t* q, int half_limbs,
                           const uint32_t* ep, const uint32_t* eq);
/* 1 when chx_lookup_key_set_crt has succeeded on k, else 0 */
int chx_lookup_key_has_crt(const chx_lkey* k);
/* Rows the row-group tables cover (0, the default: every row of the g
 * table). No counterpart in the reference, whose tables have one product per
 * row (lcu:550-588): a memory hint for callers whose plaintexts have zero
 * digits above a row, e.g. packed embeddings of chunk_len bits use
 * ceil(chunk_len / 8) rows (hamming/paillier.py:26-32), so a 4096-bit n^2 key
 * with 512-bit embeddings needs half of every group table. Digits above the
 * covered rows stay correct: those rows take one product each. Group tables
 * built for another coverage are freed and rebuilt on next use. */
int chx_lookup_key_set_table_rows(chx_ctx* ctx, chx_lkey* k, int rows);
/* c[i] = prod_j g_table[j][digits[i][j]] * prod_t noise[noise_idx[i][t]] mod n^2
 * -- encrypt_kernel (lcu:839-948), encrypt_cts_kernel (lcu:951-1011).
 * digits: count x chunks bytes; noise_idx: count x nmult bytes. */
int chx_lookup_encrypt(chx_ctx* ctx, const chx_lkey* k, const uint8_t* digits, const uint8_t* noise_idx,
                       int nmult, uint32_t* c, size_t count);
int chx_lookup_encrypt_dev(chx_ctx* ctx, const chx_lkey* k, const uint8_t* digits, const uint8_t* noise_idx,
                           int nmult, uint32_t* c, size_t count, size_t stride);
/* m[i] = L(c[i]^a mod n^2) * g_a_inv mod n -- decrypt_only_kernel (lcu:1014-1050),
 * decrypt_cts (lcu:2440-2499) */
int chx_lookup_decrypt(chx_ctx* ctx, const chx_lkey* k, const uint32_t* c, uint32_t* m, size_t count);
/* fused decrypt + odd-bit popcount -- decode_hamming_fused_kernel (lcu:1140-1192)
 * replacing K8 + odd_popcount_kernel (lcu:1279-1298) */
int chx_lookup_decode_hamming(chx_ctx* ctx, const chx_lkey* k, const uint32_t* c, const uint32_t* mask,
                              uint32_t* counts, size_t count);
/* fused re-encryption: m = decrypt under `from`, then out = the lookup encryption
 * of m's base-2^8 digits under `to` with noise indices noise_idx (count x nmult
 * bytes, below to's noise table size) -- decrypt_then_reencrypt_kernel
 * (lcu:1053-1135), decrypt_then_reencrypt (lcu:2134-2285). Both keys belong to
 * ctx; from's n may have at most 8 x to's message chunks bits (no digit of a
 * plaintext is dropped). c: count x from's width limbs, out: count x to's width.
 * The digits stay on the device; with the same noise indices the result equals
 * chx_lookup_encrypt of the decrypted plaintexts' digits. */
int chx_lookup_reencrypt(chx_ctx* ctx, const chx_lkey* from, const chx_lkey* to, const uint32_t* c,
                         const uint8_t* noise_idx, int nmult, uint32_t* out, size_t count);
/* device pointers; c and out interleaved with `stride` (at each key's width) */
int chx_lookup_reencrypt_dev(chx_ctx* ctx, const chx_lkey* from, const chx_lkey* to, const uint32_t* c,
                             const uint8_t* noise_idx, int nmult, uint32_t* out, size_t count, size_t stride);

/* ---- layout conversion (device) ---- */
int chx_aos_to_il_dev(chx_ctx* ctx, const uint32_t* aos, uint32_t* il, size_t count, int limbs, size_t stride);
int chx_il_to_aos_dev(chx_ctx* ctx, const uint32_t* il, uint32_t* aos, size_t count, int limbs, size_t stride);

/* ---- host-side helpers (no device needed; exposed so the host logic can be
 *      checked on machines without a GPU) ---- */
/* sliding-window schedule of exponent e (packed as the kernels read it);
 * returns the number of entries (may exceed cap; only cap are written) */
int chx_host_schedule(const uint32_t* e, int e_limbs, int window, uint32_t* out, int cap);
/* Montgomery constants of odd n (limbs): n0inv = -n^-1 mod 2^32, r2 = R^2 mod n */
int chx_host_mont_consts(const uint32_t* n, int limbs, uint32_t* n0inv, uint32_t* r2);
/* Consistency of a Paillier secret key given with its factors, the check
 * chx_paillier_key_create applies to such a key: p, q distinct primes (odd,
 * Miller-Rabin with 12 fixed bases) with p q = n, phi inv = 1 mod n,
 * lcm(p-1, q-1) | phi, h_p (-q) = 1 mod p, h_q (-p) = 1 mod q, q_inv q = 1
 * mod p. n, phi, inv: n_limbs limbs; the rest half_limbs. CHX_EINVAL names
 * the first failed condition (chx_last_error). No counterpart in the
 * reference, whose keys carry no factors (scheme.py:24-44). */
int chx_host_paillier_check(const uint32_t* n, int n_limbs, const uint32_t* phi, const uint32_t* inv,
                            const uint32_t* p, const uint32_t* q, int half_limbs, const uint32_t* hp,
                            const uint32_t* hq, const uint32_t* qinv);

/* ---- measurement ---- */
/* time `iters` rounds of 16 dependent-free v_mad_u64_u32 per lane on `blocks`
 * blocks of 256 lanes; returns seconds and the number of 32x32->64 MACs */
int chx_probe_mad(chx_ctx* ctx, int blocks, int iters, double* seconds, double* macs);
/* same for other instruction kinds: 0 v_mad_u64_u32 (carry-out to VCC), 1 v_mul_lo_u32,
 * 2 v_mul_hi_u32, 3 64-bit add, 4 v_fma_f64, 5 v_mad_u32_u24, 6 v_mad_u64_u32 with the
 * carry-out in an SGPR pair, 7 kind 6 plus one v_add_u32 per two MADs; *ops = the
 * operations of the kind (MADs for 6 and 7) */
int chx_probe_rate(chx_ctx* ctx, int kind, int blocks, int iters, double* seconds, double* ops);
/* calibration copy of an interleaved batch (device, [limbs][stride]) with the
 * lane-group access pattern of the exponentiation kernels: limbs 64 with tpi 4|8,
 * limbs 128 with tpi 8|16; moves exactly count*limbs*4 bytes each way */
int chx_probe_copy(chx_ctx* ctx, const uint32_t* src, uint32_t* dst, size_t count, size_t stride, int limbs,
                   int tpi);
/* HIP-event timing of the compute kernels enqueued by this context (events
 * recorded on the context's stream around every launch while enabled) */
int chx_ctx_set_timing(chx_ctx* ctx, int enable);
/* total milliseconds and launch count of kernel `name` (e.g. "crt_half",
 * "modexp", "mulmod_bcast") since the last reset; synchronises the stream */
int chx_ctx_kernel_time(chx_ctx* ctx, const char* name, double* total_ms, int* launches);
int chx_ctx_reset_timing(chx_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* CHX_H */
