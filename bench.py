#!/usr/bin/env python3
"""Paillier batched-modexp benchmark on one MI355X (gfx950).

Workload (default): CRT decryption of a batch of 65,536 ciphertexts under a
seeded key with n of 2048 bits (n^2 of 4096 bits). Each ciphertext needs two
modexps -- mod p^2 and mod q^2, 2048-bit moduli, 1024-bit exponents p-1 and
q-1 -- so one step is 131,072 modexps at n^2 = 2048 bits, then the Garner
recombination. Inputs (ciphertexts of seeded m and r) are resident in HBM
before the timed region; every output of every step is checked bit-exactly
against the seeded plaintexts after timing.

Reported beside it:
  * roofline: algorithmic 32x32->64 multiply-accumulates per crt_half launch
    (fixed convention: one L-limb modmul = 2 L^2 MACs, one e-bit modexp =
    e + ceil(e/6) modmuls) over the launch time measured with HIP events on the
    stream the kernel runs on, against the integer-multiply peak measured by
    one v_mad_u64_u32 per lane per clock (256 CUs x 64 lanes x 2.4 GHz), with a
    MAD-only probe kernel's measured rate reported beside it;
  * the plain per-call GPU form (host buffers, one call per 1,024 ciphertexts,
    synchronous) on the same batch;
  * the CPU baseline: GMP mpz_powm CRT decryption (the engine gmpy2 wraps) on
    a bounded sample, single thread and one process per host core.
"""

from __future__ import annotations

import argparse
import gc
import json
import math
import os
import random
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Paillier modexp/s @ n²=2048b, batched vs plain, % INT peak; p99 enc-kNN latency"
CLOCK_HZ = 2.4e9  # MI355X peak engine clock


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cpu_worker(args):
    cp, cs = args
    from oracle import gmp_engine as G
    dec = G.CrtDecryptor(**cp)
    return [dec(c) for c in cs]


def cpu_baseline(cp: dict, cs: list[int], want: list[int], cores: int) -> dict:
    """GMP CRT decryption of a bounded sample (checker + reported baseline)."""
    from multiprocessing import get_context

    from oracle import gmp_engine as G
    engine = "gmp " + G.version() if G.available() else "cpython-pow"
    if not G.available():
        from oracle import paillier_oracle as O
        t0 = time.perf_counter()
        got = []
        for c in cs:
            p, q = cp["p"], cp["q"]
            mp = ((pow(c % (p * p), p - 1, p * p) - 1) // p) * cp["hp"] % p
            mq = ((pow(c % (q * q), q - 1, q * q) - 1) // q) * cp["hq"] % q
            got.append(mq + q * ((mp - mq) * cp["q_inv_p"] % p))
        t1 = time.perf_counter()
        assert got == want
        rate1 = len(cs) / (t1 - t0)
        return {"engine": engine, "single": 2 * rate1, "multi": 2 * rate1, "cores": 1, "n": len(cs)}
    single_n = min(len(cs), 2048)
    dec = G.CrtDecryptor(**cp)
    t0 = time.perf_counter()
    got1 = [dec(c) for c in cs[:single_n]]
    t1 = time.perf_counter()
    assert got1 == want[:single_n], "CPU baseline disagrees with the seeded plaintexts"
    rate1 = single_n / (t1 - t0)
    chunks = [cs[i::cores] for i in range(cores)]
    with get_context("fork").Pool(cores) as pool:
        pool.map(_cpu_worker, [(cp, [cs[0]]) for _ in range(cores)])  # warm the workers
        t0 = time.perf_counter()
        res = pool.map(_cpu_worker, [(cp, ch) for ch in chunks])
        t1 = time.perf_counter()
    got = [None] * len(cs)
    for i, r in enumerate(res):
        got[i::cores] = r
    assert got == want, "CPU baseline (pool) disagrees with the seeded plaintexts"
    rate_all = len(cs) / (t1 - t0)
    return {"engine": engine, "single": 2 * rate1, "multi": 2 * rate_all, "cores": cores, "n": len(cs),
            "single_n": single_n}


def lazy_tag(key) -> str:
    """Template arguments of the exponentiation kernel the CRT halves use."""
    return {32: "28, 38, 1, 32", 64: "29, 72, 4, 64", 128: "29, 144, 8, 128"}.get(key.half, "?")


def committed_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the rocprofv3 counter summary
    committed under profiles/ (FETCH_SIZE and WRITE_SIZE runs, each scaled by
    what the counter reads for a device copy of known size)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        text = open(path).read()
        if "```json" not in text:
            continue
        try:
            blob = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
        except ValueError:
            continue
        if blob.get("kernel") == kernel:
            return blob.get("traffic_bytes"), os.path.relpath(path, ROOT)
    return None, None



def _c1_worker(job):
    n, phi, inv, items = job
    from oracle import gmp_engine as G
    n2 = n * n
    out = []
    for m, r in items:
        c = (1 + m * n) * G.powm(r, n, n2) % n2          # scheme.py:63-83 with g = n + 1
        u = G.powm(c, phi, n2)                            # scheme.py:86-104
        out.append((c, (u - 1) // n * inv % n))
    return out


def host_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def c1_leg(ctx, args, cores: int) -> dict:
    """The reference's CPU-runnable case: textbook Paillier encrypt and decrypt
    of 1,024 messages as one batch (scheme.py:63-104), seeded key, m uniform in
    [0, n), r uniform in Z_n^*; at n of 1024 bits (n^2 2048-bit, the metric's
    width) and of 2048 bits (the reference's default key_len 1024). The GPU
    runs the host entry points (PaillierKey.encrypt / decrypt: host buffers,
    PCIe included); GMP (the engine gmpy2 wraps) runs the same formulas on
    the host cores. Every ciphertext and plaintext is compared."""
    from multiprocessing import get_context

    from cuhepy_amd.engine import PaillierKey
    from cuhepy_amd.limbs import limbs_to_ints
    from cuhepy_amd.paillier import keygen as kg
    from oracle import gmp_engine as G
    out = {}
    for key_len in (512, 1024):
        rng = random.Random(args.seed + key_len)
        k = kg.paillier_key_gen(key_len, rng)
        n, phi, inv = k["pk"]["n"], k["sk"]["phi"], k["sk"]["inv"]
        key = PaillierKey(ctx, n, phi, inv, k["p"], k["q"])
        ms = [rng.randrange(n) for _ in range(1024)]
        rs = [kg.random_r(n, rng) for _ in range(1024)]
        key.decrypt(key.encrypt(ms[:64], rs[:64]))  # warm (schedules, scratch)
        ctx.sync()
        t = time.perf_counter()
        c = key.encrypt(ms, rs)
        t_enc = time.perf_counter() - t
        d = key.decrypt(c)
        t_all = time.perf_counter() - t
        cs, ds = limbs_to_ints(c), limbs_to_ints(d)
        if ds != ms:
            raise RuntimeError(f"C1 key_len {key_len}: decrypt differs from the messages")
        row = {"messages": 1024, "n_bits": n.bit_length(), "n2_bits": (n * n).bit_length(),
               "gpu_ms": round(t_all * 1e3, 3), "gpu_encrypt_ms": round(t_enc * 1e3, 3),
               "gpu_decrypt_ms": round((t_all - t_enc) * 1e3, 3)}
        if G.available() and not args.no_cpu:
            items = list(zip(ms, rs))
            single = 256 if key_len == 512 else 32
            t = time.perf_counter()
            one = _c1_worker((n, phi, inv, items[:single]))
            t1 = time.perf_counter() - t
            chunks = [items[i::cores] for i in range(cores)]
            with get_context("fork").Pool(cores) as pool:
                pool.map(_c1_worker, [(n, phi, inv, items[:1])] * cores)  # warm the workers
                t = time.perf_counter()
                res = pool.map(_c1_worker, [(n, phi, inv, ch) for ch in chunks])
                tp = time.perf_counter() - t
            got = [None] * len(items)
            for i, r in enumerate(res):
                got[i::cores] = r
            bad = sum(int(g != (a, b)) for g, a, b in zip(got, cs, ds)) + sum(
                int(g != (a, b)) for g, a, b in zip(one, cs, ds))
            if bad:
                raise RuntimeError(f"C1 key_len {key_len}: GMP and GPU differ on {bad} messages")
            row.update({"cpu_engine": "gmp " + G.version(), "cpu_ms": round(tp * 1e3, 1), "cpu_cores": cores,
                        "cpu_single_thread_ms_per_1024": round(t1 / single * 1024 * 1e3, 1),
                        "cpu_single_thread_sample": single, "gpu_over_cpu_all_cores": round(tp / t_all, 1),
                        "bit_exact": {"ciphertexts": 1024, "plaintexts": 1024, "mismatches": 0}})
        out[f"n{n.bit_length()}"] = row
        log(f"C1 (1,024 messages, n {n.bit_length()}-bit): GPU encrypt+decrypt {t_all * 1e3:.2f} ms "
            f"(host buffers)" + (f"; GMP {row['cpu_ms']:.0f} ms on {cores} processes, "
                                 f"{row['cpu_single_thread_ms_per_1024']:.0f} ms on one thread" if "cpu_ms" in row
                                 else ""))
        key.close()
    return out


def ragged_leg(ctx, args, peak: float, device_name: str) -> dict:
    """Ragged mixed-width modexp: 256 sub-batches, sizes log-uniform in
    [1, 8192], n of 1024 or 2048 bits (moduli n^2 of 2048 / 4096 bits, 50/50),
    bases uniform below n^2, exponents of full modulus width; autotuned, one
    persistent launch replayed from a hipGraph; beside it the plain form (one
    launch per sub-batch)."""
    import numpy as np

    from cuhepy_amd.autotune import tune_ragged
    from cuhepy_amd.engine import RaggedPlan
    from cuhepy_amd.ragged import DeviceWorkload, make_subbatches
    from oracle import gmp_engine as G

    t = time.perf_counter()
    subs = make_subbatches(args.ragged_subs, args.ragged_max, seed=args.seed + 3)
    wl = DeviceWorkload(ctx, subs)
    items, macs = wl.total_items(), wl.algorithmic_macs()
    n64 = sum(s.count for s in subs if s.limbs == 64)
    log(f"ragged: {len(subs)} sub-batches, {items} modexps ({n64} at 2048-bit, {items - n64} at 4096-bit moduli), "
        f"{macs:.3e} MACs, uploaded ({time.perf_counter() - t:.1f}s)")
    t = time.perf_counter()
    if args.ragged_fixed:
        v, w, wpc = (int(x) for x in args.ragged_fixed.split(","))
        best, tlog = {"variant": v, "window": w, "waves": wpc}, []
    else:
        best, tlog = tune_ragged(ctx, wl, device_name, reps=1, force=True)
    log(f"ragged autotune ({time.perf_counter() - t:.1f}s): best {best}; " +
        ", ".join(f"v{c['variant']}/w{c['window']}/{c['waves']}wpc {s * 1e3:.0f}ms" for c, s in tlog))
    plan = RaggedPlan(ctx, wl.desc, best["variant"], best["window"], best["waves"])
    plan.run(graph=True)
    ctx.sync()
    ctx.reset_timing()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.ragged_steps):
        plan.run(graph=True)
    ctx.sync()
    wall = (time.perf_counter() - t0) / args.ragged_steps
    ms, nl = ctx.kernel_time("ragged")
    ctx.set_timing(False)
    launch_s = ms / max(nl, 1) / 1e3
    persistent = [wl.result(i) for i in range(len(subs))]
    powm = G.powm if G.available() else pow
    t = time.perf_counter()
    bad = wl.spot_check(2, random.Random(args.seed), powm)
    cpu_s = time.perf_counter() - t
    checked = wl.checked
    # plain form: one non-persistent launch per sub-batch, no graph
    wl.clear_outputs()
    plan.run_plain()
    ctx.sync()
    ctx.reset_timing()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    plan.run_plain()
    ctx.sync()
    plain_wall = time.perf_counter() - t0
    pms, _ = ctx.kernel_time("ragged_plain")
    ctx.set_timing(False)
    same = all(np.array_equal(wl.result(i), persistent[i]) for i in range(len(subs)))
    plan.close()
    achieved = macs / launch_s / 1e12
    log(f"ragged: persistent+graph {launch_s * 1e3:.1f} ms ({items / launch_s:,.0f} modexp/s, {achieved:.2f} TMAC/s, "
        f"{achieved / peak:.1%}); plain {pms:.1f} ms; spot-check mismatches {len(bad)}; plain == persistent: {same}")
    if bad or not same:
        raise RuntimeError(f"ragged results wrong: {bad[:5]} same={same}")
    return {
        "workload": f"{len(subs)} sub-batches, sizes log-uniform in [1, {args.ragged_max}], n 1024/2048-bit 50/50, "
                    "full-modulus-width exponents",
        "modexps": items, "macs": macs,
        "tuned": best,
        "persistent_graph_ms": round(launch_s * 1e3, 3),
        "modexp_per_s": round(items / launch_s, 1),
        "achieved_tmac_s": round(achieved, 4),
        "frac_of_peak": round(achieved / peak, 4),
        "wall_ms_per_run": round(wall * 1e3, 3),
        "plain_per_subbatch_ms": round(pms, 3),
        "plain_wall_ms": round(plain_wall * 1e3, 3),
        "persistent_over_plain": round(pms / (launch_s * 1e3), 3),
        "bit_exact": {"vs_gmp_items": checked, "mismatches": 0, "plain_equals_persistent": same},
        "cpu_gmp_single_thread_modexp_s": round(checked / cpu_s, 1),
    }


class _TableRows:
    """g_table[i][j] as integers, converted on access from the (chunks, 256, W) limb array."""

    def __init__(self, arr):
        self.arr = arr

    def __getitem__(self, i):
        from cuhepy_amd.limbs import limbs_to_int
        row = self.arr[i]

        class _Row:
            def __getitem__(self, j):
                return limbs_to_int(row[j])

        return _Row()


def _noise_products(nmult: int, triples: bool) -> int:
    """Noise products per item: three factors per product while more than
    three are left (noise-triple table, with 5-row groups), then two, the last
    product also leaving the Montgomery form (one conversion product when
    there is no noise)."""
    if nmult == 0:
        return 1
    n, j = 0, 0
    while j < nmult:
        j += 3 if triples and j + 3 < nmult else 2 if j + 1 < nmult else 1
        n += 1
    return n


def _lookup_products(digits, items_per_wave: int, nmult: int, rows: int, groups: int = 1 << 30) -> int:
    """Table products the lookup kernel executes for this digit matrix with
    `rows` table rows per product over the first `groups` groups: per wave,
    one per row group with a nonzero digit when every digit of the group is in
    the packing alphabet, else one per row with a nonzero digit (rows past the
    table's last group go singly); the first of them is a copy (the
    accumulator starts at the Montgomery one); plus the noise products
    (_noise_products; triples with 5-row groups). All items of a wave run
    every product the wave runs."""
    import numpy as np
    count, chunks = digits.shape
    waves = -(-count // items_per_wave)
    pad = np.zeros((waves * items_per_wave, chunks), dtype=digits.dtype)
    pad[:count] = digits
    d = pad.reshape(waves, items_per_wave, chunks)
    per_wave = np.zeros(waves, dtype=np.int64)
    whole = min(chunks - chunks % rows, rows * groups)
    for j in range(0, whole, rows):
        grp = d[:, :, j:j + rows]
        packed = ~((grp & 0xAA) != 0).any(axis=(1, 2))
        nz_grp = (grp != 0).any(axis=(1, 2))
        nz_rows = (grp != 0).any(axis=1).sum(axis=1)
        per_wave += np.where(packed, nz_grp.astype(np.int64), nz_rows)
    for j in range(whole, chunks):
        per_wave += (d[:, :, j] != 0).any(axis=1)
    per_wave -= per_wave > 0  # the first active product is a copy
    # every item of a wave takes part in every product the wave runs
    real = np.full(waves, items_per_wave, dtype=np.int64)
    real[-1] = count - (waves - 1) * items_per_wave
    return int((per_wave * real).sum()) + _noise_products(nmult, rows == 5) * count


def lookup_leg(ctx, args, peak: float) -> dict:
    """Paillier-lookup table encryption: 100,000 binary vectors of 512 bits
    (Bernoulli(0.5)), key_len 1024 (n 2048-bit, n^2 4096-bit), alpha 280,
    one ciphertext per vector, 14 noise factors; digits and noise indices
    resident in HBM when the timed region starts. Reported: the row-quint
    form with noise triples (25 GB of tables, the fastest), and beside it the
    default row quads, triples, pairs and single rows; every form's output is
    the same ciphertext."""
    import numpy as np
    import torch

    from cuhepy_amd._native import OP_LOOKUP
    from cuhepy_amd.engine import Context, stride_for
    from cuhepy_amd.limbs import limbs_to_ints, width_for_bits
    from cuhepy_amd.paillier._lookup_gpu_ext import PaillierGPULookupClient
    from cuhepy_amd.paillier.packing import embeddings_to_plaintexts, plaintexts_to_digits
    from oracle import paillier_oracle as O

    dev = torch.device("cuda:0")
    t = time.perf_counter()
    client = PaillierGPULookupClient(embed_len=512, key_len=1024, rng=random.Random(args.seed + 5))
    key = client._device_key()
    k = client._keys
    mc = client._message_chunks
    log(f"lookup: key n={k['n'].bit_length()} bits, a={k['a'].bit_length()} bits, tables built on the GPU "
        f"({time.perf_counter() - t:.1f}s)")
    count = args.lookup_vectors
    nrng = np.random.default_rng(args.seed + 6)
    bits = nrng.integers(0, 2, size=(count, 512), dtype=np.uint8)
    W = width_for_bits(k["n_squared"].bit_length())
    m = embeddings_to_plaintexts(bits, 1024, W)
    digits = plaintexts_to_digits(m, mc)
    nidx = nrng.integers(0, 256, size=(count, 14), dtype=np.uint8)
    stride = stride_for(count)
    d_dig = torch.from_numpy(digits).to(dev)
    d_idx = torch.from_numpy(nidx).to(dev)
    out = torch.zeros((W, stride), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    def timed(rows: int) -> tuple[float, float]:
        """(kernel seconds per launch, wall seconds per batch) with `rows`
        8-bit table rows per product (1 = single rows, 2 = pairs, 3 = triples,
        0 = the default, row quads, 5 = quints with noise triples)."""
        ctx.set_tuning(OP_LOOKUP, W, 0, rows)
        key.encrypt_digits_dev(d_dig, d_idx, 14, out, count, stride)  # warm
        torch.cuda.synchronize()
        ctx.reset_timing()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(args.lookup_steps):
            key.encrypt_digits_dev(d_dig, d_idx, 14, out, count, stride)
        torch.cuda.synchronize()
        w = (time.perf_counter() - t0) / args.lookup_steps
        ms, nl = ctx.kernel_time("lookup_encrypt")
        ctx.set_timing(False)
        return ms / max(nl, 1) / 1e3, w

    single_s, _ = timed(1)
    pairs_s, _ = timed(2)
    triples_s, _ = timed(3)
    t = time.perf_counter()
    ctx.set_tuning(OP_LOOKUP, W, 0, 5)
    key.encrypt_digits_dev(d_dig[:64], d_idx[:64], 14, out, 64, stride)  # builds the quint and noise-triple tables
    torch.cuda.synchronize()
    build5_s = time.perf_counter() - t
    quints_s, quints_wall = timed(5)
    out5 = out.clone()
    launch_s, wall = timed(0)  # row quads (the default); its output is checked below
    ctx.set_tuning(OP_LOOKUP, W, 0, 0)
    q5_equal = bool(torch.equal(out5[:, :count], out[:, :count]))
    if not q5_equal:
        raise RuntimeError("lookup encryption with quints and noise triples differs from the row-quad form")
    del out5
    nnz = int((digits != 0).sum())
    mulmods = nnz + 13 * count
    macs = mulmods * 2 * W * W
    achieved = macs / launch_s / 1e12
    ipw = 64 // Context.supported_tpi(W, op=OP_LOOKUP)[0]
    table_rows = -(-min(client.chunk_len, 2 * client.embed_len) // 8)
    executed = _lookup_products(digits, ipw, 14, 4, -(-table_rows // 4))
    executed_tmac = executed * 2 * W * W / launch_s / 1e12
    executed5 = _lookup_products(digits, ipw, 14, 5, -(-table_rows // 5))
    executed5_tmac = executed5 * 2 * W * W / quints_s / 1e12
    entry_bytes = 4 * (144 if W == 128 else 72 if W == 64 else 38)
    quint_bytes = -(-table_rows // 5) * 16 ** 5 * entry_bytes + 256 ** 3 * entry_bytes
    log(f"lookup: quints + noise triples: {count} vectors in {quints_s * 1e3:.2f} ms "
        f"({count / quints_s:,.0f} vectors/s), {executed5 / count:.1f} products per vector, "
        f"{executed5_tmac:.2f} TMAC/s = {executed5_tmac / peak:.1%}; tables {quint_bytes / 1e9:.1f} GB built in "
        f"{build5_s:.2f} s; equal to the row-quad output: {q5_equal}")
    # round trip: every ciphertext decrypts to its plaintext (GPU decrypt)
    aos = torch.empty((count, W), dtype=torch.int32, device=dev)
    ctx.il_to_aos(out, aos, count, W, stride)
    torch.cuda.synchronize()
    c_aos = aos.cpu().numpy().view(np.uint32)
    back = key.decrypt(c_aos)
    rt_bad = int((back != m).any(axis=1).sum())
    # formula check against the oracle on a sample (injected noise indices)
    g_tab, noise = client._ensure_tables()
    nt = limbs_to_ints(noise)
    rows = _TableRows(g_tab)
    pk = {"n": k["n"], "n_squared": k["n_squared"]}
    sample = list(range(0, count, max(1, count // args.lookup_cpu_sample)))[: args.lookup_cpu_sample]
    ms_int = limbs_to_ints(m[sample])
    t = time.perf_counter()
    want = [O.lookup_encrypt(mi, pk, rows, nt, mc, nidx[i].tolist()) for mi, i in zip(ms_int, sample)]
    cpu_s = time.perf_counter() - t
    got = limbs_to_ints(c_aos[sample])
    f_bad = sum(int(a != b) for a, b in zip(got, want))
    # CPU baseline with GMP (the engine gmpy2 wraps), the reference's loop over every digit row
    gmp_lookup = None
    from oracle import gmp_engine as G
    if G.available():
        enc = G.LookupEncryptor(rows, nt, k["n_squared"], mc)
        gsample = list(range(0, count, max(1, count // args.lookup_gmp_sample)))[: args.lookup_gmp_sample]
        gms = limbs_to_ints(m[gsample])
        for mi in gms:
            enc.prepare(mi)
        t = time.perf_counter()
        gout = [enc(mi, nidx[i].tolist()) for mi, i in zip(gms, gsample)]
        g_s = time.perf_counter() - t
        g_bad = sum(int(a != b) for a, b in zip(gout, limbs_to_ints(c_aos[gsample])))
        if g_bad:
            raise RuntimeError(f"GMP lookup encryption differs from the GPU on {g_bad} vectors")
        gmp_lookup = {"engine": "gmp " + G.version() + " (ctypes), one thread", "vectors": len(gsample),
                      "vectors_per_s": round(len(gsample) / g_s, 2),
                      "mulmods_per_vector": mc + 14, "equal_to_gpu": len(gsample)}
        log(f"lookup: GMP CPU baseline {len(gsample) / g_s:.1f} vectors/s single thread "
            f"({mc + 14} mulmods per vector), {len(gsample)} equal to the GPU ciphertexts")
    log(f"lookup: {count} vectors in {launch_s * 1e3:.2f} ms ({count / launch_s:,.0f} vectors/s); row quads "
        f"execute {executed / count:.1f} products per vector: {executed_tmac:.2f} TMAC/s = "
        f"{executed_tmac / peak:.1%} of the issue rate (the reference's one product per nonzero digit, "
        f"{mulmods / count:.1f} per vector, would need {achieved:.2f} TMAC/s for this rate); triples "
        f"{triples_s * 1e3:.2f} ms, pairs {pairs_s * 1e3:.2f} ms, single rows {single_s * 1e3:.2f} ms; round-trip "
        f"mismatches {rt_bad}; oracle mismatches {f_bad}/{len(sample)}; CPU {len(sample) / cpu_s:.1f} vectors/s")
    if rt_bad or f_bad:
        raise RuntimeError(f"lookup encryption wrong: round trip {rt_bad}, oracle {f_bad}")
    # the client path end to end (validation, packing, transfer, kernels,
    # output conversion) on a slice, for each input/output form
    sl = min(count, args.lookup_client_vectors)
    client_path = {"vectors": sl}
    as_lists = bits[:sl].tolist()
    for name, fn, arg in (("lists_in_ints_out", client.encrypt, as_lists),
                          ("numpy_in_ints_out", client.encrypt, bits[:sl]),
                          ("numpy_in_bytes_out", client.encrypt_bytes, bits[:sl]),
                          ("lists_in_bytes_out", client.encrypt_bytes, as_lists)):
        fn(arg[:64])  # first-call set-up outside the timing
        gc.collect()  # as the reference's time_call (benchmarks/bfv_client_matrix.py:40-44)
        gc.disable()
        t = time.perf_counter()
        fn(arg)
        client_path[name + "_vectors_per_s"] = round(sl / (time.perf_counter() - t), 1)
        gc.enable()
    client_s = sl / client_path["lists_in_ints_out_vectors_per_s"]
    log("lookup: client path end to end, vectors/s: " + ", ".join(
        f"{k[:-14]} {v:,.0f}" for k, v in client_path.items() if k.endswith("_per_s")))
    reenc = _reencrypt_block(ctx, args, key, out, d_dig, d_idx, m, count, stride, W)
    return {
        "workload": f"{count} binary vectors x 512 bits (Bernoulli 0.5), key_len 1024 (n^2 4096-bit), alpha 280, "
                    "14 noise factors per ciphertext",
        "vectors": count,
        "form": "row quints with noise triples (chx_ctx_set_tuning window 5): 5-row group tables over the 130 "
                "rows 512-bit packed embeddings use, three noise factors per product, "
                f"{quint_bytes / 1e9:.1f} GB of tables; outputs equal to the default row-quad form",
        "kernel_ms": round(quints_s * 1e3, 3),
        "vectors_per_s": round(count / quints_s, 1),
        "executed_products": executed5, "products_per_vector": round(executed5 / count, 2),
        "achieved_tmac_s": round(executed5_tmac, 4),
        "frac_of_peak": round(executed5_tmac / peak, 4),
        "reference_algorithm": {"mulmods": mulmods, "macs": macs, "mulmod_per_s": round(mulmods / quints_s, 1),
                                "equivalent_tmac_s": round(macs / quints_s / 1e12, 4)},
        "table_bytes": quint_bytes, "table_build_s": round(build5_s, 3),
        "wall_ms_per_batch": round(quints_wall * 1e3, 3),
        "row_groups": {"default_row_quads": {
                           "kernel_ms": round(launch_s * 1e3, 3),
                           "vectors_per_s": round(count / launch_s, 1),
                           "products_per_vector": round(executed / count, 2),
                           "achieved_tmac_s": round(executed_tmac, 4),
                           "frac_of_peak": round(executed_tmac / peak, 4),
                           "wall_ms_per_batch": round(wall * 1e3, 3),
                           "reference_algorithm_equivalent_tmac_s": round(achieved, 4)},
                       "triples_kernel_ms": round(triples_s * 1e3, 3),
                       "pairs_kernel_ms": round(pairs_s * 1e3, 3),
                       "single_rows_kernel_ms": round(single_s * 1e3, 3),
                       "quints_speedup_over_quads": round(launch_s / quints_s, 3),
                       "quints_speedup_over_single_rows": round(single_s / quints_s, 3)},
        "client_path_vectors_per_s": round(sl / client_s, 1),
        "client_path": client_path,
        "bit_exact": {"round_trip_vectors": count, "oracle_vectors": len(sample), "mismatches": 0},
        "cpu_port_single_thread_vectors_s": round(len(sample) / cpu_s, 2),
        "cpu_gmp": gmp_lookup,
        "reencrypt": reenc,
    }


def _reencrypt_block(ctx, args, key, c_il, d_dig, d_idx, m, count: int, stride: int, W: int) -> dict:
    """Fused re-encryption (decrypt_then_reencrypt, lcu:1053-1135) of the
    lookup leg's ciphertexts from its key to a second key_len-1024 key: the
    decryption c^a mod n^2 (a 280-bit), the device digit extraction and the
    row-quad encryption under the second key, all on the device. Checked over
    every vector: it equals the second key's encryption of the plaintexts'
    digits with the same noise indices, and decrypts to the plaintexts."""
    import numpy as np
    import torch

    from cuhepy_amd.engine import LookupKey
    from cuhepy_amd.paillier._lookup_gpu_ext import PaillierGPULookupClient

    dev = torch.device("cuda:0")
    t = time.perf_counter()
    kb = PaillierGPULookupClient(embed_len=512, key_len=1024, rng=random.Random(args.seed + 7))._device_key()
    log(f"reencrypt: second key and its tables built ({time.perf_counter() - t:.1f}s)")
    out = torch.zeros((W, stride), dtype=torch.int32, device=dev)
    ref = torch.zeros((W, stride), dtype=torch.int32, device=dev)
    key.reencrypt_dev(c_il, kb, d_idx, 14, out, count, stride)  # warm: the second key's quad table
    torch.cuda.synchronize()
    names = ("lazy_modexp", "decrypt_finish", "il_digits", "lookup_encrypt")
    ctx.reset_timing()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.lookup_steps):
        key.reencrypt_dev(c_il, kb, d_idx, 14, out, count, stride)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.lookup_steps
    per = {}
    for nm in names:  # per step (the decryption's finish may be two launches)
        per[nm] = ctx.kernel_time(nm)[0] / args.lookup_steps
    ctx.set_timing(False)
    kernel_ms = sum(per.values())
    kb.encrypt_digits_dev(d_dig, d_idx, 14, ref, count, stride)
    torch.cuda.synchronize()
    differ = int((out[:, :count] != ref[:, :count]).any(dim=0).sum())
    aos = torch.empty((count, W), dtype=torch.int32, device=dev)
    ctx.il_to_aos(out, aos, count, W, stride)
    torch.cuda.synchronize()
    cts_b = aos.cpu().numpy().view(np.uint32)
    dec_ms = {}
    for form, k in (("half_width_crt", kb), ("full_width", LookupKey(ctx, kb.n, kb.a, kb.g_a_inv, [], []))):
        assert k.crt == (form == "half_width_crt")
        ctx.reset_timing()
        ctx.set_timing(True)
        got = k.decrypt(cts_b)
        ctx.set_timing(False)
        dec_ms[form] = sum(ctx.kernel_time(nm)[0] for nm in ("lazy_modexp", "decrypt_finish"))
        if form == "half_width_crt":
            back = got
        elif not np.array_equal(got, back):
            raise RuntimeError("half-width and full-width lookup decryptions differ")
    rt_bad = int((back != m).any(axis=1).sum())
    log(f"lookup decrypt of {count} ciphertexts: half width (CRT) {dec_ms['half_width_crt']:.2f} ms, "
        f"full width {dec_ms['full_width']:.2f} ms of kernels, results equal")
    log(f"reencrypt: {count} vectors in {kernel_ms:.2f} ms of kernels ({count / kernel_ms * 1e3:,.0f} vectors/s; "
        f"decrypt {per['lazy_modexp']:.2f} + {per['decrypt_finish']:.3f} ms, digits {per['il_digits']:.3f} ms, "
        f"encrypt {per['lookup_encrypt']:.2f} ms), wall {wall * 1e3:.2f} ms; differ from the two-step "
        f"encryption: {differ}; round-trip mismatches {rt_bad}")
    if differ or rt_bad:
        raise RuntimeError(f"re-encryption wrong: {differ} differ from the two-step path, {rt_bad} round trip")
    return {"vectors": count, "kernel_ms": round(kernel_ms, 3), "vectors_per_s": round(count / kernel_ms * 1e3, 1),
            "decrypt_ms": round(per["lazy_modexp"] + per["decrypt_finish"], 3),
            "digits_ms": round(per["il_digits"], 4), "encrypt_ms": round(per["lookup_encrypt"], 3),
            "wall_ms_per_batch": round(wall * 1e3, 3), "decrypt_form": "half_width_crt" if key.crt else "full_width",
            "lookup_decrypt_kernel_ms": {k: round(v, 3) for k, v in dec_ms.items()},
            "bit_exact": {"vs_two_step_vectors": count, "round_trip_vectors": count, "mismatches": 0,
                          "half_width_equals_full_width": True}}


def serve_leg(ctx, args, peak: float) -> dict:
    """Served encrypted kNN: 1,000,000 lookup-encrypted 512-bit binary vectors
    (key_len 1024, n^2 4096-bit, 512 B per ciphertext) resident in HBM; each
    query (an encrypted 512-bit vector) is multiplied into every index entry
    (encode_hamming_server). Reported: the single-query kernel against both
    rooflines, saturation throughput, and p50/p99 latency (arrival to device
    completion) of Poisson arrivals at a sweep of rates, for continuous
    batching and for the plain one-launch-per-query FIFO form."""
    import numpy as np
    import torch

    from cuhepy_amd.engine import Modulus, stride_for
    from cuhepy_amd.limbs import limbs_to_ints, width_for_bits
    from cuhepy_amd.paillier._lookup_gpu_ext import PaillierGPULookupClient
    from cuhepy_amd.paillier.packing import embeddings_to_plaintexts, plaintexts_to_digits
    from cuhepy_amd.serve import KnnServer, latency_stats, poisson_arrivals

    dev = torch.device("cuda:0")
    t = time.perf_counter()
    client = PaillierGPULookupClient(embed_len=512, key_len=1024, rng=random.Random(args.seed + 7))
    key = client._device_key()
    k = client._keys
    n2 = k["n_squared"]
    W = width_for_bits(n2.bit_length())
    mc = client._message_chunks
    count = args.serve_index
    stride = stride_for(count)
    index = torch.empty((W, stride), dtype=torch.int32, device=dev)
    nrng = np.random.default_rng(args.seed + 8)
    batch = 100_000
    tmp = torch.empty((W, stride_for(batch)), dtype=torch.int32, device=dev)
    keep_bits = None
    for off in range(0, count, batch):
        cnt = min(batch, count - off)
        bits = nrng.integers(0, 2, size=(cnt, 512), dtype=np.uint8)
        if keep_bits is None:
            keep_bits = bits[:256].copy()
        digits = plaintexts_to_digits(embeddings_to_plaintexts(bits, 1024, W), mc)
        nidx = nrng.integers(0, 256, size=(cnt, 14), dtype=np.uint8)
        d_dig, d_idx = torch.from_numpy(digits).to(dev), torch.from_numpy(nidx).to(dev)
        torch.cuda.synchronize()
        key.encrypt_digits_dev(d_dig, d_idx, 14, tmp, cnt, tmp.shape[1])
        torch.cuda.synchronize()
        index[:, off:off + cnt] = tmp[:, :cnt]
    torch.cuda.synchronize()
    del tmp
    log(f"serve: index of {count} encrypted 512-bit vectors resident ({W * stride * 4 / 2**20:.0f} MiB, "
        f"{time.perf_counter() - t:.1f}s)")
    mod = Modulus(ctx, n2, W)
    qrng = np.random.default_rng(args.seed + 9)

    def queries(nq):
        qb = qrng.integers(0, 2, size=(nq, 512), dtype=np.uint8)
        return qb, [int(c[0]) for c in client.encrypt(qb.tolist())]

    # 1) one query alone, whole index in one launch: the kernel against its rooflines
    qb, ys = queries(1)
    plain = KnnServer(ctx, mod, index, count, chunk=count, max_active=1, mode="scan")
    r = plain.submit(ys[0], keep=True)
    plain.drain()
    for _ in range(10):  # warm-up: result buffers pooled, clocks up
        plain.submit(ys[0])
    plain.drain()
    ctx.reset_timing()
    ctx.set_timing(True)
    reps = 5
    for _ in range(reps):
        plain.submit(ys[0])
    plain.drain()
    ms, nl = ctx.kernel_time("mulmod_multi")
    ctx.set_timing(False)
    kern_s = ms / max(nl, 1) / 1e3
    macs = count * 2 * W * W
    hbm_bytes = count * 2 * W * 4
    # bit-exact: sampled entries against exact integers, decoded distances against the plaintexts
    aos = torch.empty((256, W), dtype=torch.int32, device=dev)
    idx_aos = torch.empty((256, W), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    ctx.il_to_aos(r.result, aos, 256, W, stride)
    ctx.il_to_aos(index, idx_aos, 256, W, stride)
    ctx.sync()
    got = limbs_to_ints(aos.cpu().numpy().view(np.uint32))
    xs = limbs_to_ints(idx_aos.cpu().numpy().view(np.uint32))
    exact_bad = sum(int(g != x * ys[0] % n2) for g, x in zip(got, xs))
    dist = client.decode_hamming_client([[c] for c in got])
    want = [int((b != qb[0]).sum()) for b in keep_bits]
    dist_bad = sum(int(a != b) for a, b in zip(dist, want))
    plain.release(r)
    if exact_bad or dist_bad:
        raise RuntimeError(f"served evaluation wrong: {exact_bad} products, {dist_bad} distances")
    log(f"serve: one query = {count} mulmods in {kern_s * 1e3:.3f} ms ({count / kern_s:,.0f} mulmod/s, "
        f"{macs / kern_s / 1e12:.2f} TMAC/s = {macs / kern_s / 1e12 / peak:.1%} of the integer peak, "
        f"{hbm_bytes / kern_s / 1e9:.0f} GB/s = {hbm_bytes / kern_s / 8e12:.1%} of HBM); 256 entries exact, "
        f"256 distances decoded correctly")

    # 1b) the pointwise form (per-item operands: encode_hamming_server of two
    # ciphertext lists): index x rolled index, lazy-engine multiply
    other = torch.roll(index[:, :count], 1, dims=1).contiguous()
    pw_out = torch.empty((W, stride), dtype=torch.int32, device=dev)
    other_full = torch.zeros((W, stride), dtype=torch.int32, device=dev)
    other_full[:, :count] = other
    del other
    pw_forms = {}
    for form, kname in ((2, "mulmod"), (0, "pw_mulmod")):  # two Montgomery products; one product + fold
        ctx.set_tuning(1, W, 0, form)
        for _ in range(3):  # warm-up (the first call builds the fold table)
            mod.mulmod_dev(index, other_full, pw_out, count, stride)
        torch.cuda.synchronize()
        ctx.reset_timing()
        ctx.set_timing(True)
        for _ in range(reps):
            mod.mulmod_dev(index, other_full, pw_out, count, stride)
        torch.cuda.synchronize()
        pms, pnl = ctx.kernel_time(kname)
        ctx.set_timing(False)
        if pnl != reps:
            raise RuntimeError(f"pointwise form {form}: {pnl} launches of {kname} timed, expected {reps}")
        pw_forms[kname] = pms / pnl / 1e3
    ctx.set_tuning(1, W, 0, 0)
    pw_s = pw_forms["pw_mulmod"]
    pw_two_s = pw_forms["mulmod"]
    pw_aos = torch.empty((256, W), dtype=torch.int32, device=dev)
    oth_aos = torch.empty((256, W), dtype=torch.int32, device=dev)
    ctx.il_to_aos(pw_out, pw_aos, 256, W, stride)
    ctx.il_to_aos(other_full, oth_aos, 256, W, stride)
    ctx.sync()
    pw_got = limbs_to_ints(pw_aos.cpu().numpy().view(np.uint32))
    oth = limbs_to_ints(oth_aos.cpu().numpy().view(np.uint32))
    pw_bad = sum(int(g != x * y % n2) for g, x, y in zip(pw_got, xs, oth))
    del pw_out, other_full
    if pw_bad:
        raise RuntimeError(f"pointwise multiply wrong: {pw_bad} of 256 products")
    log(f"serve: pointwise form (per-item operands, k_pw_mulmod) {count} mulmods in {pw_s * 1e3:.3f} ms "
        f"({count / pw_s:,.0f} mulmod/s, {macs / pw_s / 1e12:.2f} TMAC/s = {macs / pw_s / 1e12 / peak:.1%}, "
        f"{count * 3 * W * 4 / pw_s / 1e9:.0f} GB/s); two-product form {pw_two_s * 1e3:.3f} ms; "
        f"256 entries exact")

    # 1c) CPU baseline: the reference's evaluate loop int(mpz(a) * y % n^2) over
    # a sample of the index (benchmarks/bfv_client_matrix.py:166-169), GMP
    cpu_eval = None
    from oracle import gmp_engine as G
    if G.available() and not args.no_cpu:
        ns = min(args.serve_cpu_sample, count)
        s_aos = torch.empty((ns, W), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        ctx.il_to_aos(index, s_aos, ns, W, stride)
        ctx.sync()
        sx = limbs_to_ints(s_aos.cpu().numpy().view(np.uint32))
        ev = G.Evaluator(ys[0], n2)
        t = time.perf_counter()
        sout = [ev(a) for a in sx]
        cpu_s = time.perf_counter() - t
        if sout[:256] != got:
            raise RuntimeError("CPU evaluate loop differs from the served products")
        cpu_eval = {"engine": "gmp " + G.version() + " (ctypes), one thread, the loop int(mpz(a) * y % n^2)",
                    "entries": ns, "mulmod_per_s": round(ns / cpu_s, 1),
                    "us_per_entry": round(cpu_s / ns * 1e6, 3),
                    "ms_per_query_over_index": round(count * cpu_s / ns * 1e3, 1),
                    "reference_published_us_per_entry": 5.3,
                    "reference_published_source": "benchmarks/results/native_bfv_1024.json:97 (gmpy2, one thread)"}
        log(f"serve: CPU evaluate loop (GMP) {ns / cpu_s:,.0f} mulmod/s single thread = "
            f"{count * cpu_s / ns:.2f} s per query over the index; equal to the GPU on 256 entries")

    # 1d) the kernel alone: B queries in one stamped launch against B single
    # launches (host synchronised around each form; best of four)
    from cuhepy_amd.limbs import ints_to_limbs
    _, kys = queries(32)
    kops = [torch.empty((mod.operand_words,), dtype=torch.int32, device=dev) for _ in kys]
    for y, op in zip(kys, kops):
        mod.operands(ints_to_limbs([y], W), op)
    kouts = [torch.empty((W, stride), dtype=torch.int32, device=dev) for _ in kys]
    kdone = torch.zeros(64, dtype=torch.int32, device=dev)
    kst = torch.zeros(32, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    ctx.sync()
    batching_kernel = {}
    for B in (1, 4, 32):
        ts, tb = [], []
        for _ in range(4):
            ctx.sync()
            t0 = time.perf_counter()
            for q in range(B):
                mod.mulmod_multi_dev(index, stride, 0, count, [kops[q]], [kouts[q]], stride)
            ctx.sync()
            ts.append((time.perf_counter() - t0) / B)
            t0 = time.perf_counter()
            mod.mulmod_multi_stamped_dev(index, stride, 0, count, kops[:B], kouts[:B], stride, kdone, kst)
            ctx.sync()
            tb.append((time.perf_counter() - t0) / B)
        batching_kernel[str(B)] = {"single_launches_ms_per_query": round(min(ts) * 1e3, 4),
                                   "one_launch_ms_per_query": round(min(tb) * 1e3, 4)}
    del kouts
    log(f"serve: kernel-level batching, ms per query (B single launches | one launch of B): "
        + ", ".join(f"B={b}: {v['single_launches_ms_per_query']:.3f} | {v['one_launch_ms_per_query']:.3f}"
                    for b, v in batching_kernel.items()))

    # 2) saturation: 128 queries submitted at once, batched (FIFO launches of
    # up to 32 admitted queries, 64 admitted at most) and plain (one launch per
    # query); the two forms alternate, twice each, and each keeps its best run
    srv = KnnServer(ctx, mod, index, count, max_active=64, mode="fifo")
    _, ys = queries(128)
    for server in (srv, plain):  # warm-up: every result buffer of the pools allocated
        for y in ys:
            server.submit(y)
        server.drain()
    sat_runs = {"batched": [], "plain": []}
    launches = 0
    for _ in range(2):
        for name, server in (("batched", srv), ("plain", plain)):
            it0 = server.iterations
            gc.collect()
            gc.disable()
            t0 = time.perf_counter()
            for y in ys:
                server.submit(y)
            server.drain()
            sat_runs[name].append(round(len(ys) / (time.perf_counter() - t0), 2))
            gc.enable()
            if name == "batched":
                launches = server.iterations - it0
    sat, sat_plain = max(sat_runs["batched"]), max(sat_runs["plain"])
    log(f"serve: saturation {sat:.1f} queries/s batched ({launches} launches for {len(ys)} queries; stamp clock "
        f"{srv.clock_hz / 1e6:.2f} MHz measured), {sat_plain:.1f} plain; runs {sat_runs}")

    # 3) Poisson arrivals at a sweep of rates (fractions of the plain
    # saturation): the same arrival trace for both forms, twice each in
    # alternating order, latencies pooled per form
    sweep = []
    rng = random.Random(args.seed + 10)
    _, pool_ys = queries(args.serve_queries)
    for k, frac in enumerate((0.25, 0.5, 0.75, 0.9, 0.97)):
        rate = frac * sat_plain
        offs = poisson_arrivals(rate, args.serve_queries, rng)
        row = {"rate_qps": round(rate, 2)}
        pooled = {"batched": [], "plain": []}
        walls = {"batched": 0.0, "plain": 0.0}
        replays = {"batched": [], "plain": []}
        order = (("batched", srv), ("plain", plain))
        for rep in range(2):
            for name, server in (order if (k + rep) % 2 == 0 else order[::-1]):
                server.completed.clear()
                it0, lp0 = server.iterations, server.launched_pairs
                gc.collect()  # no collector pause inside a replay (as the reference's time_call,
                gc.disable()  # benchmarks/bfv_client_matrix.py:40-44): near saturation one pause
                t0 = time.perf_counter()  # of a few ms becomes a backlog that takes seconds to drain
                reqs = server.run_trace(list(zip(offs, pool_ys)))
                wall = time.perf_counter() - t0
                gc.enable()
                pooled[name] += reqs
                walls[name] += wall
                rs = latency_stats(reqs)
                nl = server.iterations - it0
                replays[name].append({"p50_ms": round(rs["p50_ms"], 3), "p99_ms": round(rs["p99_ms"], 3),
                                      "launches": nl,
                                      "mean_batch": round((server.launched_pairs - lp0) / count / max(nl, 1), 2)})
        for name in ("batched", "plain"):
            st = latency_stats(pooled[name])
            row[name] = {"p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                         "mean_ms": round(st["mean_ms"], 3), "queries": st["n"],
                         "throughput_qps": round(len(pooled[name]) / walls[name], 2),
                         "replays": replays[name]}
        sweep.append(row)
        log(f"serve: {rate:7.1f} q/s: batched p50 {row['batched']['p50_ms']:.2f} p99 {row['batched']['p99_ms']:.2f} "
            f"ms | plain p50 {row['plain']['p50_ms']:.2f} p99 {row['plain']['p99_ms']:.2f} ms "
            f"({row['plain']['queries']} queries each); replays {replays}")
    return {
        "workload": f"{count} lookup-encrypted 512-bit vectors (key_len 1024, n^2 4096-bit) resident in HBM; "
                    "each query = one mulmod per index entry (encode_hamming_server)",
        "index_entries": count,
        "single_query_kernel_ms": round(kern_s * 1e3, 4),
        "mulmod_per_s": round(count / kern_s, 1),
        "roofline_valu": {"achieved": round(macs / kern_s / 1e12, 4), "peak": round(peak, 4),
                          "unit": "TMAC/s", "frac": round(macs / kern_s / 1e12 / peak, 4)},
        "roofline_hbm": {"achieved": round(hbm_bytes / kern_s / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
                         "frac": round(hbm_bytes / kern_s / 8e12, 4), "bytes_per_query": hbm_bytes},
        "pointwise": {"kernel": "k_pw_mulmod (per-item operands: one product, one fold, two REDC rows)",
                      "kernel_ms": round(pw_s * 1e3, 4), "mulmod_per_s": round(count / pw_s, 1),
                      "two_product_kernel_ms": round(pw_two_s * 1e3, 4),
                      "achieved_tmac_s": round(macs / pw_s / 1e12, 4),
                      "frac_of_peak": round(macs / pw_s / 1e12 / peak, 4),
                      "hbm_gb_s": round(count * 3 * W * 4 / pw_s / 1e9, 1), "entries_checked": 256},
        "cpu_baseline": cpu_eval,
        "saturation_qps_batched": round(sat, 2),
        "saturation_qps_plain": round(sat_plain, 2),
        "saturation_runs": sat_runs,
        "batching_kernel": batching_kernel,
        "saturation_queries": len(ys),
        "batched": "fifo: each launch takes the admitted queries not yet launched (up to 32; up to 64 admitted "
                   "and prepared) over the whole index, tiles query by query; a query completes at its last tile "
                   "(device clock stamp); a launch of one query is the plain launch",
        "plain": "one launch per query, whole index",
        "max_active": 64, "queries_per_launch_max": 32, "queries_per_rate": 2 * args.serve_queries,
        "sweep_method": "one Poisson trace per rate, replayed twice per form in alternating order, latencies pooled; "
                        "Python's collector off inside each replay",
        "latency_endpoint": "device-complete from arrival (plain: event after the query's launch; batched: the "
                            "stamp of the query's last tile, anchored to the event after its launch)",
        "poisson_sweep": sweep,
        "bit_exact": {"entries_checked": 256, "distances_checked": 256, "mismatches": 0},
    }

def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=65536, help="ciphertexts per step")
    ap.add_argument("--key-len", type=int, default=1024, help="prime bits (n has 2*key_len bits)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-sample", type=int, default=65536,
                    help="ciphertexts in the CPU-baseline sample (all processes; one thread times 2,048 of them)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-chunk", type=int, default=1024)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--tpi", type=int, default=0, help="force lanes per bignum (0 = automatic)")
    ap.add_argument("--window", type=int, default=0, help="force the window width (0 = automatic)")
    ap.add_argument("--legs", default="c1,ragged,lookup,serve",
                    help="extra workloads reported in the same line: c1, ragged, lookup, serve (comma list, or none)")
    ap.add_argument("--serve-index", type=int, default=1_000_000)
    ap.add_argument("--serve-queries", type=int, default=1000)
    ap.add_argument("--serve-cpu-sample", type=int, default=20000)
    ap.add_argument("--ragged-subs", type=int, default=256)
    ap.add_argument("--ragged-max", type=int, default=8192)
    ap.add_argument("--ragged-steps", type=int, default=3)
    ap.add_argument("--ragged-fixed", default="", metavar="VARIANT,WINDOW,WAVES_PER_CU",
                    help="skip the autotuner and use this launch shape")
    ap.add_argument("--lookup-vectors", type=int, default=100000)
    ap.add_argument("--lookup-steps", type=int, default=5)
    ap.add_argument("--lookup-cpu-sample", type=int, default=24)
    ap.add_argument("--lookup-gmp-sample", type=int, default=400)
    ap.add_argument("--lookup-client-vectors", type=int, default=20000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (counter calibration for profiles)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; every unit of work is independent and "
                                   "the multi-GPU split is a host-side scatter of those units",
                          "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import torch

    from cuhepy_amd.engine import Context, PaillierKey, crt_constants, stride_for
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    from cuhepy_amd.paillier import keygen as kg

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        return 1
    dev = torch.device("cuda:0")
    log(f"device {torch.cuda.get_device_name(0)}")

    rng = random.Random(args.seed)
    t = time.perf_counter()
    k = kg.paillier_key_gen(args.key_len, rng)
    n, phi, inv, p, q = k["pk"]["n"], k["sk"]["phi"], k["sk"]["inv"], k["p"], k["q"]
    log(f"key n={n.bit_length()} bits ({time.perf_counter() - t:.1f}s)")

    ctx = Context.default(0)
    ctx.use_torch_stream()
    key = PaillierKey(ctx, n, phi, inv, p, q)
    W, H = key.width, key.half
    count = args.batch
    stride = stride_for(count)
    if args.tpi or args.window:
        ctx.set_tuning(0, H, args.tpi, args.window)

    # seeded inputs: m uniform in [0, n), r uniform in Z_n^*
    t = time.perf_counter()
    ms = [rng.randrange(n) for _ in range(count)]
    rs = []
    while len(rs) < count:
        r = rng.randrange(1, n)
        if math.gcd(r, n) == 1:
            rs.append(r)
    m_aos = ints_to_limbs(ms, W)
    r_aos = ints_to_limbs(rs, W)
    log(f"inputs generated ({time.perf_counter() - t:.1f}s)")

    def to_il(aos: np.ndarray, limbs: int) -> torch.Tensor:
        src = torch.from_numpy(aos.view(np.int32)).to(dev)
        il = torch.zeros((limbs, stride), dtype=torch.int32, device=dev)
        ctx.aos_to_il(src, il, count, limbs, stride)
        return il

    def from_il(il: torch.Tensor, limbs: int) -> np.ndarray:
        aos = torch.empty((count, limbs), dtype=torch.int32, device=dev)
        ctx.il_to_aos(il, aos, count, limbs, stride)
        torch.cuda.synchronize()
        return aos.cpu().numpy().view(np.uint32)

    m_il, r_il = to_il(m_aos, W), to_il(r_aos, W)
    c_il = torch.zeros((W, stride), dtype=torch.int32, device=dev)
    t = time.perf_counter()
    key.encrypt_dev(m_il, r_il, c_il, count, stride)
    torch.cuda.synchronize()
    log(f"encrypted {count} ciphertexts on the GPU ({time.perf_counter() - t:.2f}s)")
    out_il = torch.zeros((H, stride), dtype=torch.int32, device=dev)

    # warmup
    for _ in range(args.warmup):
        key.decrypt_crt_dev(c_il, out_il, count, stride)
    torch.cuda.synchronize()
    log("warmup done")

    ctx.reset_timing()
    ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        key.decrypt_crt_dev(c_il, out_il, count, stride)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    half_ms, half_n = ctx.kernel_time("lazy_modexp")
    fin_ms, fin_n = ctx.kernel_time("crt_finish")
    ctx.set_timing(False)
    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    modexp_per_step = 2 * count
    value = modexp_per_step * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step, {value:,.0f} modexp/s "
        f"(exponentiation {half_ms / max(half_n, 1):.2f} ms, CRT finish {fin_ms / max(fin_n, 1):.2f} ms per step)")

    # full-size bit-exact check: CRT decrypt must return the seeded plaintexts
    got = from_il(out_il, H)
    want_aos = ints_to_limbs(ms, H)
    mism = int((got != want_aos).any(axis=1).sum())
    log(f"bit-exact check over {count} plaintexts: {mism} mismatches")
    if mism:
        print(json.dumps({"error": f"{mism} CRT decryptions differ from the seeded plaintexts"}))
        return 1

    # exact half-width encryption and decryption of this key (r^n and c^phi
    # mod p^2 and mod q^2, recombined modulo n^2) beside the same key without
    # factors (full width): kernel ms per batch, outputs compared in full
    full_key = PaillierKey(ctx, n, phi, inv)
    exact = {"items": count, "n_bits": n.bit_length()}
    outs = {}
    for form, kk in (("half_width", key), ("full_width", full_key)):
        e_il = torch.zeros((W, stride), dtype=torch.int32, device=dev)
        d_il = torch.zeros((W, stride), dtype=torch.int32, device=dev)
        for what, run, fin in (("encrypt", lambda: kk.encrypt_dev(m_il, r_il, e_il, count, stride), "encrypt_finish"),
                               ("decrypt", lambda: kk.decrypt_dev(c_il, d_il, count, stride), "decrypt_finish")):
            torch.cuda.synchronize()
            ctx.reset_timing()
            ctx.set_timing(True)
            run()
            torch.cuda.synchronize()
            exact[f"{what}_{form}_ms"] = round(ctx.kernel_time("lazy_modexp")[0] + ctx.kernel_time(fin)[0], 3)
            ctx.set_timing(False)
        outs[form] = (from_il(e_il, W), from_il(d_il, W))
    same = all(np.array_equal(a, b) for a, b in zip(outs["half_width"], outs["full_width"]))
    dec_ok = bool(np.array_equal(outs["half_width"][1], m_aos))
    log(f"exact half width vs full width, {count} items: encrypt {exact['encrypt_half_width_ms']:.2f} vs "
        f"{exact['encrypt_full_width_ms']:.2f} ms, decrypt {exact['decrypt_half_width_ms']:.2f} vs "
        f"{exact['decrypt_full_width_ms']:.2f} ms; outputs equal: {same}; decryptions = plaintexts: {dec_ok}")
    if not (same and dec_ok):
        print(json.dumps({"error": "half-width and full-width encryption or decryption differ"}))
        return 1
    exact["bit_exact"] = {"half_equals_full_items": count, "decrypt_equals_plaintexts": count, "mismatches": 0}
    # host entry point (key.decrypt: host buffers) of the key with factors:
    # all units take the textbook CRT pass only; with one non-unit (c = p) that
    # item alone is decrypted again in the exact form
    c_host = from_il(c_il, W)
    key.decrypt(c_host[:64])
    for name, batch in (("units", c_host), ("one_nonunit", c_host.copy())):
        if name == "one_nonunit":
            batch[0] = ints_to_limbs([p], W)[0]
        t = time.perf_counter()
        dh = key.decrypt(batch)
        exact[f"host_decrypt_{name}_ms"] = round((time.perf_counter() - t) * 1e3, 3)
        ok = bool(np.array_equal(dh[1:], m_aos[1:]))
        if name == "units":
            ok = ok and bool(np.array_equal(dh[:1], m_aos[:1]))
        else:
            ok = ok and bool(np.array_equal(dh[:1], full_key.decrypt(batch[:1])))
        if not ok:
            print(json.dumps({"error": f"host decrypt ({name}) differs"}))
            return 1
    log(f"host decrypt of {count} (PCIe included): units {exact['host_decrypt_units_ms']:.1f} ms, "
        f"with one non-unit {exact['host_decrypt_one_nonunit_ms']:.1f} ms")
    del full_key

    # roofline of the dominant kernel
    ebits = max((p - 1).bit_length(), (q - 1).bit_length())
    macs_per_modexp = (ebits + math.ceil(ebits / 6)) * 2 * H * H
    macs_per_launch = modexp_per_step * macs_per_modexp
    avg_launch_s = half_ms / max(half_n, 1) / 1e3
    achieved = macs_per_launch / avg_launch_s / 1e12
    props = torch.cuda.get_device_properties(0)
    cus = props.multi_processor_count
    # Peak: one v_mad_u64_u32 per lane per clock (the full VALU issue rate:
    # 64 lanes per CU per clock at the MI355X's 2.4 GHz engine clock). The
    # probe below (MAD-only chains, SGPR carry-out as compiled code uses it)
    # measures how close a pure MAD stream gets; mixed streams issue at the
    # same total rate, so the kernels are priced against the issue rate.
    peak = cus * 64 * CLOCK_HZ / 1e12
    s_probe, macs_probe = ctx.probe_rate(6, cus * 16, 20000)
    peak_probe = macs_probe / s_probe / 1e12
    log(f"lazy_modexp avg {avg_launch_s * 1e3:.2f} ms over {half_n} launches -> {achieved:.3f} TMAC/s = "
        f"{achieved / peak:.1%} of {peak:.2f} (peak issue rate); v_mad_u64_u32 probe {peak_probe:.3f} T/s on {cus} CUs")

    traffic, traffic_src = committed_traffic(f"k_lazy_modexp<{lazy_tag(key)}>")

    # plain per-call GPU form on the same batch (host buffers, synchronous calls)
    plain = None
    if not args.no_plain:
        c_aos = from_il(c_il, W)
        chunk = args.plain_chunk
        key.decrypt_crt(c_aos[:chunk])  # warm
        t = time.perf_counter()
        outs = [key.decrypt_crt(c_aos[i:i + chunk]) for i in range(0, count, chunk)]
        plain_s = time.perf_counter() - t
        plain_out = np.concatenate(outs)
        if not (plain_out == want_aos).all():
            print(json.dumps({"error": "plain per-call form differs from the seeded plaintexts"}))
            return 1
        plain = modexp_per_step / plain_s
        log(f"plain per-call form ({chunk}/call): {plain:,.0f} modexp/s")

    try:
        host_cores_visible = len(os.sched_getaffinity(0))
    except AttributeError:
        host_cores_visible = os.cpu_count() or 1
    cpu = None
    if not args.no_cpu:
        c_ints = limbs_to_ints(from_il(c_il, W)[: args.cpu_sample])
        cp = {"p": p, "q": q}
        cc = crt_constants(p, q)
        cp.update({"hp": cc["hp"], "hq": cc["hq"], "q_inv_p": cc["qinv"]})
        cores = min(16, host_cores_visible)
        t = time.perf_counter()
        cpu = cpu_baseline(cp, c_ints, ms[: args.cpu_sample], cores)
        log(f"cpu baseline ({cpu['engine']}): {cpu['single']:,.0f} modexp/s single thread, "
            f"{cpu['multi']:,.0f} modexp/s on {cpu['cores']} processes ({time.perf_counter() - t:.1f}s)")

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "modexp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded key, m uniform in [0,n), r uniform in Z_n*)",
        "config": {
            "workload": "Paillier CRT decrypt batch: 65,536 ciphertexts, n 2048-bit (n^2 4096-bit), "
                        "halves mod p^2 and q^2 at 2048 bits, 1024-bit exponents",
            "ciphertexts": count,
            "modexp_per_step": modexp_per_step,
            "n_bits": n.bit_length(),
            "half_modulus_bits": (p * p).bit_length(),
            "exponent_bits": ebits,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu-int",
            "achieved": round(achieved, 4),
            "peak": round(peak, 4),
            "unit": "TMAC/s (32x32->64 multiply-accumulates)",
            "frac": round(achieved / peak, 4),
            "peak_basis": f"{cus} CUs x 64 lanes x {CLOCK_HZ / 1e9:.1f} GHz, one v_mad_u64_u32 per lane per clock; "
                          f"MAD-only probe measured {peak_probe:.2f} T/s",
            "probe": round(peak_probe, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": "k_lazy_modexp (CRT halves, one launch)",
            "launch_ms": round(avg_launch_s * 1e3, 3),
            "macs_per_launch": macs_per_launch,
        },
        "plain_per_call_modexp_s": round(plain, 1) if plain else None,
        "batched_over_plain": round(value / plain, 3) if plain else None,
        "decrypts_per_s": round(value / 2, 1),
        "exact_encrypt_decrypt": exact,
        "bit_exact": True,
    }
    if cpu:
        line["cpu_baseline"] = {
            "value": round(cpu["multi"], 1),
            "unit": "modexp/s",
            "cores": cpu["cores"],
            "kind": "port",
            "sample": f"{cpu['n']} CRT decryptions of the same ciphertexts with {cpu['engine']} mpz_powm "
                      f"(one process per core); single thread {cpu['single']:.1f} modexp/s on "
                      f"{cpu['single_n']} of them",
            "single_thread": round(cpu["single"], 1),
            "model": host_model(),
            "host_cores_visible": host_cores_visible,
            "pool_note": "at most 16 processes: the host cores a one-GPU box allots",
        }
    legs = [x for x in args.legs.split(",") if x and x != "none"]
    if "c1" in legs:
        line["c1_encrypt_decrypt"] = c1_leg(ctx, args, min(16, host_cores_visible))
    if "ragged" in legs:
        line["ragged"] = ragged_leg(ctx, args, peak, torch.cuda.get_device_name(0))
    if "lookup" in legs:
        line["lookup_encrypt"] = lookup_leg(ctx, args, peak)
    if "serve" in legs:
        line["served_knn"] = serve_leg(ctx, args, peak)
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration * (1 << 20) // 4, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        # the same bytes with the lane-group pattern of the exponentiation and
        # mulmod kernels (k_copy_groups), one launch per (limbs, lanes) shape;
        # (128, 1): one lane per item, 64 consecutive items per limb row (the
        # fixed and pointwise kernels' index loads)
        for limbs, tpi in ((64, 4), (128, 8), (128, 1)):
            cnt = args.copy_calibration * (1 << 20) // (4 * limbs)
            st = stride_for(cnt)
            a = torch.ones((limbs, st), dtype=torch.int32, device=dev)
            b = torch.empty_like(a)
            torch.cuda.synchronize()
            ctx.probe_copy(a, b, cnt, st, limbs, tpi)
            torch.cuda.synchronize()
            del a, b
        log(f"copy calibration: 3 device copies of {args.copy_calibration} MiB, lane-group copies (64, 4), (128, 8), (128, 1)")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
