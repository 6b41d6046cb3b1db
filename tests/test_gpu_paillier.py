"""GPU parity of Paillier encrypt / decrypt / CRT decrypt / decode and the
lookup scheme against the oracle, on the golden vectors and seeded batches."""

import math
import random

import numpy as np
import pytest

from conftest import H
from oracle import paillier_oracle as O

pytestmark = pytest.mark.gpu

OP_MODEXP, OP_MULMOD, OP_LOOKUP = 0, 1, 2


@pytest.fixture(autouse=True)
def _reset_tuning(ctx):
    yield
    for limbs in (32, 64, 128):
        for op in (0, 1, 2, 3):
            ctx.set_tuning(op, limbs, 0, 0)


def _key_from_golden(ctx, g):
    from cuhepy_amd.engine import PaillierKey
    return PaillierKey(ctx, H(g["n"]), H(g["phi"]), H(g["inv"]), H(g["p"]), H(g["q"]))


@pytest.mark.parametrize("case", ["paillier_w64", "paillier_w128"])
def test_encrypt_decrypt_golden(ctx, golden, case):
    from cuhepy_amd.limbs import limbs_to_ints
    g = golden[case]
    key = _key_from_golden(ctx, g)
    ms, rs, cs = [H(v) for v in g["m"]], [H(v) for v in g["r"]], [H(v) for v in g["c"]]
    assert limbs_to_ints(key.encrypt(ms, rs)) == cs
    assert limbs_to_ints(key.decrypt(cs)) == ms
    assert limbs_to_ints(key.decrypt_crt(cs)) == ms
    edge = [H(v) for v in g["edge_c"]]
    assert limbs_to_ints(key.decrypt(edge)) == [H(v) for v in g["edge_dec"]]
    assert limbs_to_ints(key.decrypt_crt(edge)) == [H(v) for v in g["edge_crt"]]


@pytest.mark.parametrize("key_len", [256, 512, 1024])
def test_seeded_batches_every_variant(ctx, key_len):
    """encrypt / decrypt / CRT on seeded batches for every built lane-group size,
    with the key given its factors (half-width forms) and without them
    (full-width exponentiation and finishes)."""
    from cuhepy_amd.engine import Context, PaillierKey
    from cuhepy_amd.limbs import limbs_to_ints
    keys = O.key_gen(key_len, 300 + key_len)
    n = keys["pk"]["n"]
    key = PaillierKey(ctx, n, keys["sk"]["phi"], keys["sk"]["inv"], keys["p"], keys["q"])
    plain = PaillierKey(ctx, n, keys["sk"]["phi"], keys["sk"]["inv"])
    rng = random.Random(key_len)
    count = 97
    ms = [rng.randrange(n) for _ in range(count)]
    rs = [O.random_r(n, rng) for _ in range(count)]
    want_c = [O.encrypt(m, keys["pk"], r) for m, r in zip(ms, rs)]
    for k, widths in ((plain, (key.width,)), (key, (key.width, key.half) if key.crt else (key.width,))):
        for w in widths:  # the lane counts that steer this key's launches
            exp_tpis = Context.supported_tpi(w, op=OP_MODEXP)
            fin_tpis = [t for t in Context.supported_tpi(w, op=OP_MULMOD) if k is plain or w == key.width or t >= 4]
            for i in range(max(len(exp_tpis), len(fin_tpis))):
                ctx.set_tuning(OP_MODEXP, w, exp_tpis[i % len(exp_tpis)], 0)
                ctx.set_tuning(OP_MULMOD, w, fin_tpis[i % len(fin_tpis)], 0)
                tag = f"{'factors' if k.crt else 'no factors'} width {w}: exp tpi={exp_tpis[i % len(exp_tpis)]} " \
                      f"finish tpi={fin_tpis[i % len(fin_tpis)]}"
                assert limbs_to_ints(k.encrypt(ms, rs)) == want_c, f"encrypt {tag}"
                assert limbs_to_ints(k.decrypt(want_c)) == ms, f"decrypt {tag}"
            ctx.set_tuning(OP_MODEXP, w, 0, 0)
            ctx.set_tuning(OP_MULMOD, w, 0, 0)
    ctx.set_tuning(OP_MODEXP, key.width, 0, 0)
    ctx.set_tuning(OP_MULMOD, key.width, 0, 0)
    if not key.crt:  # 16-limb halves are not built: CRT is refused, plain decrypt covers the key
        with pytest.raises(ValueError):
            key.decrypt_crt(want_c)
        return
    exp_tpis = Context.supported_tpi(key.half, op=OP_MODEXP)
    fin_tpis = Context.supported_tpi(key.half, op=OP_MULMOD)
    for i in range(max(len(exp_tpis), len(fin_tpis))):
        ctx.set_tuning(OP_MODEXP, key.half, exp_tpis[i % len(exp_tpis)], 0)
        ctx.set_tuning(OP_MULMOD, key.half, fin_tpis[i % len(fin_tpis)], 0)
        assert limbs_to_ints(key.decrypt_crt(want_c)) == ms, f"crt variant {i}"


@pytest.mark.parametrize("key_len", [512, 1024])
def test_crt_full_width_ciphertext_reduction(ctx, key_len):
    """Each CRT half reduces the whole ciphertext c = c_lo + 2^(32 H) c_hi
    modulo p^2 (q^2) on entry: halves of all-ones and zero limbs, multiples of
    p^2 and q^2 (non-units), and values at the top of [0, n^2)."""
    from cuhepy_amd.engine import Context, PaillierKey
    from cuhepy_amd.limbs import limbs_to_ints
    keys = O.key_gen(key_len, 500 + key_len)
    n, p, q = keys["pk"]["n"], keys["p"], keys["q"]
    key = PaillierKey(ctx, n, keys["sk"]["phi"], keys["sk"]["inv"], p, q)
    n2, half = n * n, 32 * key.half
    lo = (1 << half) - 1
    rng = random.Random(key_len)
    cs = [lo, 1 << half, (1 << half) + lo, p * p, q * q, p * p - 1, q * q + 1, 7 * p * p, n2 - 1,
          ((n2 - 1) >> half << half) - 1, ((((n2 - 1) >> half) - 1) << half) | (lo >> 1),
          n2 - 1 - rng.randrange(1 << 64)]
    cs += [rng.randrange(n2 >> 1, n2) for _ in range(20)]
    assert all(0 <= c < n2 for c in cs)
    want = [O.decrypt_crt(c, keys) for c in cs]
    for tpi in Context.supported_tpi(key.half, op=OP_MODEXP):
        ctx.set_tuning(OP_MODEXP, key.half, tpi, 0)
        assert limbs_to_ints(key.decrypt_crt(cs)) == want, f"exp tpi={tpi}"
    ctx.set_tuning(OP_MODEXP, key.half, 0, 0)


def test_crt_rejects_factors_too_wide_for_the_halves(ctx):
    """p^2 and q^2 must fit in the CRT half width (H limbs, so p, q below
    2^(16 H)): a 1100-bit p with a 948-bit q (n of 2048 bits, H = 64) is
    refused instead of being truncated; the same n without factors still
    decrypts through the full-width path."""
    from cuhepy_amd.engine import PaillierKey
    from cuhepy_amd.limbs import limbs_to_ints
    rng = random.Random(77)
    p, q = O.get_prime(1100, rng), O.get_prime(948, rng)
    n = p * q
    phi = (p - 1) * (q - 1)
    inv = pow(phi, -1, n)
    assert 4090 < (n * n).bit_length() <= 4096
    with pytest.raises(ValueError):
        PaillierKey(ctx, n, phi, inv, p, q)
    key = PaillierKey(ctx, n, phi, inv)
    pk = {"n": n, "n_squared": n * n, "g": n + 1}
    m, r = rng.randrange(n), O.random_r(n, rng)
    c = O.encrypt(m, pk, r)
    assert limbs_to_ints(key.decrypt([c])) == [m]


def test_decode_hamming_fused(ctx, golden):
    """decrypt + odd-bit popcount on the GPU = plaintext Hamming distance."""
    from cuhepy_amd.engine import Modulus, PaillierKey
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    from cuhepy_amd.paillier.packing import embeddings_to_plaintexts, odd_mask
    keys = O.key_gen(512, 77)
    n, n2 = keys["pk"]["n"], keys["pk"]["n_squared"]
    key = PaillierKey(ctx, n, keys["sk"]["phi"], keys["sk"]["inv"], keys["p"], keys["q"])
    rng = random.Random(5)
    embed_len = 400
    A = np.array([[rng.getrandbits(1) for _ in range(embed_len)] for _ in range(33)], dtype=np.uint8)
    B = np.array([[rng.getrandbits(1) for _ in range(embed_len)] for _ in range(33)], dtype=np.uint8)
    ma = embeddings_to_plaintexts(A, 512, key.width)
    mb = embeddings_to_plaintexts(B, 512, key.width)
    ra = [O.random_r(n, rng) for _ in range(len(ma))]
    rb = [O.random_r(n, rng) for _ in range(len(mb))]
    ca, cb = key.encrypt(ma, ra), key.encrypt(mb, rb)
    prod = Modulus(ctx, n2, key.width).mulmod(ca, cb)
    counts = key.decode_hamming(prod, odd_mask(1024))
    want = [O.hamming(a, b) for a, b in zip(A.tolist(), B.tolist())]
    assert counts.tolist() == want
    # the oracle path over the same ciphertexts agrees
    assert [O.decode_hamming([O.decrypt(c, keys)], 1024) for c in limbs_to_ints(prod)] == want


def test_decrypt_rejects_out_of_range(ctx, golden):
    key = _key_from_golden(ctx, golden["paillier_w64"])
    with pytest.raises(ValueError):
        key.decrypt([key.n2])
    with pytest.raises(ValueError):
        key.encrypt([key.n], [1])


def _lookup_key(ctx, g):
    from cuhepy_amd.engine import LookupKey
    n, gg = H(g["n"]), H(g["g"])
    table = O.precompute_g_table(gg, n, g["key_len"])
    noise = O.precompute_noise_table(gg, n, [H(x) for x in g["noise_r"]])
    key = LookupKey(ctx, n, H(g["a"]), H(g["g_a_inv"]), [table[i] for i in range(len(table))], noise)
    return key, table, noise


def test_lookup_golden(ctx, golden):
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    from cuhepy_amd.paillier.packing import plaintexts_to_digits
    g = golden["lookup_w32"]
    key, _, _ = _lookup_key(ctx, g)
    ms = [H(v) for v in g["m"]]
    digits = plaintexts_to_digits(ints_to_limbs(ms, key.width), key.chunks)
    idx = np.array(g["noise_idx"], dtype=np.uint8)
    got = limbs_to_ints(key.encrypt_digits(digits, idx))
    assert got == [H(v) for v in g["c"]]
    assert limbs_to_ints(key.decrypt(got)) == ms


def test_lookup_seeded_every_variant(ctx):
    from cuhepy_amd.engine import Context, LookupKey
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    from cuhepy_amd.paillier.packing import plaintexts_to_digits
    keys = O.lookup_key_gen(512, 280, 4242, with_tables=True)
    n = keys["pk"]["n"]
    table = keys["g_table"]
    key = LookupKey(ctx, n, keys["sk"]["a"], keys["sk"]["g_a_inv"], [table[i] for i in range(len(table))],
                    keys["noise_table"])
    rng = random.Random(11)
    count = 75
    # sparse and dense digit patterns (rows where every item has digit 0 are skipped)
    ms = [0, n - 1] + [rng.randrange(n) for _ in range(count - 12)] + [rng.getrandbits(64) for _ in range(10)]
    idx = np.array([[rng.randrange(256) for _ in range(14)] for _ in ms], dtype=np.uint8)
    want = [O.lookup_encrypt(m, keys["pk"], table, keys["noise_table"], keys["message_chunks"], ix.tolist())
            for m, ix in zip(ms, idx)]
    digits = plaintexts_to_digits(ints_to_limbs(ms, key.width), key.chunks)
    exp_tpis = Context.supported_tpi(key.width, op=OP_MODEXP)
    lk_tpis = Context.supported_tpi(key.width, op=OP_LOOKUP)
    mul_tpis = Context.supported_tpi(key.width, op=OP_MULMOD)
    for i in range(max(len(exp_tpis), len(lk_tpis), len(mul_tpis))):
        tpi = lk_tpis[i % len(lk_tpis)]
        ctx.set_tuning(OP_LOOKUP, key.width, tpi, 0)
        ctx.set_tuning(OP_MULMOD, key.width, mul_tpis[i % len(mul_tpis)], 0)
        ctx.set_tuning(OP_MODEXP, key.width, exp_tpis[i % len(exp_tpis)], 0)
        got = limbs_to_ints(key.encrypt_digits(digits, idx))
        assert got == want, f"tpi={tpi}"
        assert limbs_to_ints(key.decrypt(got)) == ms, f"decrypt tpi={tpi}"


def test_lookup_row_groups_packed_and_mixed(ctx):
    """Packed embeddings take the row-group tables (one product per two, three
    or four 8-bit digits; 128 rows leave two single rows after 42 triples);
    waves with a message outside the packing alphabet fall back to single
    rows. All must equal the oracle: quads (window 0/4), triples (3), pairs
    (2), single rows (1)."""
    from cuhepy_amd.engine import LookupKey
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    from cuhepy_amd.paillier.packing import plaintexts_to_digits
    keys = O.lookup_key_gen(512, 280, 4242, with_tables=True)
    n = keys["pk"]["n"]
    table = keys["g_table"]
    key = LookupKey(ctx, n, keys["sk"]["a"], keys["sk"]["g_a_inv"], [table[i] for i in range(len(table))],
                    keys["noise_table"])
    rng = random.Random(12)
    embed_len = 512
    embs = [[rng.getrandbits(1) for _ in range(embed_len)] for _ in range(70)]
    embs += [[0] * embed_len, [1] * embed_len, [1] + [0] * (embed_len - 1)]
    packed = [O.embed_to_plaintexts(e, embed_len, 512)[0] for e in embs]
    # every 16th message is an arbitrary residue: its wave uses single rows
    mixed = [rng.randrange(n) if i % 16 == 5 else m for i, m in enumerate(packed)]
    for ms in (packed, mixed):
        idx = np.array([[rng.randrange(256) for _ in range(14)] for _ in ms], dtype=np.uint8)
        want = [O.lookup_encrypt(m, keys["pk"], table, keys["noise_table"], keys["message_chunks"], ix.tolist())
                for m, ix in zip(ms, idx)]
        digits = plaintexts_to_digits(ints_to_limbs(ms, key.width), key.chunks)
        if ms is packed:
            assert not (digits & 0xAA).any()  # the packing alphabet: the pair path applies
        for rows in (0, 4, 3, 2, 1):
            ctx.set_tuning(OP_LOOKUP, key.width, 0, rows)
            got = limbs_to_ints(key.encrypt_digits(digits, idx))
            assert got == want, f"rows per product {rows}"
        assert limbs_to_ints(key.decrypt(got)) == ms
    # noise factors go two per product: an odd count leaves one single. The
    # oracle takes the reference's 14 factors; dividing its result by the
    # noise entries left out gives the expected value for fewer.
    ctx.set_tuning(OP_LOOKUP, key.width, 0, 0)
    n2, nt = keys["pk"]["n_squared"], keys["noise_table"]
    idx14 = np.array([[rng.randrange(256) for _ in range(14)] for _ in packed], dtype=np.uint8)
    full = [O.lookup_encrypt(m, keys["pk"], table, nt, keys["message_chunks"], ix.tolist())
            for m, ix in zip(packed, idx14)]
    digits = plaintexts_to_digits(ints_to_limbs(packed, key.width), key.chunks)
    for nm in (13, 1):
        want = []
        for c, ix in zip(full, idx14):
            for j in ix[nm:]:
                c = c * pow(nt[int(j) % len(nt)], -1, n2) % n2
            want.append(c)
        got = limbs_to_ints(key.encrypt_digits(digits, np.ascontiguousarray(idx14[:, :nm])))
        assert got == want, f"{nm} noise factors"


def test_lookup_modulus_at_the_top_of_the_width(ctx):
    """n = 2^2048 - 159, so n^2 lies just below 2^4096 (128 limbs): the last
    product of the lookup encryption (a plain noise factor; R' = 2^4104 leaves
    little room above n^2) lands at or above 2^4096 for about 0.1% of the
    items, before the final subtraction, whose carry word must then be the
    bit above the limbs. 10,000 items give ~10 such items. The encryption
    formula holds for any odd n, any g and any number of 8-bit rows: 8 rows
    (messages below 2^64, packed and arbitrary) keep the oracle cheap; quads
    and single rows against it."""
    from cuhepy_amd.engine import LookupKey
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    from cuhepy_amd.paillier.packing import plaintexts_to_digits
    n = (1 << 2048) - 159
    n2 = n * n
    assert n2.bit_length() == 4096 and n2 > (1 << 4096) - (1 << 4060)
    rng = random.Random(4096)
    g = rng.randrange(2, n2)
    table = O.precompute_g_table(g, n, 32)  # 8 rows
    noise = [rng.randrange(1, n2) for _ in range(256)]
    key = LookupKey(ctx, n, 1, 1, [table[i] for i in range(len(table))], noise)
    assert key.width == 128 and key.chunks == 8
    spread = lambda v: sum(((v >> b) & 1) << (2 * b) for b in range(32))  # noqa: E731
    ms = [spread(rng.getrandbits(32)) for _ in range(4000)]  # packed: every digit in the alphabet
    ms += [rng.getrandbits(64) for _ in range(3000)] + [0] * 3000
    idx = np.array([[rng.randrange(256) for _ in range(14)] for _ in ms], dtype=np.uint8)
    pk = {"n": n, "n_squared": n2}
    want = [O.lookup_encrypt(m, pk, table, noise, 8, ix.tolist()) for m, ix in zip(ms, idx)]
    digits = plaintexts_to_digits(ints_to_limbs(ms, key.width), key.chunks)
    for rows in (0, 1):
        ctx.set_tuning(OP_LOOKUP, key.width, 0, rows)
        got = limbs_to_ints(key.encrypt_digits(digits, idx))
        bad = [i for i, (x, y) in enumerate(zip(got, want)) if x != y]
        assert not bad, f"rows per product {rows}: {len(bad)} items differ, first {bad[:4]}"


def test_plain_encrypt_decrypt_1024_messages():
    """The reference's CPU-runnable case: key_len 1024-bit n (512-bit primes,
    n^2 2048-bit), keypair from seeded primes, 1024 messages m uniform in [0, n)
    and r uniform in Z_n*, one batch: every ciphertext equals (1+n)^m r^n mod
    n^2 computed with GMP (the engine gmpy2 wraps) and every decryption, plain
    and CRT, returns m (scheme.py:63-104)."""
    from cuhepy_amd.engine import Context, PaillierKey
    from cuhepy_amd.limbs import limbs_to_ints
    from oracle import gmp_engine as G
    powm = G.powm if G.available() else pow
    keys = O.key_gen(512, 20241)
    n, n2 = keys["pk"]["n"], keys["pk"]["n_squared"]
    assert n.bit_length() == 1024
    key = PaillierKey(Context.default(0), n, keys["sk"]["phi"], keys["sk"]["inv"], keys["p"], keys["q"])
    rng = random.Random(1024)
    ms = [rng.randrange(n) for _ in range(1024)]
    rs = [O.random_r(n, rng) for _ in range(1024)]
    cs = limbs_to_ints(key.encrypt(ms, rs))
    assert cs == [(1 + m * n) * powm(r, n, n2) % n2 for m, r in zip(ms, rs)]
    assert cs[:8] == [O.encrypt(m, keys["pk"], r) for m, r in zip(ms[:8], rs[:8])]
    assert limbs_to_ints(key.decrypt(cs)) == ms
    assert limbs_to_ints(key.decrypt_crt(cs)) == ms


def test_lookup_reencrypt_on_device(ctx):
    """Fused re-encryption (lcu:1053-1135): decrypt under key A, digits taken
    on the device, encryption under key B's tables. With the same noise
    indices every result equals the oracle's lookup encryption of the
    oracle's decryption, for packed messages (row groups) and arbitrary ones
    (single rows), at every lookup window, through the host and the device
    entry points."""
    import torch
    from cuhepy_amd.engine import LookupKey
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    ka_o = O.lookup_key_gen(512, 280, 4242, with_tables=True)
    kb_o = O.lookup_key_gen(512, 280, 4343, with_tables=True)
    na, nb = ka_o["pk"]["n"], kb_o["pk"]["n"]

    def mk(k):
        t = k["g_table"]
        return LookupKey(ctx, k["pk"]["n"], k["sk"]["a"], k["sk"]["g_a_inv"], [t[i] for i in range(len(t))],
                         k["noise_table"])

    ka, kb = mk(ka_o), mk(kb_o)
    assert ka.chunks == kb.chunks == 128
    rng = random.Random(31)
    embs = [[rng.getrandbits(1) for _ in range(512)] for _ in range(40)]
    packed = [O.embed_to_plaintexts(e, 512, 512)[0] for e in embs]
    top = min(na, nb)
    arbitrary = [0, 1, top - 1] + [rng.randrange(top) for _ in range(37)]
    for ms in (arbitrary, packed):  # arbitrary first: key B builds no group table for them
        ia = np.array([[rng.randrange(256) for _ in range(14)] for _ in ms], dtype=np.uint8)
        cts = [O.lookup_encrypt(m, ka_o["pk"], ka_o["g_table"], ka_o["noise_table"], ka_o["message_chunks"],
                                ix.tolist()) for m, ix in zip(ms, ia)]
        assert [O.lookup_decrypt(c, ka_o) for c in cts] == ms
        ib = np.array([[rng.randrange(256) for _ in range(14)] for _ in ms], dtype=np.uint8)
        want = [O.lookup_encrypt(m, kb_o["pk"], kb_o["g_table"], kb_o["noise_table"], kb_o["message_chunks"],
                                 ix.tolist()) for m, ix in zip(ms, ib)]
        for rows in (0, 3, 1):
            ctx.set_tuning(OP_LOOKUP, kb.width, 0, rows)
            got = limbs_to_ints(ka.reencrypt(cts, kb, ib))
            assert got == want, f"rows per product {rows}"
        # device entry point: interleaved limbs with a padded stride
        dev = torch.device("cuda:0")
        count = len(ms)
        stride = (count + 63) // 64 * 64 + 64  # whole 64-item columns, one spare
        aos = torch.from_numpy(ints_to_limbs(cts, ka.width).view(np.int32).copy()).to(dev)
        il = torch.zeros((ka.width, stride), dtype=torch.int32, device=dev)
        out_il = torch.full((kb.width, stride), -1, dtype=torch.int32, device=dev)
        out_aos = torch.zeros((count, kb.width), dtype=torch.int32, device=dev)
        nidx = torch.from_numpy(ib.copy()).to(dev)
        torch.cuda.synchronize()
        ctx.aos_to_il(aos, il, count, ka.width, stride)
        ka.reencrypt_dev(il, kb, nidx, 14, out_il, count, stride)
        ctx.il_to_aos(out_il, out_aos, count, kb.width, stride)
        ctx.sync()
        assert limbs_to_ints(out_aos.cpu().numpy().view(np.uint32)) == want
        assert (out_il[:, count:] == -1).all()  # the padding past count is not written
        assert limbs_to_ints(kb.decrypt(want[:5])) == ms[:5]


def test_lookup_reencrypt_rejects_narrow_targets(ctx):
    """A target whose message chunks cannot hold every digit of the source's
    plaintexts is refused (no digit is dropped), as are bad noise indices."""
    from cuhepy_amd.engine import LookupKey
    ka_o = O.lookup_key_gen(512, 280, 4242, with_tables=True)
    n = ka_o["pk"]["n"]
    t = ka_o["g_table"]
    ka = LookupKey(ctx, n, ka_o["sk"]["a"], ka_o["sk"]["g_a_inv"], [t[i] for i in range(len(t))],
                   ka_o["noise_table"])
    narrow = LookupKey(ctx, n, 1, 1, [t[i] for i in range(60)], ka_o["noise_table"])  # 480 bits of digits
    c = [O.lookup_encrypt(5, ka_o["pk"], t, ka_o["noise_table"], ka_o["message_chunks"], [0] * 14)]
    with pytest.raises(ValueError, match="message chunks"):
        ka.reencrypt(c, narrow, np.zeros((1, 14), dtype=np.uint8))
    small_noise = LookupKey(ctx, n, 1, 1, [t[i] for i in range(len(t))], ka_o["noise_table"][:16])
    with pytest.raises(ValueError, match="noise index"):
        ka.reencrypt(c, small_noise, np.full((1, 14), 16, dtype=np.uint8))


@pytest.mark.parametrize("key_len", [512, 1024])
def test_lookup_crt_decrypt_matches_full_width(ctx, key_len):
    """Half-width lookup decryption (c^a mod p^2 and mod q^2, recombined to
    c^a mod n^2 before L) against the oracle's full-width formula
    (lookup.py:257-278) and the key's own full-width form: valid ciphertexts
    and arbitrary c in [0, n^2) -- zero, one, multiples of p^2 and q^2
    (non-units), all-ones halves, the top of the range -- at every lane count
    of the half-width exponentiation and of the finish; the fused odd-bit
    count as well."""
    from cuhepy_amd.engine import Context, LookupKey
    from cuhepy_amd.limbs import limbs_to_ints
    keys = O.lookup_key_gen(key_len, 280, 900 + key_len, with_tables=False)
    n, n2, p, q = keys["pk"]["n"], keys["pk"]["n_squared"], keys["p"], keys["q"]
    g, g_n = keys["pk"]["g"], keys["pk"]["g_n"]
    a, gai = keys["sk"]["a"], keys["sk"]["g_a_inv"]
    full = LookupKey(ctx, n, a, gai, [], [])
    crt = LookupKey(ctx, n, a, gai, [], [])
    assert not crt.crt
    crt.set_crt(p, q)
    assert crt.crt and not full.crt
    W = crt.width
    half = 16 * W  # bits per half
    rng = random.Random(key_len)
    ms = [0, 1, n - 1] + [rng.randrange(n) for _ in range(21)]
    valid = [pow(g, m, n2) * pow(g_n, O.lookup_random_r(n, rng), n2) % n2 for m in ms]
    lo = (1 << half) - 1
    odd = [0, 1, p * p, q * q, 5 * p * p, 3 * q * q, lo, (1 << half) + lo, n2 - 1, n2 - 2, n
           ] + [rng.randrange(n2) for _ in range(13)]
    cs = valid + odd
    assert all(0 <= c < n2 for c in cs)
    want = [O.lookup_decrypt(c, keys) for c in cs]
    assert want[:len(ms)] == ms
    assert limbs_to_ints(full.decrypt(cs)) == want
    exp_tpis = Context.supported_tpi(W // 2, op=OP_MODEXP)
    fin_tpis = Context.supported_tpi(W, op=OP_MULMOD)
    for i in range(max(len(exp_tpis), len(fin_tpis))):
        ctx.set_tuning(OP_MODEXP, W // 2, exp_tpis[i % len(exp_tpis)], 0)
        ctx.set_tuning(OP_MULMOD, W, fin_tpis[i % len(fin_tpis)], 0)
        assert limbs_to_ints(crt.decrypt(cs)) == want, f"variant {i}"
    # the textbook route's finish (k_lookup_crt_fast) at every lane count of
    # the half width, the invalid and non-unit items through the exact one
    ctx.set_tuning(OP_MODEXP, W // 2, 0, 0)
    ctx.set_tuning(OP_MULMOD, W, 0, 0)
    for t in Context.supported_tpi(W // 2, op=OP_MULMOD):
        ctx.set_tuning(OP_MULMOD, W // 2, t, 0)
        assert limbs_to_ints(crt.decrypt(cs)) == want, f"fast finish tpi {t}"
        assert limbs_to_ints(crt.decrypt(cs[:len(ms)])) == ms, f"valid only, fast finish tpi {t}"
    ctx.set_tuning(OP_MULMOD, W // 2, 0, 0)
    mask = sum(1 << b for b in range(1, 32 * W, 2))
    counts = crt.decode_hamming(cs, mask)
    assert [int(x) for x in counts] == [bin(m & mask).count("1") for m in want]


def test_lookup_crt_refuses_bad_factors(ctx):
    """set_crt checks p q = n and leaves the key in its full-width form when
    refused; a second call is refused too."""
    from cuhepy_amd.engine import LookupKey
    from cuhepy_amd.limbs import limbs_to_ints
    keys = O.lookup_key_gen(512, 280, 1905, with_tables=False)
    n, p, q = keys["pk"]["n"], keys["p"], keys["q"]
    key = LookupKey(ctx, n, keys["sk"]["a"], keys["sk"]["g_a_inv"], [], [])
    with pytest.raises(ValueError):
        key.set_crt(p + 2, q)
    assert not key.crt
    c = random.Random(5).randrange(n * n)
    assert limbs_to_ints(key.decrypt([c])) == [O.lookup_decrypt(c, keys)]
    key.set_crt(q, p)
    assert key.crt
    with pytest.raises(ValueError):
        key.set_crt(p, q)
    assert limbs_to_ints(key.decrypt([c])) == [O.lookup_decrypt(c, keys)]


@pytest.mark.parametrize("key_len", [512, 1024])
def test_paillier_exact_crt_decrypt_matches_full_width(ctx, key_len):
    """A key with p and q encrypts, decrypts and decodes at half width (c^phi mod p^2
    and mod q^2, recombined to c^phi mod n^2 before L): equal to the
    oracle's scheme.py:86-104 formula and to a key without factors for
    valid ciphertexts and for non-units (multiples of p, p^2, q^2, zero),
    where the textbook CRT formula (decrypt_crt) is not bound to agree; at
    every lane count of the half-width exponentiation and of the finish."""
    from cuhepy_amd.engine import Context, PaillierKey
    from cuhepy_amd.limbs import limbs_to_ints
    keys = O.key_gen(key_len, 700 + key_len)
    n, p, q = keys["pk"]["n"], keys["p"], keys["q"]
    phi, inv = keys["sk"]["phi"], keys["sk"]["inv"]
    crt = PaillierKey(ctx, n, phi, inv, p, q)
    full = PaillierKey(ctx, n, phi, inv)
    assert crt.exact_crt and not full.exact_crt
    n2, W = n * n, crt.width
    rng = random.Random(key_len + 3)
    ms = [0, 1, n - 1] + [rng.randrange(n) for _ in range(13)]
    cs = [O.encrypt(m, keys["pk"], O.random_r(n, rng)) for m in ms]
    lo = (1 << (16 * W)) - 1
    cs += [0, 1, p, 3 * q, p * p, q * q, 11 * p * p, lo, n2 - 1, n] + [rng.randrange(n2) for _ in range(6)]
    want = [O.decrypt(c, keys) for c in cs]
    assert want[:len(ms)] == ms
    assert limbs_to_ints(full.decrypt(cs)) == want
    exp_tpis = Context.supported_tpi(W // 2, op=OP_MODEXP)
    fin_tpis = Context.supported_tpi(W, op=OP_MULMOD)
    for i in range(max(len(exp_tpis), len(fin_tpis))):
        ctx.set_tuning(OP_MODEXP, W // 2, exp_tpis[i % len(exp_tpis)], 0)
        ctx.set_tuning(OP_MULMOD, W, fin_tpis[i % len(fin_tpis)], 0)
        assert limbs_to_ints(crt.decrypt(cs)) == want, f"variant {i}"
    mask = sum(1 << b for b in range(1, 32 * W, 2))
    assert [int(x) for x in crt.decode_hamming(cs, mask)] == [bin(m & mask).count("1") for m in want]
    # host entry points: units alone take the textbook CRT pass only; the
    # non-units of a batch are listed by the finish and only they are
    # decrypted again in the exact form (a batch with more than a quarter of
    # non-units takes the exact form whole); at every lane count of the
    # half-width finish that lists them
    nv = len(ms)
    nonunit = [math.gcd(c, n) != 1 for c in cs]
    few = cs[:nv] + [c for c, b in zip(cs, nonunit) if b][:3]  # 3 of 19: listed and redone
    many = cs[nv:]  # more than a quarter non-units: one exact pass
    assert sum(nonunit[nv:]) * 4 > len(many)
    few_want = [O.decrypt(c, keys) for c in few]
    for t in Context.supported_tpi(W // 2, op=OP_MULMOD):
        if t < 4:  # the finish splits a lane group into p- and q-halves
            continue
        ctx.set_tuning(OP_MULMOD, W // 2, t, 0)
        for batch, bwant, passes in ((cs[:nv], ms, 0), (few, few_want, 1), (many, want[nv:], 1), (cs, want, 1)):
            ctx.reset_timing()
            ctx.set_timing(True)
            got = limbs_to_ints(crt.decrypt(batch))
            ctx.set_timing(False)
            assert got == bwant, f"finish tpi {t}"
            assert ctx.kernel_time("crt_finish")[1] == 1 and ctx.kernel_time("decrypt_finish")[1] == passes
            assert [int(x) for x in crt.decode_hamming(batch, mask)] == [bin(m & mask).count("1") for m in bwant]
    ctx.set_tuning(OP_MULMOD, W // 2, 0, 0)
    # encryption: r^n mod p^2 and mod q^2 at half width, recombined; r = 1,
    # n - 1, non-units (multiples of p and q), random
    rs = [1, n - 1, p, 2 * q, 3 * p] + [O.random_r(n, rng) for _ in range(len(ms) - 5)]
    g = keys["pk"]["g"]
    # the oracle refuses non-unit r (scheme.py draws units); for those the closed form g^m r^n mod n^2
    enc_want = [O.encrypt(m, keys["pk"], r) if math.gcd(r, n) == 1 else pow(g, m, n2) * pow(r, n, n2) % n2
                for m, r in zip(ms, rs)]
    assert limbs_to_ints(full.encrypt(ms, rs)) == enc_want
    for i in range(max(len(exp_tpis), len(fin_tpis))):
        ctx.set_tuning(OP_MODEXP, W // 2, exp_tpis[i % len(exp_tpis)], 0)
        ctx.set_tuning(OP_MULMOD, W, fin_tpis[i % len(fin_tpis)], 0)
        assert limbs_to_ints(crt.encrypt(ms, rs)) == enc_want, f"encrypt variant {i}"
