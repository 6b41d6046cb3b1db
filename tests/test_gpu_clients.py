"""The drop-in client classes, mirroring tests/unit/test_GPU_paillier_clients.py
of the reference (encrypt -> evaluate -> decode = plaintext Hamming distance,
key reload, CPU <-> GPU interop, table dump/reload), plus ciphertext-level
bit-exactness against the oracle where the randomness is injected."""

import json
import random

import numpy as np
import pytest

from oracle import paillier_oracle as O

pytestmark = pytest.mark.gpu

_EMBED_LEN = 8
_KEY_LEN = 1024
_VECTORS = [  # test_GPU_paillier_clients.py:12-16
    [0, 1, 0, 1, 1, 0, 1, 0],
    [1, 1, 0, 0, 1, 0, 0, 1],
    [1, 0, 1, 0, 1, 1, 0, 0],
]


def _hamming(lhs, rhs):
    return sum(int(a != b) for a, b in zip(lhs, rhs))


def _classes():
    from cuhepy_amd.paillier._gpu_ext import PaillierGPUClient
    from cuhepy_amd.paillier._lookup_gpu_ext import PaillierGPULookupClient
    return {"paillier": PaillierGPUClient, "paillier_lookup": PaillierGPULookupClient}


def _to_int(c):
    return int(c, 16) if isinstance(c, str) else int(c)


def _reload(client, cls, lookup):
    config = json.loads(client.stringify_config())
    kwargs = {"embed_len": _EMBED_LEN, "key_len": _KEY_LEN}
    if lookup:
        kwargs["alpha_len"] = config["alpha_len"]
    clone = cls(**kwargs, skip_key_gen=True)
    clone.load_stringified_keys(client.stringify_pk(), client.stringify_sk())
    if lookup:
        clone.load_config(config, client.dump_tables())
    else:
        clone.load_config(config)
    return clone


@pytest.fixture(scope="module")
def clients():
    cls = _classes()
    rng = random.Random(2718)
    return {"paillier": cls["paillier"](embed_len=_EMBED_LEN, key_len=_KEY_LEN, rng=rng),
            "paillier_lookup": cls["paillier_lookup"](embed_len=_EMBED_LEN, key_len=_KEY_LEN, rng=rng)}


@pytest.mark.parametrize("name", ["paillier", "paillier_lookup"])
def test_distinct_keypairs(name):
    cls = _classes()[name]
    first = cls(embed_len=_EMBED_LEN, key_len=512)
    second = cls(embed_len=_EMBED_LEN, key_len=512)
    assert json.loads(first.stringify_pk()) != json.loads(second.stringify_pk())
    assert json.loads(first.stringify_sk()) != json.loads(second.stringify_sk())


@pytest.mark.parametrize("name", ["paillier", "paillier_lookup"])
def test_encrypt_hamming_and_key_roundtrip(clients, name):
    client = clients[name]
    lookup = name == "paillier_lookup"
    cls = type(client)
    config = json.loads(client.stringify_config())
    assert config["embed_len"] == _EMBED_LEN and config["key_len"] == _KEY_LEN

    one = client.encrypt([_VECTORS[0]])[0]
    batch = client.encrypt(_VECTORS)
    assert len(batch) == len(_VECTORS) and len(one) == len(batch[0]) == client.chunk_num == 1
    pk = json.loads(client.stringify_pk())
    pk_arg = {"n_squared": format(int(pk["n_squared"]), "x")} if not lookup else pk
    enc = cls.encode_hamming_server(one, batch[1], pk_arg)
    assert client.decode_hamming_client([enc]) == [_hamming(_VECTORS[0], _VECTORS[1])]
    encs = [cls.encode_hamming_server(batch[0], batch[1], pk_arg),
            cls.encode_hamming_server(batch[1], batch[2], pk_arg)]
    assert client.decode_hamming_client(encs) == [_hamming(_VECTORS[0], _VECTORS[1]),
                                                  _hamming(_VECTORS[1], _VECTORS[2])]
    # client-side evaluation gives the same ciphertexts as the server-side one
    assert client.encode_hamming_client(batch[0], batch[1]) == encs[0]

    clone = _reload(client, cls, lookup)
    enc2 = cls.encode_hamming_server(clone.encrypt([_VECTORS[0]])[0], clone.encrypt([_VECTORS[1]])[0], pk_arg)
    assert json.loads(clone.stringify_pk()) == pk
    assert json.loads(clone.stringify_sk()) == json.loads(client.stringify_sk())
    assert json.loads(clone.stringify_config()) == config
    assert clone.decode_hamming_client([enc2]) == [_hamming(_VECTORS[0], _VECTORS[1])]


def _oracle_keys(client, lookup):
    pk, sk = json.loads(client.stringify_pk()), json.loads(client.stringify_sk())
    n = int(pk["n"])
    if lookup:
        return {"pk": {"g": int(pk["g"]), "n": n, "n_squared": n * n, "g_n": int(pk["g_n"])},
                "sk": {"phi": int(sk["phi"]), "a": int(sk["a"]), "g_a_inv": int(sk["g_a_inv"])}}
    return {"pk": {"g": n + 1, "n": n, "n_squared": n * n}, "sk": {"phi": int(sk["phi"]), "inv": int(sk["inv"])}}


@pytest.mark.parametrize("name", ["paillier", "paillier_lookup"])
def test_cross_device_interop(clients, name):
    """Ciphertexts from the CPU scheme (oracle) decode on the GPU and GPU
    ciphertexts decrypt with the CPU scheme (test_GPU_paillier_clients.py:99-128)."""
    client = clients[name]
    lookup = name == "paillier_lookup"
    keys = _oracle_keys(client, lookup)
    n, n2 = keys["pk"]["n"], keys["pk"]["n_squared"]
    rng = random.Random(3)
    chunk_len = 2 * _KEY_LEN
    expected = _hamming(_VECTORS[0], _VECTORS[1])
    # CPU -> GPU
    ma = O.embed_to_plaintexts(_VECTORS[0], _EMBED_LEN, _KEY_LEN)[0]
    mb = O.embed_to_plaintexts(_VECTORS[1], _EMBED_LEN, _KEY_LEN)[0]
    if lookup:
        tables = client.dump_tables()
        gt = {int(i): [int(v) for v in row] for i, row in tables["g_table"].items()}
        nt = [int(v) for v in tables["noise_table"]]
        mc = json.loads(client.stringify_config())["message_chunks"]
        ca = O.lookup_encrypt(ma, keys["pk"], gt, nt, mc, [rng.randrange(256) for _ in range(14)])
        cb = O.lookup_encrypt(mb, keys["pk"], gt, nt, mc, [rng.randrange(256) for _ in range(14)])
    else:
        ca = O.encrypt(ma, keys["pk"], O.random_r(n, rng))
        cb = O.encrypt(mb, keys["pk"], O.random_r(n, rng))
    prod = O.add(ca, cb, keys["pk"])
    arg = [prod] if lookup else [format(prod, "x")]
    assert client.decode_hamming_client([arg]) == [expected]
    # GPU -> CPU
    ga = _to_int(client.encrypt([_VECTORS[0]])[0][0])
    gb = _to_int(client.encrypt([_VECTORS[1]])[0][0])
    s = O.add(ga, gb, keys["pk"])
    dec = O.lookup_decrypt(s, keys) if lookup else O.decrypt(s, keys)
    assert O.decode_hamming([dec], chunk_len) == expected


def test_plain_encrypt_cts_bit_exact(clients):
    """Injected r: GPU ciphertexts equal scheme.py's formula bit for bit."""
    client = clients["paillier"]
    keys = _oracle_keys(client, False)
    n = keys["pk"]["n"]
    rng = random.Random(9)
    ms = [0, 1, n - 1] + [rng.randrange(n) for _ in range(29)]
    rs = [O.random_r(n, rng) for _ in ms]
    got = client.encrypt_cts(ms, rs)
    assert got == [O.encrypt(m, keys["pk"], r) for m, r in zip(ms, rs)]
    assert client.decrypt_cts(got) == ms
    assert client.decrypt_cts_crt(got) == ms


def test_lookup_encrypt_cts_bit_exact_and_tables(clients):
    """Injected noise indices: GPU lookup ciphertexts equal lookup.py's
    formula; the dumped tables match their definition."""
    client = clients["paillier_lookup"]
    keys = _oracle_keys(client, True)
    n, n2, g = keys["pk"]["n"], keys["pk"]["n_squared"], keys["pk"]["g"]
    tables = client.dump_tables()
    gt = {int(i): [int(v) for v in row] for i, row in tables["g_table"].items()}
    nt = [int(v) for v in tables["noise_table"]]
    mc = json.loads(client.stringify_config())["message_chunks"]
    assert mc == 256 and len(gt) == mc and len(nt) == 256
    for i, j in ((0, 0), (0, 1), (3, 77), (255, 255), (128, 2)):
        assert gt[i][j] == pow(g, j * 2 ** (8 * i), n2)
    g_n = keys["pk"]["g_n"]
    assert g_n == pow(g, n, n2)
    # noise entries are n-th-power residues of g: noise^(a) = 1 (order divides a)
    assert all(pow(v, keys["sk"]["a"], n2) == 1 for v in nt[:8])
    rng = random.Random(12)
    ms = [0, n - 1, 1 << 1000] + [rng.randrange(n) for _ in range(13)]
    idx = np.array([[rng.randrange(256) for _ in range(14)] for _ in ms], dtype=np.uint8)
    got = [int(c) for c in client.encrypt_cts(ms, noise_idx=idx)]
    assert got == [O.lookup_encrypt(m, keys["pk"], gt, nt, mc, ix.tolist()) for m, ix in zip(ms, idx)]
    assert client.decrypt_cts(got) == ms


def test_lookup_reencrypt(clients):
    """decrypt_then_reencrypt: key A -> key B keeps the plaintext."""
    from cuhepy_amd.paillier._lookup_gpu_ext import PaillierGPULookupClient
    a = clients["paillier_lookup"]
    b = PaillierGPULookupClient(embed_len=_EMBED_LEN, key_len=512, rng=random.Random(77))
    ka = {"pk": json.loads(a.stringify_pk()), "sk": json.loads(a.stringify_sk())}
    kb = {"pk": json.loads(b.stringify_pk()), "sk": json.loads(b.stringify_sk()),
          "message_chunks": json.loads(b.stringify_config())["message_chunks"], **b.dump_tables()}
    nb = int(kb["pk"]["n"])
    rng = random.Random(5)
    ms = [rng.randrange(nb) for _ in range(9)]
    cts = a.encrypt_cts(ms)
    out = a.decrypt_then_reencrypt(cts, ka, kb)
    assert b.decrypt_cts(out) == ms
    assert a._decrypt_key().crt and b._decrypt_key().crt  # phi gives the factors: half-width decryption


def test_readme_pair_distance_256():
    """README.md:59-76 at embed_len 512 / key_len 1024 with the lookup client."""
    from cuhepy_amd.paillier._lookup_gpu_ext import PaillierGPULookupClient
    client = PaillierGPULookupClient(embed_len=512, key_len=1024, rng=random.Random(1))
    a = [i % 2 for i in range(512)]
    b = [(i // 2) % 2 for i in range(512)]
    ca, cb = client.encrypt([a, b])
    enc = PaillierGPULookupClient.encode_hamming_server(ca, cb, json.loads(client.stringify_pk()))
    assert client.decode_hamming_client([enc]) == [256]


def test_invalid_inputs(clients):
    client = clients["paillier"]
    with pytest.raises(RuntimeError):
        client.encrypt([[0, 1, 2, 0, 0, 0, 0, 0]])
    with pytest.raises(RuntimeError):
        client.encrypt([[0, 1]])
    with pytest.raises(RuntimeError):
        client.decode_hamming_client([["1", "2"]])
    assert client.encrypt([]) == []
    assert client.decode_hamming_client([]) == []


def test_lookup_tables_survive_the_keys_file_format(tmp_path):
    """keys.save(include_tables=True) writes {"type", "pk", "sk", "config",
    "tables": dump_tables()} as JSON and keys.load hands the tables back to
    load_config (keys.py:57-120): a JSON round trip turns the g_table keys into
    strings; the reloaded client encrypts exactly as the original with the same
    noise indices and decrypts its ciphertexts."""
    from cuhepy_amd.paillier._lookup_gpu_ext import PaillierGPULookupClient
    client = PaillierGPULookupClient(embed_len=_EMBED_LEN, key_len=512, rng=random.Random(404))
    blob = {"type": "PaillierLookupClient", "pk": client.stringify_pk(), "sk": client.stringify_sk(),
            "config": json.loads(client.stringify_config()), "tables": client.dump_tables()}
    path = tmp_path / "key.json"
    path.write_text(json.dumps(blob))
    back = json.loads(path.read_text())
    assert all(isinstance(k, str) for k in back["tables"]["g_table"])
    cfg = back["config"]
    clone = PaillierGPULookupClient(embed_len=cfg["embed_len"], key_len=cfg["key_len"], alpha_len=cfg["alpha_len"],
                                    skip_key_gen=True)
    clone.load_stringified_keys(back["pk"], back["sk"])
    clone.load_config(cfg, back["tables"])
    n = int(json.loads(client.stringify_pk())["n"])
    rng = random.Random(5)
    ms = [rng.randrange(n) for _ in range(40)] + [0, 1, n - 1]
    idx = np.array([[rng.randrange(256) for _ in range(14)] for _ in ms], dtype=np.uint8)
    a = [int(c) for c in client.encrypt_cts(ms, noise_idx=idx)]
    b = [int(c) for c in clone.encrypt_cts(ms, noise_idx=idx)]
    assert a == b
    assert clone.decrypt_cts(a) == ms
    assert clone.dump_tables()["noise_table"] == client.dump_tables()["noise_table"]
