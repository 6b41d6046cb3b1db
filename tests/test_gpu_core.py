"""GPU parity of the core kernels against the exact-integer oracle.

Every built (width, lanes-per-bignum) variant is forced in turn through the
context tuning knob, so each kernel instantiation is checked bit-exactly.
"""

import random

import numpy as np
import pytest

from conftest import H
from oracle import paillier_oracle as O

pytestmark = pytest.mark.gpu

OP_MODEXP, OP_MULMOD, OP_LOOKUP = 0, 1, 2
# shape (RB, LP, waves) of k_pw_mulmod per width, and the host's fold-row rule
PW_SHAPE = {32: (28, 38, 2), 64: (29, 72, 4), 128: (29, 144, 8)}
from test_pw_model import fold_rows  # noqa: E402


@pytest.fixture(autouse=True)
def _reset_tuning(ctx):
    yield
    for limbs in (32, 64, 128):
        for op in (0, 1, 2, 3):
            ctx.set_tuning(op, limbs, 0, 0)


def _variants(limbs):
    from cuhepy_amd.engine import Context
    return Context.supported_tpi(limbs)


def _mod(rng, limbs, top=True):
    n = rng.getrandbits(32 * limbs) | 1
    if top:
        n |= 1 << (32 * limbs - 1)
    return n


@pytest.mark.parametrize("limbs", [32, 64, 128])
def test_mulmod_every_variant(ctx, limbs):
    from cuhepy_amd.engine import Modulus
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    rng = random.Random(100 + limbs)
    R = 1 << (32 * limbs)
    for tpi in _variants(limbs):
        ctx.set_tuning(OP_MULMOD, limbs, tpi, 0)
        for top in (True, False):
            n = _mod(rng, limbs, top)
            mod = Modulus(ctx, n, limbs)
            count = 131  # not a multiple of any group size: exercises the tail
            a = [0, 1, n - 1, R - 1, n] + [rng.randrange(R) for _ in range(count - 5)]
            b = [n - 1, R - 1, n - 1, R - 1, 2] + [rng.randrange(R) for _ in range(count - 5)]
            got = limbs_to_ints(mod.mulmod(ints_to_limbs(a, limbs), ints_to_limbs(b, limbs)))
            want = [x * y % n for x, y in zip(a, b)]
            bad = [i for i, (g, w) in enumerate(zip(got, want)) if g != w]
            assert not bad, f"tpi={tpi} top={top} first bad index {bad[:5]}"


@pytest.mark.parametrize("limbs", [32, 64, 128])
def test_mulmod_lazy_large_batches(ctx, limbs):
    """Batches of 8,192 or more items: the one-reduction multiply (k_pw_mulmod:
    product, fold by the modulus's table, two REDC rows) and the two-product
    form (window 2; both lazy configurations of the width). Full-width
    moduli, moduli a few and 20 bits short of the width (more fold rows),
    64 bits short, 3/4 of the width; half the width and 9 L bits (the fold
    rows do not fit: the two-product form); near 2^(32 L); operands up to
    2^(32 L) - 1."""
    from cuhepy_amd.engine import Context, Modulus
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    rng = random.Random(300 + limbs)
    R = 1 << (32 * limbs)
    count = 9000
    moduli = [_mod(rng, limbs, True), _mod(rng, limbs, False), R - 2 * rng.randrange(1, 1 << 20) - 1,
              (1 << (32 * limbs - 20)) + 2 * rng.getrandbits(32 * limbs - 30) + 1,
              (1 << (32 * limbs - 64)) + 2 * rng.getrandbits(32 * limbs - 70) + 1,
              (1 << (24 * limbs)) + 2 * rng.getrandbits(24 * limbs - 8) + 1,  # 3/4 of the width
              (1 << (16 * limbs)) + 2 * rng.getrandbits(16 * limbs - 8) + 1,  # half: two-product form
              (1 << (9 * limbs)) + 2 * rng.getrandbits(9 * limbs - 8) + 1]  # too short for the fold rows
    assert fold_rows(moduli[0], *PW_SHAPE[limbs]) and not fold_rows(moduli[-1], *PW_SHAPE[limbs])
    forms = [(0, 0)] + [(2, t) for t in Context.supported_tpi(limbs, op=OP_MODEXP)]
    for window, tpi in forms:
        ctx.set_tuning(OP_MODEXP, limbs, tpi, 0)
        ctx.set_tuning(OP_MULMOD, limbs, 0, window)
        for k, n in enumerate(moduli):
            mod = Modulus(ctx, n, limbs)
            a = [0, 1, n - 1, R - 1, n] + [rng.randrange(R) for _ in range(count - 5)]
            b = [n - 1, R - 1, n - 1, R - 1, 2] + [rng.randrange(R) for _ in range(count - 5)]
            ctx.reset_timing()
            ctx.set_timing(True)
            got = limbs_to_ints(mod.mulmod(ints_to_limbs(a, limbs), ints_to_limbs(b, limbs)))
            ctx.set_timing(False)
            one_fold = window == 0 and fold_rows(n, *PW_SHAPE[limbs]) is not None
            assert ctx.kernel_time("pw_mulmod")[1] == int(one_fold), (window, k)
            bad = [i for i, (g, x, y) in enumerate(zip(got, a, b)) if g != x * y % n]
            assert not bad, f"window={window} tpi={tpi} n bits {n.bit_length()}: first bad index {bad[:5]}"


@pytest.mark.parametrize("limbs", [32, 64, 128])
def test_modexp_every_variant(ctx, limbs):
    from cuhepy_amd.engine import Modulus
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    rng = random.Random(200 + limbs)
    from cuhepy_amd.engine import Context
    for tpi in Context.supported_tpi(limbs, op=OP_MODEXP):
        for window in (0, 2):
            ctx.set_tuning(OP_MODEXP, limbs, tpi, window)
            n = _mod(rng, limbs)
            mod = Modulus(ctx, n, limbs)
            count = 70
            bases = [0, 1, n - 1] + [rng.randrange(n) for _ in range(count - 3)]
            e = rng.getrandbits(min(1024, 32 * limbs)) | 1
            got = limbs_to_ints(mod.modexp(ints_to_limbs(bases, limbs), e))
            want = [pow(b, e, n) for b in bases]
            bad = [i for i, (g, w) in enumerate(zip(got, want)) if g != w]
            assert not bad, f"tpi={tpi} window={window} first bad {bad[:5]}"


def test_modexp_small_and_zero_exponents(ctx):
    from cuhepy_amd.engine import Modulus
    from cuhepy_amd.limbs import ints_to_limbs, limbs_to_ints
    rng = random.Random(7)
    n = _mod(rng, 64)
    mod = Modulus(ctx, n, 64)
    bases = [rng.randrange(n) for _ in range(40)]
    for e in (0, 1, 2, 3, 65537, (1 << 100)):
        got = limbs_to_ints(mod.modexp(ints_to_limbs(bases, 64), e))
        assert got == [pow(b, e, n) for b in bases], e


@pytest.mark.parametrize("case", ["paillier_w64", "paillier_w128"])
def test_modexp_golden(ctx, golden, case):
    from cuhepy_amd.engine import Modulus
    from cuhepy_amd.limbs import limbs_to_ints, width_for_bits
    g = golden[case]
    n2 = H(g["n_squared"])
    mod = Modulus(ctx, n2, width_for_bits(n2.bit_length()))
    bases = [H(b) for b in g["modexp_bases"]]
    for e_hex, row in zip(g["modexp_exps"], g["modexp"]):
        got = limbs_to_ints(mod.modexp(bases, H(e_hex)))
        assert got == [H(v) for v in row]
    a = [H(v) for v in g["mul_a"]]
    b = [H(v) for v in g["mul_b"]]
    assert limbs_to_ints(mod.mulmod(a, b)) == [H(v) for v in g["mul"]]


def test_empty_batches(ctx):
    from cuhepy_amd.engine import Modulus
    mod = Modulus(ctx, (1 << 2047) + 1, 64)
    assert mod.mulmod([], []).shape == (0, 64)
    assert mod.modexp([], 5).shape == (0, 64)


def test_invalid_inputs_raise(ctx):
    from cuhepy_amd.engine import Modulus
    n = (1 << 1023) + 12345677
    mod = Modulus(ctx, n, 32)
    with pytest.raises(ValueError):
        mod.modexp([n], 3)  # base must be below the modulus
    with pytest.raises(ValueError):
        Modulus(ctx, 1 << 100, 32)  # even modulus
    with pytest.raises(ValueError):
        mod.mulmod([1 << 1024], [1])  # does not fit the width


@pytest.mark.parametrize("limbs", [32, 64, 128])
def test_modexp_each_every_variant(ctx, limbs):
    """Per-item exponents (fixed window): zero, one, small and full-width."""
    from cuhepy_amd.engine import Context, Modulus
    from cuhepy_amd.limbs import limbs_to_ints
    rng = random.Random(300 + limbs)
    OP_RAGGED = 3
    for tpi in Context.supported_tpi(limbs, op=OP_RAGGED):
        for window in (0, 3):
            ctx.set_tuning(OP_RAGGED, limbs, tpi, window)
            n = _mod(rng, limbs)
            mod = Modulus(ctx, n, limbs)
            count = 67
            bases = [0, 1, n - 1, rng.randrange(n)] + [rng.randrange(n) for _ in range(count - 4)]
            exps = [5, 0, 1, 2] + [rng.getrandbits(rng.choice([1, 17, 300, 32 * limbs])) for _ in range(count - 4)]
            got = limbs_to_ints(mod.modexp_each(bases, exps))
            want = [pow(b, e, n) for b, e in zip(bases, exps)]
            bad = [i for i, (g, w) in enumerate(zip(got, want)) if g != w]
            assert not bad, f"tpi={tpi} window={window} first bad {bad[:5]}"


def test_oversize_batches_rejected(ctx):
    """Counts or strides beyond the 2^28 bound are refused before any launch
    (kernels index items with 32-bit integers)."""
    import torch
    from cuhepy_amd.engine import Modulus
    mod = Modulus(ctx, (1 << 2047) + 12345, 64)
    t = torch.zeros((64, 64), dtype=torch.int32, device="cuda:0")
    big = 1 << 29
    with pytest.raises(ValueError):
        mod.mulmod_dev(t, t, t, big, big)
    with pytest.raises(ValueError):
        mod.modexp_dev(t, t, big, big, 65537)
    with pytest.raises(ValueError):
        ctx.aos_to_il(t, t, 64, 64, big)
