"""CPU model of the multiply-free Montgomery digits of the lazy engine
(Lazy::UNIT in cuhepy_amd/csrc/lazy.hpp, used by k_lazy_modexp, the per-item
and ragged exponentiations and the lookup encryption).

With k = -N^-1 mod 2^RB the modulus N' = kN is -1 mod 2^RB, so every REDC
digit is the low RB bits of the running value (no multiply). The chain of
products runs modulo N' with R' = 2^(RB LP); operands only need to be
congruent mod N to their Montgomery forms and below 2N' (below 4N' for the
CRT input sum), and the last product runs modulo N itself. This model checks,
at every built shape where the engine uses it, that the digits clear the low
positions, that every product stays below 2N' (the bound R' > 16 N' gives),
and that the final product modulo N gives the exact power; a planted mistake
(digits computed as if N' were N) must fail.
"""

import random

import pytest

# (RB, LP, L32) of the lazy configurations whose UNIT flag is set (every
# built one: lazy_cfgs in chx.hip), and two earlier 28- and 27-bit shapes
UNIT_SHAPES = [(28, 38, 32), (29, 72, 64), (29, 144, 128), (28, 76, 64), (27, 160, 128)]
# an earlier shape where it is not (R' has no room for N')
NON_UNIT = (27, 152, 128)


def unit(rb, lp, l32):
    return rb * lp >= 32 * l32 + rb + 4


def redc(t, N, rb, lp, digit):
    """t R'^-1 mod N for t < R' N: lp digit rows of rb bits (right to left)."""
    M = (1 << rb) - 1
    for _ in range(lp):
        m = digit(t & M)
        t += m * N
        assert t & M == 0, "the digit must clear the low position"
        t >>= rb
    return t


class Engine:
    def __init__(self, N, rb, lp, l32, use_kn=True, planted=False):
        self.N, self.rb, self.lp = N, rb, lp
        self.R = 1 << (rb * lp)
        M = (1 << rb) - 1
        n0 = (-pow(N, -1, 1 << rb)) % (1 << rb)
        self.digit_n = lambda v: (v * n0) & M  # the digit of N itself
        if use_kn:
            self.Np = n0 * N  # N' = kN = -1 mod 2^rb
            assert (self.Np + 1) & M == 0
            # planted: the digit as if the modulus were N (n0 multiply dropped)
            self.digit = (lambda v: (v * n0) & M) if planted else (lambda v: v)
        else:
            self.Np = N
            self.digit = lambda v: (v * n0) & M
        self.R2 = self.R * self.R % N  # constants built mod N serve unchanged
        self.RW = (1 << (32 * l32)) * self.R2 % N

    def mul(self, a, b, last=False):
        # the last product runs modulo N with its own digits
        mod, digit = (self.N, self.digit_n) if last else (self.Np, self.digit)
        r = redc(a * b, mod, self.rb, self.lp, digit)
        assert r < 2 * mod, "product above 2N'"
        return r


def model_modexp(x, e, N, rb, lp, l32, wide_hi=None, **kw):
    E = Engine(N, rb, lp, l32, **kw)
    xm = E.mul(x, E.R2)
    if wide_hi is not None:  # CRT input: c_lo R'^2 + c_hi 2^(32 L32) R'^2, below 4N'
        xm += E.mul(wide_hi, E.RW)
        assert xm < 4 * E.Np
    acc = E.R % E.N  # Montgomery one, built mod N
    for bit in bin(e)[2:]:
        acc = E.mul(acc, acc)
        if bit == "1":
            acc = E.mul(acc, xm)
    out = E.mul(acc, 1, last=True)
    assert out <= E.N
    return out % E.N


@pytest.mark.parametrize("rb,lp,l32", UNIT_SHAPES)
def test_unit_digit_modexp_exact(rb, lp, l32):
    assert unit(rb, lp, l32)
    rng = random.Random(rb * lp + l32)
    for _ in range(3):
        p = rng.getrandbits(16 * l32) | 1 | (1 << (16 * l32 - 1))
        N = p * p  # a CRT half modulus of 32 l32 bits
        e = rng.getrandbits(64) | 1
        for x in (0, 1, N - 1, (1 << (32 * l32)) - 1, rng.randrange(N)):
            assert model_modexp(x, e, N, rb, lp, l32) == pow(x, e, N)
        # double-width input of the CRT halves: c = c_lo + 2^(32 l32) c_hi
        c_lo, c_hi = rng.getrandbits(32 * l32), rng.getrandbits(32 * l32)
        c = c_lo + (c_hi << (32 * l32))
        assert model_modexp(c_lo, e, N, rb, lp, l32, wide_hi=c_hi) == pow(c, e, N)


def test_unit_flag_matches_the_built_shapes():
    assert all(unit(*s) for s in UNIT_SHAPES)
    assert not unit(*NON_UNIT)


def test_unit_digit_needs_kn():
    """Digits computed without the multiply but reduced against N itself do
    not clear the low position; digits of N applied to N' do not either."""
    rb, lp, l32 = UNIT_SHAPES[1]
    rng = random.Random(5)
    p = rng.getrandbits(16 * l32) | 1 | (1 << (16 * l32 - 1))
    N = p * p
    with pytest.raises(AssertionError):
        E = Engine(N, rb, lp, l32, use_kn=False)
        E.digit = lambda v: v  # no n0 multiply, modulus N
        E.mul(rng.randrange(N), rng.randrange(N))
    with pytest.raises(AssertionError):
        Engine(N, rb, lp, l32, planted=True).mul(rng.randrange(N), rng.randrange(N))
