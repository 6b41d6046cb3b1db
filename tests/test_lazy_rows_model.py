"""Lane-level CPU model of the lazy engine's Montgomery rows (Lazy::mul_staged,
Lazy::sqr_staged, Lazy::row, Lazy::row_sq, Lazy::reduce and Lazy::settle in
cuhepy_amd/csrc/lazy.hpp), used by k_lazy_modexp, the per-item and ragged
exponentiations, the lookup encryption and the pointwise fallback.

A group of TPI lanes holds K = LP / TPI positions per lane in 64-bit
registers; position p of row s of a block lives in register (p + s) mod K.
Each row adds a_i * b (or the symmetric squaring terms) and m_i * N, then the
lane's bottom register carries its bits above RB into the lane's next
register and passes only its low RB bits one lane down, where they become
that lane's top position. So a register accumulates for at most K rows before
it is split again; this bound (not LP rows) decides whether a shape needs
carry folds (Lazy::FOLD) and whether it may use the symmetric squaring rows
(Lazy::SYM).

The model follows the kernel register by register and checks:
- every product against x y R'^-1 mod N (N' = kN on UNIT shapes), with the
  result below 2N', over chains of products fed with settled positions;
- the largest register value of a worst-case run (settled operand positions
  at their bound, digits at their maximum) against 2^64, at every built
  shape and at the bound the kernel's constants are derived from;
- planted mistakes that must fail: a bottom register passed down whole (the
  accumulation then spans LP rows and overflows at 29-bit positions), and a
  dropped carry into the next register.
"""

import random

import pytest

# (RB, LP, TPI, L32) of the lazy configurations (lazy_cfgs in chx.hip)
BUILT = [(28, 38, 1, 32), (28, 38, 2, 32), (28, 40, 8, 32), (29, 72, 4, 64), (29, 72, 8, 64),
         (29, 80, 16, 64), (29, 144, 8, 128), (29, 144, 16, 128)]
# the shapes these replaced (27- and 28-bit positions at 2048- and 4096-bit
# moduli), still valid configurations of the engine
EARLIER = [(28, 76, 4, 64), (28, 80, 8, 64), (27, 152, 8, 128), (27, 160, 16, 128)]

TWO64 = 1 << 64


def settled_bound(rb):
    """Positions after settle are below D = 2^RB + 2^(64-2RB) + 1."""
    return (1 << rb) + (1 << (64 - 2 * rb)) + 1


def row_limits(rb):
    """Lazy::FMAX and Lazy::FMAX_SQ, restated in exact integers: the most rows
    a register may accumulate. It enters below 2^RB (low bits passed down) and
    takes one carry below 2^(64-RB); a multiply row adds two products, a
    squaring row at most three products' worth, each below
    P = 2^(2RB) + 2^(66-RB) (settled operand positions)."""
    p = (1 << (2 * rb)) + (1 << (66 - rb))
    room = TWO64 - 2 * (1 << (64 - rb))
    return room // (2 * p), room // (3 * p)


def unit(rb, lp, l32):
    return rb * lp >= 32 * l32 + rb + 4


def to_positions(v, rb, lp):
    m = (1 << rb) - 1
    return [(v >> (rb * p)) & m for p in range(lp)]


def value(pos, rb):
    return sum(x << (rb * p) for p, x in enumerate(pos))


class Rows:
    """One group of TPI lanes running the engine's rows."""

    def __init__(self, rb, lp, tpi, l32, pass_whole=False, drop_carry=False):
        self.rb, self.lp, self.tpi, self.l32 = rb, lp, tpi, l32
        self.K = lp // tpi
        self.mask = (1 << rb) - 1
        self.pass_whole = pass_whole
        self.drop_carry = drop_carry
        self.vmax = 0

    def _track(self, v):
        if v > self.vmax:
            self.vmax = v

    def product(self, a, b, N, n0, U, sq=False, digit=None):
        """settle(a * b R'^-1) as positions; a, b, N positions (LP each).
        With sq, a == b and the symmetric squaring rows run. digit forces the
        row digit (worst-case runs), else it is computed as the kernel does."""
        rb, K, T, mask = self.rb, self.K, self.tpi, self.mask
        V = [[0] * K for _ in range(T)]
        bl = [[b[r * K + p] for p in range(K)] for r in range(T)]
        Nl = [[N[r * K + p] for p in range(K)] for r in range(T)]
        for blk in range(T):
            for S in range(K):
                i = blk * K + S
                if sq:
                    x, x2 = a[i], (a[i] << 1) & 0xFFFFFFFF
                    assert a[i] << 1 < 1 << 32
                    for r in range(T):
                        V[r][(2 * S) % K] += x * bl[r][S]
                        for p in range(S + 1, K):
                            V[r][(p + S) % K] += x2 * bl[r][p]
                else:
                    for r in range(T):
                        for p in range(K):
                            V[r][(p + S) % K] += a[i] * bl[r][p]
                low = V[0][S] & 0xFFFFFFFF
                if digit is not None:
                    m = digit
                elif U:
                    m = low & mask
                else:
                    m = ((low * n0) & 0xFFFFFFFF) & mask
                for r in range(T):
                    for p in range(K):
                        q = (p + S) % K
                        V[r][q] += m * Nl[r][p]
                        self._track(V[r][q])
                v0 = [V[r][S] for r in range(T)]
                if digit is None:
                    assert v0[0] & mask == 0, "the digit clears rank 0's bottom position"
                if self.pass_whole:
                    # planted: every bottom register goes one lane down whole;
                    # rank 0's (low bits cleared) shifts into its next register
                    V[0][(S + 1) % K] += v0[0] >> rb
                    for r in range(T):
                        V[r][S] = v0[r + 1] if r + 1 < T else 0
                else:
                    for r in range(T):
                        if not self.drop_carry:
                            V[r][(S + 1) % K] += v0[r] >> rb
                        # low 32 bits through the DPP move, masked; the top
                        # rank gets the next group's rank 0 bits, which the
                        # digit cleared
                        V[r][S] = ((v0[r + 1] & 0xFFFFFFFF) & mask) if r + 1 < T else 0
                for r in range(T):
                    for q in range(K):
                        self._track(V[r][q])
                        assert V[r][q] < TWO64, "accumulator overflow"
        return self._settle(V, digit is None)

    def _settle(self, V, exact):
        rb, K, T, mask = self.rb, self.K, self.tpi, self.mask
        X = [[0] * K for _ in range(T)]
        carry = [0] * T
        for r in range(T):
            c = 0
            for p in range(K):
                t = V[r][p] + c
                self._track(t)
                assert t < TWO64, "settle overflow"
                X[r][p] = t & mask
                c = t >> rb
            carry[r] = c
        if exact:
            assert carry[T - 1] == 0, "the value must stay below R'"
        for r in range(1, T):
            t = X[r][0] + carry[r - 1]
            X[r][0] = t & mask
            X[r][1] += t >> rb
            assert X[r][1] < 1 << 32
        return [X[r][p] for r in range(T) for p in range(self.K)]


def setup(rb, lp, l32, rng):
    N = rng.getrandbits(32 * l32) | (1 << (32 * l32 - 1)) | 1
    k = (-pow(N, -1, 1 << rb)) % (1 << rb)
    U = unit(rb, lp, l32)
    Nm = k * N if U else N  # the modulus of the chain (N' = kN on UNIT shapes)
    n0 = (-pow(Nm, -1, 1 << rb)) % (1 << rb)
    if U:
        assert n0 == 1
    return Nm, n0, U


def chain(shape, seed, steps=5, **kw):
    rb, lp, tpi, l32 = shape
    rng = random.Random(seed)
    Nm, n0, U = setup(rb, lp, l32, rng)
    R = 1 << (rb * lp)
    Rinv = pow(R, -1, Nm)
    eng = Rows(rb, lp, tpi, l32, **kw)
    Np = to_positions(Nm, rb, lp)
    x = to_positions(rng.randrange(2 * Nm), rb, lp)
    y = to_positions(rng.randrange(2 * Nm), rb, lp)
    for s in range(steps):
        sq = s % 2 == 0
        vx, vy = value(x, rb), value(y, rb)
        out = eng.product(x, x if sq else y, Np, n0, U, sq=sq)
        v = value(out, rb)
        want = (vx * (vx if sq else vy) * Rinv) % Nm
        assert v % Nm == want, "product is not x y R'^-1 mod N"
        assert v < 2 * Nm, "product above 2N'"
        if sq:
            x = out
        else:
            y = out
    return eng


@pytest.mark.parametrize("shape", BUILT + EARLIER)
def test_rows_give_the_montgomery_product(shape):
    chain(shape, seed=sum(shape))


@pytest.mark.parametrize("shape", BUILT + EARLIER)
def test_worst_case_registers_stay_below_2_64(shape):
    """Settled operand positions at their bound and the largest digit in
    every row: the largest register value, against the row limit the kernel's
    constants use."""
    rb, lp, tpi, l32 = shape
    K = lp // tpi
    fmax, fmax_sq = row_limits(rb)
    assert K <= fmax, "the shape would need carry folds"
    d = settled_bound(rb) - 1
    mask = (1 << rb) - 1
    top = [d] * lp
    N = [mask] * lp
    eng = Rows(rb, lp, tpi, l32)
    eng.product(top, top, N, 1, True, sq=False, digit=mask)
    p = (1 << (2 * rb)) + (1 << (66 - rb))
    assert eng.vmax < (1 << rb) + (1 << (64 - rb)) + 2 * K * p <= TWO64
    if K <= fmax_sq:
        eng = Rows(rb, lp, tpi, l32)
        eng.product(top, top, N, 1, True, sq=True, digit=mask)
        assert eng.vmax < (1 << rb) + (1 << (64 - rb)) + 3 * K * p <= TWO64


def test_built_shapes_need_no_folds_and_square_symmetrically():
    """Lazy::FOLD is false and Lazy::SYM true at every built shape; the 29-bit
    shapes have K = 18 or 9 rows per register (limits 31 and 21), the 1024-bit
    shapes keep 28 bits (one lane per modulus holds all 38 rows, and 29 bits
    would allow 31)."""
    for rb, lp, tpi, l32 in BUILT:
        K = lp // tpi
        fmax, fmax_sq = row_limits(rb)
        assert K <= fmax and K <= fmax_sq
        assert rb * lp >= 32 * l32 + 2, "R' > 4N"
    assert row_limits(29) == (31, 21)
    assert row_limits(28)[0] >= 38


def test_29_bit_shapes_use_the_unit_digit():
    """R' > 16 N' at (29, 72) and (29, 144): the digit multiply goes at the
    2048- and 4096-bit moduli."""
    assert unit(29, 72, 64) and unit(29, 144, 128)
    assert not unit(27, 152, 128)


def test_planted_whole_register_pass_overflows_at_29_bits():
    """A bottom register passed down whole accumulates for LP rows: at
    (29, 144, 8) the worst case passes 2^64, which the model must catch."""
    rb, lp, tpi, l32 = 29, 144, 8, 128
    d = settled_bound(rb) - 1
    mask = (1 << rb) - 1
    eng = Rows(rb, lp, tpi, l32, pass_whole=True)
    with pytest.raises(AssertionError):
        eng.product([d] * lp, [d] * lp, [mask] * lp, 1, True, digit=mask)


def test_planted_dropped_carry_fails():
    with pytest.raises(AssertionError):
        chain((29, 72, 4, 64), seed=7, steps=2, drop_carry=True)
