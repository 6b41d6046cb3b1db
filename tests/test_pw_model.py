"""CPU model of the one-reduction pointwise multiply (k_pw_mulmod in
cuhepy_amd/csrc/lazy.hpp) at every built shape: the product chunks with the
kernel's register ring and row skips (every LDS row it reads checked in
range, every chunk sum checked against the schoolbook product), the chunk
carries, the rows of the folded positions, the fold by the per-modulus table
Z_t = 2^(RB (t + 2)) mod N, the accumulator bound (< 2^64 at every
multiply-add) and the bound V < 2N 2^(2RB) that the finish needs; then the
finish (tests/test_fixed_model.py) gives a b mod N. Planted mistakes (a fold
that starts too high, a dropped chunk carry) must fail.
"""

import random

import pytest

from test_fixed_model import finish_model

# (RB, LP, waves, L32): the shape of each width (pw_shape in chx.hip)
SHAPES = [(29, 144, 8, 128), (29, 72, 4, 64), (28, 38, 2, 32)]


def settled_bound(RB):
    """Positions of T after the chunk carries: a carry (ripple carry plus what
    the folds moved above the chunk, below 3 2^(64-RB)) enters at position 0
    and its top part at position 1."""
    return (1 << RB) + 3 * (1 << (64 - 2 * RB)) + 1


def product_fold_blocks(RB, LP, T):
    """PwSeg<RB, LP, TPI>::PB: product blocks of K rows between folds (a row adds
    a_i b_j < 2^(2RB); a folded position holds less than 2^RB + 2^(64-RB))."""
    K = LP // T
    rows = ((1 << 64) - (1 << RB) - (1 << (64 - RB))) // (1 << (2 * RB)) - 1
    return max(1, rows // K) if rows < T * K else T


def fold_phase_rows(RB):
    """PwSeg<...>::FS: fold-table rows between folds (a row adds a settled T
    position times a canonical Z position; the finish's two REDC products
    follow the last segment), a multiple of 4."""
    row = settled_bound(RB) * (1 << RB)
    base = (1 << RB) + (1 << (64 - RB)) + 2 * (1 << (2 * RB))
    return (((1 << 64) - base) // row - 1) // 4 * 4


def fold(acc, hi, RB):
    M = (1 << RB) - 1
    hi += acc[-1] >> RB
    for p in range(len(acc) - 1, 0, -1):
        acc[p] = (acc[p] & M) + (acc[p - 1] >> RB)
    acc[0] &= M
    return hi


def fold_rows(N, RB, LP, T):
    """Host rule: h0 = floor((bits(N) - 2) / RB) so 2^(RB h0) <= N / 2; nz =
    2 LP - h0 rounded up to 4 rows; h = 2 LP - nz. None: the shape cannot
    hold the rows (nz > 2 LP - 12 T - 2, the first row of the finish's areas)."""
    h0 = (N.bit_length() - 2) // RB
    nz = (2 * LP - h0 + 3) // 4 * 4
    if nz > 2 * LP - 12 * T - 2 or 2 * LP - nz < 0:
        return None
    return 2 * LP - nz, nz


def pw_model(N, a, b, RB, LP, T, L32, h_shift=0, drop_carry=False):
    K = LP // T
    M = (1 << RB) - 1
    LIM = 1 << 64
    pos = lambda v, n: [(v >> (RB * p)) & M for p in range(n)]
    A, B = pos(a, LP), pos(b, LP)
    rows = fold_rows(N, RB, LP, T)
    assert rows is not None
    h, nz = rows
    h, nz = h + h_shift, nz - h_shift  # planted: rows [h, h + h_shift) shifted instead of folded
    assert 2 ** (RB * h) <= N // 2 or h_shift

    def chunk(c, low):
        acc = [0] * K
        ring = [None] * K
        i0 = 0 if low else (c - T) * K
        j0 = c * K if low else LP
        ai, bq = i0, j0 - (K - 1)
        if low:
            for p in range(1, K):
                assert 0 <= j0 + p < LP
                ring[p] = B[j0 + p]

        def block(mode):
            for s in range(K):
                assert 0 <= ai + s < LP
                a_s = A[ai + s]
                slot = (K - s) % K
                if mode == 0 or (mode == 1 and s == 0) or (mode == 2 and s > 0):
                    r = bq + (K - 1 - s)
                    assert 0 <= r < LP, (c, mode, s, r)
                    ring[slot] = B[r]
                for p in range(K):
                    if (mode == 1 and p < s) or (mode == 2 and p >= s):
                        continue
                    v = ring[(p - s) % K]
                    assert v is not None
                    acc[p] += a_s * v
                    assert acc[p] < LIM

        nblocks = c + 1 if low else 2 * T - c
        pb = product_fold_blocks(RB, LP, T)
        hi = 0
        done = 0

        def after_block():
            nonlocal done, hi
            done += 1
            if done % pb == 0 and done < nblocks:
                hi = fold(acc, hi, RB)

        if low:
            for _ in range(c):
                block(0)
                after_block()
                ai += K
                bq -= K
            block(1)
            after_block()
        else:
            block(2)
            after_block()
            for _ in range(1, 2 * T - c):
                ai += K
                bq -= K
                block(0)
                after_block()
        # the chunk holds exactly the schoolbook sums (the folds keep the value)
        want = sum(sum(A[i] * B[t - i] for i in range(max(0, t - LP + 1), min(t, LP - 1) + 1)) << (RB * (t - c * K))
                   for t in range(c * K, c * K + K))
        assert sum(v << (RB * p) for p, v in enumerate(acc)) + (hi << (RB * K)) == want
        return acc, hi

    def ripple(acc_hi):
        acc, hi = acc_hi
        x, c = [], 0
        for v in acc:
            t = v + c
            x.append(t & M)
            c = t >> RB
        return x, c + hi

    X, C = {}, {}
    for w in range(T):
        X[w], C[w] = ripple(chunk(w, True))
        X[T + w], C[T + w] = ripple(chunk(T + w, False))
        assert C[w] < 3 << (64 - RB) and C[T + w] < 3 << (64 - RB)
    assert C[2 * T - 1] == 0  # a b < 2^(RB 2 LP)
    for c in range(1, 2 * T):  # carry-in at the bottom position (settled positions)
        if drop_carry and c == T:
            continue
        t = X[c][0] + C[c - 1]
        X[c][0] = t & M
        X[c][1] += t >> RB
    Tpos = [X[c][p] for c in range(2 * T) for p in range(K)]
    assert sum(v << (RB * t) for t, v in enumerate(Tpos)) == a * b or drop_carry
    assert all(v < settled_bound(RB) for v in Tpos)
    # LDS rows: ts [0, nz), finish areas [2LP - 12T - 2, 2LP - 6T), bd [2LP - 6T, 2LP - 4T), cx [2LP - 4T, 2LP)
    assert nz <= 2 * LP - 12 * T - 2
    Zp = [pos(pow(2, RB * (h + r + 2), N), LP) for r in range(nz)]
    fs = fold_phase_rows(RB)
    acc, his = [], []
    for w in range(T):
        ch = []
        for p in range(K):
            t = w * K + p - 2
            ch.append(Tpos[t] if 0 <= t < h else 0)
        hi = 0
        for r0 in range(0, nz, 4):  # rows in batches of 4, folded every fs rows
            for r in range(r0, r0 + 4):
                for p in range(K):
                    ch[p] += Tpos[h + r] * Zp[r][w * K + p]
                    assert ch[p] < LIM
            if (r0 + 4) % fs == 0 and r0 + 4 < nz:
                hi = fold(ch, hi, RB)
        acc.append(ch)
        his.append(hi)
    V = sum(v << (RB * (w * K + p)) for w in range(T) for p, v in enumerate(acc[w])) + \
        sum(his[w] << (RB * (w * K + K)) for w in range(T))
    assert V % N == a * b * 2 ** (2 * RB) % N or drop_carry or h_shift
    if not h_shift:
        assert V < 2 * N * 2 ** (2 * RB)
    return finish_model(N, acc, RB, LP, T, L32, hi=his)


def _moduli(rng, L32):
    top = 32 * L32
    n = rng.getrandbits(top // 2) | 1 | (3 << (top // 2 - 2))
    return [n * n,                                   # n^2 of a key with the top two bits set
            (1 << top) - 159,                          # just below 2^(32 L32)
            (1 << (top - 1)) + 2 * rng.getrandbits(top - 40) + 1,  # one bit short
            (1 << (top - 20)) + 1,                     # 20 bits short of the width
            (1 << (top * 3 // 4)) + 2 * rng.getrandbits(top * 3 // 4 - 8) + 1]  # 3/4 of the width: more fold rows


@pytest.mark.parametrize("RB,LP,T,L32", SHAPES)
def test_pw_model_exact(RB, LP, T, L32):
    rng = random.Random(LP)
    top = (1 << (32 * L32)) - 1
    for N in _moduli(rng, L32):
        if fold_rows(N, RB, LP, T) is None:
            continue
        for a in (0, 1, N - 1, N, top, rng.randrange(N)):
            for b in (1, N - 1, top, rng.getrandbits(32 * L32)):
                assert pw_model(N, a, b, RB, LP, T, L32) == a * b % N


@pytest.mark.parametrize("RB,LP,T,L32", SHAPES)
def test_pw_fold_rows_cover_the_built_widths(RB, LP, T, L32):
    """Moduli from the full width down to 4/5 of it fit the LDS rows; the
    host falls back to the two-product kernel for shorter ones."""
    for bits in range(32 * L32 * 4 // 5, 32 * L32 + 1):
        assert fold_rows((1 << bits) - 1, RB, LP, T) is not None, bits
    assert fold_rows((1 << (16 * L32)) + 1, RB, LP, T) is None


def test_pw_planted_mistakes_fail():
    rng = random.Random(9)
    RB, LP, T, L32 = SHAPES[1]
    N = _moduli(rng, L32)[0]
    a, b = rng.randrange(N), rng.randrange(N)
    assert pw_model(N, a, b, RB, LP, T, L32) == a * b % N
    # a dropped chunk carry: wrong product
    assert pw_model(N, a, b, RB, LP, T, L32, drop_carry=True) != a * b % N
    # a fold start one row too high leaves T_low above N: the bound fails
    with pytest.raises(AssertionError):
        for _ in range(8):
            a, b = rng.randrange(N), rng.randrange(N)
            assert pw_model(N, a, b, RB, LP, T, L32, h_shift=4) == a * b % N


def test_pw_fold_intervals():
    """The 29-bit shapes fold the product every 3 blocks of 18 rows and the
    fold-table rows every 60 rows; the 28-bit shape never folds."""
    assert product_fold_blocks(29, 144, 8) == 3 and product_fold_blocks(29, 72, 4) == 3
    assert product_fold_blocks(28, 38, 2) == 2
    assert fold_phase_rows(29) == 60 and fold_phase_rows(28) >= 2 * 38


def test_pw_model_needs_the_product_folds():
    """A missing fold in the product phase overflows at 29 bits."""
    rng = random.Random(4)
    RB, LP, T, L32 = SHAPES[0]
    N = _moduli(rng, L32)[0]
    x = (1 << (32 * L32)) - 1
    orig = product_fold_blocks
    globals()["product_fold_blocks"] = lambda RB_, LP_, T_: T_
    try:
        with pytest.raises(AssertionError):
            pw_model(N, x, x, RB, LP, T, L32)
    finally:
        globals()["product_fold_blocks"] = orig
