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

# (RB, LP, waves, L32): the configuration of each width with 19 positions per chunk
SHAPES = [(27, 152, 8, 128), (28, 76, 4, 64), (28, 38, 2, 32)]


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

        if low:
            for _ in range(c):
                block(0)
                ai += K
                bq -= K
            block(1)
        else:
            block(2)
            for _ in range(1, 2 * T - c):
                ai += K
                bq -= K
                block(0)
        for p in range(K):  # the chunk holds exactly the schoolbook sums
            t = c * K + p
            assert acc[p] == sum(A[i] * B[t - i] for i in range(max(0, t - LP + 1), min(t, LP - 1) + 1))
        return acc

    def ripple(acc):
        x, c = [], 0
        for v in acc:
            t = v + c
            x.append(t & M)
            c = t >> RB
        return x, c

    X, C = {}, {}
    for w in range(T):
        X[w], C[w] = ripple(chunk(w, True))
        X[T + w], C[T + w] = ripple(chunk(T + w, False))
    assert C[2 * T - 1] == 0  # a b < 2^(RB 2 LP)
    for c in range(1, 2 * T):  # carry-in at the bottom position (settled positions)
        if drop_carry and c == T:
            continue
        t = X[c][0] + C[c - 1]
        X[c][0] = t & M
        X[c][1] += t >> RB
    Tpos = [X[c][p] for c in range(2 * T) for p in range(K)]
    assert sum(v << (RB * t) for t, v in enumerate(Tpos)) == a * b or drop_carry
    assert all(v < (1 << RB) + (1 << (64 - 2 * RB)) + 1 for v in Tpos)
    # LDS rows: ts [0, nz), finish areas [2LP - 12T - 2, 2LP - 6T), bd [2LP - 6T, 2LP - 4T), cx [2LP - 4T, 2LP)
    assert nz <= 2 * LP - 12 * T - 2
    Zp = [pos(pow(2, RB * (h + r + 2), N), LP) for r in range(nz)]
    acc = []
    for w in range(T):
        ch = []
        for p in range(K):
            t = w * K + p - 2
            v = Tpos[t] if 0 <= t < h else 0
            for r in range(nz):
                v += Tpos[h + r] * Zp[r][w * K + p]
            assert v < LIM
            ch.append(v)
        acc.append(ch)
    V = sum(v << (RB * (w * K + p)) for w in range(T) for p, v in enumerate(acc[w]))
    assert V % N == a * b * 2 ** (2 * RB) % N or drop_carry or h_shift
    if not h_shift:
        assert V < 2 * N * 2 ** (2 * RB)
    return finish_model(N, acc, RB, LP, T, L32)


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
