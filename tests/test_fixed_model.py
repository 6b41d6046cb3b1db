"""CPU model of the served fixed-operand multiply (k_fixed_mulmod in
cuhepy_amd/csrc/lazy.hpp): the chunked reduction, the carry lookahead over
chunks, the limb assignment and the borrow lookahead of the conditional
subtraction, at every built (RB, LP, chunks, limbs) shape, against x*y mod N.

The kernel's GPU parity is checked by tests/test_gpu_serve.py; this model pins
the arithmetic of its layout (which positions and limbs each chunk owns, the
bounds that keep them exact) on the CPU, with a planted mistake that must fail.
"""

import random

import pytest

# the served shapes (serve_cfg in chx.hip) and the pointwise multiply's shapes,
# which share the finish: (RB, LP, chunks, limbs, bits of n)
SHAPES = [(29, 144, 8, 128, 2048), (29, 72, 4, 64, 1024), (28, 38, 1, 32, 512),
          (27, 152, 8, 128, 2048), (28, 76, 4, 64, 1024)]


def fixed_rows(LP):
    return 4 if LP % 4 == 0 else 2 if LP % 2 == 0 else 1


def fold_interval(RB, LP):
    """FixedSeg<RB, LP>::SEG, restated: the most rows between folds such that
    SEG rows of x_i Y_p (x_i < 2^RB, settled Y_p < 2^RB + 2^(64-2RB) + 1) on top
    of a folded position (< 2^RB + 2^(64-RB)) and the finish's two REDC
    products (< 2^(2RB) each) stay below 2^64, as a whole number of row
    batches dividing LP (LP itself when every row fits)."""
    row = (1 << RB) * ((1 << RB) + (1 << (64 - 2 * RB)) + 1)
    base = (1 << RB) + (1 << (64 - RB)) + 2 * (1 << (2 * RB))
    maxr = ((1 << 64) - base) // row - 1
    if maxr >= LP:
        return LP
    s = maxr
    while not (LP % s == 0 and s % fixed_rows(LP) == 0):
        s -= 1
    return s


def row_loop_model(xs, Yp, RB, LP, T, seg):
    """k_fixed_mulmod's row loop for every chunk: V's positions accumulated row
    by row, folded every `seg` rows (each keeps its low RB bits plus the bits
    of the one below above RB; the top one's high bits go to hi). Every
    accumulator is checked against 2^64 after every multiply-add."""
    K = LP // T
    M = (1 << RB) - 1
    acc = [[0] * K for _ in range(T)]
    hi = [0] * T
    for s0 in range(0, LP, seg):
        for i in range(s0, s0 + seg):
            for W in range(T):
                a = acc[W]
                for p in range(K):
                    a[p] += xs[i] * Yp[i][W * K + p]
                    assert a[p] < 1 << 64, "accumulator overflow"
        if s0 + seg < LP:
            for W in range(T):
                a = acc[W]
                hi[W] += a[K - 1] >> RB
                for p in range(K - 1, 0, -1):
                    a[p] = (a[p] & M) + (a[p - 1] >> RB)
                a[0] &= M
    return acc, hi


def fixed_mulmod_model(N, x, y, RB, LP, T, L32, drop_top_term=False, seg=None):
    M = (1 << RB) - 1

    def positions(v, n):
        return [(v >> (RB * p)) & M for p in range(n)]

    xs = positions(x, LP)
    Yp = [positions((y << (RB * (i + 2))) % N, LP) for i in range(LP)]
    # B: chunk W sums positions [W K, W K + K) of V = sum_i x_i Y_i
    acc, hi = row_loop_model(xs, Yp, RB, LP, T, fold_interval(RB, LP) if seg is None else seg)
    return finish_model(N, acc, RB, LP, T, L32, drop_top_term, hi)


def finish_model(N, acc, RB, LP, T, L32, drop_top_term=False, hi=None):
    """fixed_finish_w: chunk W's K accumulators of V (acc[W][p], V < 2N 2^(2RB))
    and what its folds moved above its top position (hi[W])
    -> canonical V / 2^(2RB) mod N."""
    K = LP // T
    M = (1 << RB) - 1
    n0 = (-pow(N, -1, 1 << RB)) % (1 << RB)
    Np = [(N >> (RB * p)) & M for p in range(LP)]
    N32 = [(N >> (32 * j)) & 0xFFFFFFFF for j in range(L32)]
    acc = [list(a) for a in acc]
    hi = [0] * T if hi is None else hi
    # 1: REDC digits from positions 0 and 1
    m1 = (acc[0][0] * n0) & M
    t0 = acc[0][0] + m1 * Np[0]
    m2 = ((acc[0][1] + m1 * Np[1] + (t0 >> RB)) * n0) & M
    # 2: (m1 + m2 2^RB) N
    for W in range(T):
        for p in range(K):
            q = W * K + p
            acc[W][p] += m1 * Np[q] + (m2 * Np[q - 1] if q >= 1 else 0)
            assert acc[W][p] < 1 << 64
    # 3: exact positions; the top chunk keeps everything at positions >= LP
    X = [[0] * K for _ in range(T)]
    cw = [0] * T
    ctop = 0
    for W in range(T):
        c = 0
        for p in range(K):
            t = acc[W][p] + c
            X[W][p], c = t & M, t >> RB
        if W < T - 1:
            cw[W] = c + hi[W]
            assert cw[W] < 1 << 37
        else:
            ctop = c + hi[W] + (0 if drop_top_term else m2 * Np[LP - 1])
    gp = [0] * T
    for W in range(T):
        cin = cw[W - 1] if W > 0 else 0
        allm = M
        for p in range(K):
            t = X[W][p] + cin
            X[W][p], cin = t & M, t >> RB
            allm &= X[W][p]
        if W < T - 1:
            assert cin <= 1  # one bit leaves a chunk after its second ripple
            gp[W] = cin | (2 if allm == M else 0)
        else:
            ctop += cin
    for W in range(T):
        cin = 0
        for w in range(W):
            cin = (gp[w] & 1) | ((gp[w] >> 1) & cin)
        for p in range(K):
            t = X[W][p] + cin
            X[W][p], cin = t & M, t >> RB
        if W == T - 1:
            ctop += cin
    assert ctop < 1 << 47
    # 4 and 5: limbs of R = V / 2^(2 RB) per chunk, R - N with a borrow lookahead
    out = [0] * L32
    flags, limbs = [0] * T, {}
    for W in range(T):
        up = (X[W + 1][0], X[W + 1][1]) if W < T - 1 else (ctop & M, ctop >> RB)

        def pos(q):
            o = q - W * K
            return X[W][o] if o < K else up[o - K] if o < K + 2 else 0

        B0, B1 = RB * W * K - 2 * RB, RB * (W * K + K) - 2 * RB
        J0 = 0 if B0 <= 0 else (B0 + 31) // 32
        J1 = L32 if W == T - 1 else min((B1 + 31) // 32, L32)
        rl = []
        for j in range(J0, J1):
            b = 32 * j + 2 * RB
            q, o = b // RB, b % RB
            v = (pos(q) >> o) | (pos(q + 1) << (RB - o))
            if 2 * RB - o < 64:
                v |= pos(q + 2) << (2 * RB - o)
            rl.append(v & 0xFFFFFFFF)
        bo, eq = 0, 1
        for k, j in enumerate(range(J0, J1)):
            bo = 1 if rl[k] - N32[j] - bo < 0 else 0
            eq &= rl[k] == N32[j]
        tb = 0
        if W == T - 1:
            b = 32 * L32 + 2 * RB
            tb = (pos(b // RB) >> (b % RB)) & 1
        flags[W] = bo | (eq << 1) | (tb << 2)
        limbs[W] = (J0, J1, rl)
    for W in range(T):
        chain, bin_ = 0, 0
        for w in range(T):
            if w == W:
                bin_ = chain
            chain = (flags[w] & 1) | ((flags[w] >> 1) & 1 & chain)
        sub = (flags[T - 1] >> 2) & 1 or chain == 0
        J0, J1, rl = limbs[W]
        b = bin_
        for k, j in enumerate(range(J0, J1)):
            d = rl[k] - N32[j] - b
            b = 1 if d < 0 else 0
            out[j] = (d & 0xFFFFFFFF) if sub else rl[k]
    return sum(v << (32 * j) for j, v in enumerate(out))


def _modulus(rng, nbits):
    n = rng.getrandbits(nbits) | 1 | (1 << (nbits - 1))
    return n * n


@pytest.mark.parametrize("RB,LP,T,L32,nbits", SHAPES)
def test_fixed_model_exact(RB, LP, T, L32, nbits):
    rng = random.Random(RB * LP)
    N = _modulus(rng, nbits)
    top = (1 << (32 * L32)) - 1
    for x in (0, 1, N - 1, N, N + 1, top, rng.randrange(N)):
        for y in (1, N - 1, top, rng.randrange(N)):
            assert fixed_mulmod_model(N, x, y, RB, LP, T, L32) == x * y % N, (x.bit_length(), y.bit_length())


def test_fixed_model_needs_the_top_term():
    """The top chunk's m2 N_(LP-1) lies above every chunk; at the 27-bit shape
    (the pointwise multiply's finish) a 4096-bit n^2 reaches position LP-1, so
    dropping the term breaks the result. (At 29 bits position 143 starts at bit
    4147, above every n^2 of a 2048-bit n.)"""
    rng = random.Random(3)
    RB, LP, T, L32, nbits = SHAPES[3]
    N = _modulus(rng, nbits)
    bad = 0
    for _ in range(4):
        x, y = rng.randrange(N), rng.randrange(N)
        try:
            bad += fixed_mulmod_model(N, x, y, RB, LP, T, L32, drop_top_term=True) != x * y % N
        except AssertionError:
            bad += 1
    assert bad == 4


def test_fold_intervals():
    """The served shapes fold where the accumulator bound requires: twice per
    entry at 4096 bits, once at 2048 bits, never at 27 and 28 bits."""
    assert fold_interval(29, 144) == 48 and fold_interval(29, 72) == 36
    for RB, LP in ((28, 38), (27, 152), (28, 76)):
        assert fold_interval(RB, LP) == LP
    for RB, LP, _, _, _ in SHAPES:
        seg = fold_interval(RB, LP)
        row = (1 << RB) * ((1 << RB) + (1 << (64 - 2 * RB)) + 1)
        base = (1 << RB) + (1 << (64 - RB)) + 2 * (1 << (2 * RB))
        assert seg * row + base < 1 << 64


def test_fixed_model_needs_the_folds():
    """Without folds, 144 rows of 29-bit positions overflow the 64-bit
    accumulators for a full-width entry (the planted mistake of a missing fold)."""
    rng = random.Random(9)
    RB, LP, T, L32, nbits = SHAPES[0]
    N = _modulus(rng, nbits)
    x = (1 << (32 * L32)) - 1
    with pytest.raises(AssertionError, match="overflow"):
        fixed_mulmod_model(N, x, rng.randrange(N), RB, LP, T, L32, seg=LP)
