"""Proves the input-magnitude bounds under which the HIP kernels run the
lifting operators with plain i32 arithmetic (recon.hip: kPctBound = 2^25,
kPotBound = 2^23): every intermediate of every operator, propagated as an
interval from inputs in [-B, B-1], stays inside i32.

The operators are written once here, generically over the number type; on
Python ints they are checked against the oracle (so the intervals follow
the same lifting steps), and on intervals they give the bound."""
import numpy as np
import pytest

import oracle_lib as O

I32_MIN, I32_MAX = -(2**31), 2**31 - 1


class Iv:
    """Closed integer interval; records the widest magnitude it reaches."""
    peak = 0

    def __init__(self, lo, hi):
        self.lo, self.hi = lo, hi
        Iv.peak = max(Iv.peak, abs(lo), abs(hi))

    def __add__(self, o):
        o = o if isinstance(o, Iv) else Iv(o, o)
        return Iv(self.lo + o.lo, self.hi + o.hi)

    __radd__ = __add__

    def __sub__(self, o):
        o = o if isinstance(o, Iv) else Iv(o, o)
        return Iv(self.lo - o.hi, self.hi - o.lo)

    def __rsub__(self, o):
        return Iv(o, o) - self

    def __mul__(self, k):
        assert isinstance(k, int) and k >= 0
        return Iv(self.lo * k, self.hi * k)

    def __rshift__(self, k):
        return Iv(self.lo >> k, self.hi >> k)

    def __neg__(self):
        return Iv(-self.hi, -self.lo)


# ---- operators, generic over int / Iv --------------------------------------
def t2x2h(v, a, b, c, d, r):
    v[a] = v[a] + v[d]
    v[b] = v[b] - v[c]
    first = (v[a] - v[b] + r) >> 1
    second = v[c]
    v[c] = first - v[d]
    v[d] = first - second
    v[a] = v[a] - v[d]
    v[b] = v[b] + v[c]


def inv_todd(v, i0, i1, i2, i3):
    v[i1] = v[i1] + v[i3]
    v[i0] = v[i0] - v[i2]
    v[i3] = v[i3] - (v[i1] >> 1)
    v[i2] = v[i2] + ((v[i0] + 1) >> 1)
    v[i0] = v[i0] - ((v[i1] * 3 + 4) >> 3)
    v[i1] = v[i1] + ((v[i0] * 3 + 4) >> 3)
    v[i2] = v[i2] - ((v[i3] * 3 + 4) >> 3)
    v[i3] = v[i3] + ((v[i2] * 3 + 4) >> 3)
    v[i2] = v[i2] - ((v[i1] + 1) >> 1)
    v[i3] = ((v[i0] + 1) >> 1) - v[i3]
    v[i1] = v[i1] + v[i2]
    v[i0] = v[i0] - v[i3]


def inv_todd_odd(v, i0, i1, i2, i3):
    v[i3] = v[i3] + v[i0]
    v[i2] = v[i2] - v[i1]
    first, second = v[i3] >> 1, v[i2] >> 1
    v[i0] = v[i0] - first
    v[i1] = v[i1] + second
    v[i0] = v[i0] - ((v[i1] * 3 + 3) >> 3)
    v[i1] = v[i1] + ((v[i0] * 3 + 3) >> 2)
    v[i0] = v[i0] - ((v[i1] * 3 + 4) >> 3)
    v[i1] = v[i1] - second
    v[i0] = v[i0] + first
    v[i2] = v[i2] + v[i1]
    v[i3] = v[i3] - v[i0]
    v[i1] = -v[i1]
    v[i2] = -v[i2]


PERM = [0, 8, 4, 13, 2, 15, 3, 14, 1, 12, 5, 9, 7, 11, 6, 10]


def inverse_core(x):
    v = [None] * 16
    for i in range(16):
        v[PERM[i]] = x[i]
    t2x2h(v, 0, 1, 4, 5, 1)
    inv_todd(v, 2, 3, 6, 7)
    inv_todd(v, 8, 12, 9, 13)
    inv_todd_odd(v, 10, 11, 14, 15)
    t2x2h(v, 0, 3, 12, 15, 0)
    t2x2h(v, 5, 6, 9, 10, 0)
    t2x2h(v, 1, 2, 13, 14, 0)
    t2x2h(v, 4, 7, 8, 11, 0)
    return v


def inverse_chroma420(x):
    v = list(x)
    t2x2h(v, 0, 1, 2, 3, 0)
    v[1], v[2] = v[2], v[1]
    return v


def inverse_chroma422(x):
    v = list(x)
    v[0] = v[0] - ((v[4] + 1) >> 1)
    v[4] = v[4] + v[0]
    t2x2h(v, 0, 1, 2, 3, 0)
    v[1], v[2] = v[2], v[1]
    t2x2h(v, 4, 6, 5, 7, 0)
    v[5], v[6] = v[6], v[5]
    return v


def inv_rotate(v, a, b):
    v[a] = v[a] - ((v[b] + 1) >> 1)
    v[b] = v[b] + ((v[a] + 1) >> 1)


def inv_scale(v, a, b):
    v[a] = v[a] + v[b]
    v[b] = (v[a] >> 1) - v[b]
    v[a] = v[a] + ((v[b] * 3) >> 3)
    v[b] = v[b] + ((v[a] * 3) >> 4)
    v[b] = v[b] + (v[a] >> 7)
    v[b] = v[b] - (v[a] >> 10)


def inverse_overlap4(x):
    v = list(x)
    v[0] = v[0] + v[3]
    v[1] = v[1] + v[2]
    v[3] = v[3] - ((v[0] + 1) >> 1)
    v[2] = v[2] - ((v[1] + 1) >> 1)
    inv_scale(v, 0, 3)
    inv_scale(v, 1, 2)
    v[0] = v[0] + ((v[3] * 3 + 4) >> 3)
    v[1] = v[1] + ((v[2] * 3 + 4) >> 3)
    v[3] = v[3] - (v[0] >> 1)
    v[2] = v[2] - (v[1] >> 1)
    v[0] = v[0] + v[3]
    v[1] = v[1] + v[2]
    v[3] = -v[3]
    v[2] = -v[2]
    inv_rotate(v, 2, 3)
    v[3] = v[3] + ((v[0] + 1) >> 1)
    v[2] = v[2] + ((v[1] + 1) >> 1)
    v[0] = v[0] - v[3]
    v[1] = v[1] - v[2]
    return v


def hadamard_post(v, a, b, c, d):
    v[b] = v[b] - v[c]
    v[a] = v[a] + ((v[d] * 3 + 4) >> 3)
    v[d] = v[d] - (v[b] >> 1)
    v[c] = ((v[a] - v[b]) >> 1) - v[c]
    v[c], v[d] = v[d], v[c]
    v[a] = v[a] - v[d]
    v[b] = v[b] + v[c]


def todd_odd_post(v, i0, i1, i2, i3):
    v[i3] = v[i3] + v[i0]
    v[i2] = v[i2] - v[i1]
    first, second = v[i3] >> 1, v[i2] >> 1
    v[i0] = v[i0] - first
    v[i1] = v[i1] + second
    v[i0] = v[i0] - ((v[i1] * 3 + 6) >> 3)
    v[i1] = v[i1] + ((v[i0] * 3 + 2) >> 2)
    v[i0] = v[i0] - ((v[i1] * 3 + 4) >> 3)
    v[i1] = v[i1] - second
    v[i0] = v[i0] + first
    v[i2] = v[i2] + v[i1]
    v[i3] = v[i3] - v[i0]


GROUPS = [(0, 3, 12, 15), (1, 2, 13, 14), (4, 7, 8, 11), (5, 6, 9, 10)]


def inverse_overlap4x4(x):
    v = list(x)
    for g in GROUPS:
        t2x2h(v, *g, 0)
    for a, b in [(13, 12), (9, 8), (7, 3), (6, 2)]:
        inv_rotate(v, a, b)
    todd_odd_post(v, 10, 11, 14, 15)
    for a, b in [(0, 15), (1, 14), (4, 11), (5, 10)]:
        inv_scale(v, a, b)
    for g in GROUPS:
        hadamard_post(v, *g)
    return v


def inv2(v, a, b):
    v[b] = v[b] + ((v[a] + 2) >> 2)
    v[a] = v[a] + ((v[b] + 1) >> 1)
    v[a] = v[a] + (v[b] >> 5)
    v[a] = v[a] + (v[b] >> 9)
    v[a] = v[a] + (v[b] >> 13)
    v[b] = v[b] + ((v[a] + 2) >> 2)


def inverse_overlap2(x):
    v = list(x)
    inv2(v, 0, 1)
    return v


def inverse_overlap2x2(x):
    v = list(x)
    v[0] = v[0] + v[3]
    v[1] = v[1] + v[2]
    v[3] = v[3] - ((v[0] + 1) >> 1)
    v[2] = v[2] - ((v[1] + 1) >> 1)
    inv2(v, 0, 1)
    v[3] = v[3] + ((v[0] + 1) >> 1)
    v[2] = v[2] + ((v[1] + 1) >> 1)
    v[0] = v[0] - v[3]
    v[1] = v[1] - v[2]
    return v


OPS = [
    (inverse_core, 16, "jxro_inverse_core_transform", 25),
    (inverse_chroma420, 4, "jxro_inverse_chroma_420", 25),
    (inverse_chroma422, 8, "jxro_inverse_chroma_422", 25),
    (inverse_overlap4, 4, "jxro_inverse_overlap_4", 23),
    (inverse_overlap4x4, 16, "jxro_inverse_overlap_4x4", 23),
    (inverse_overlap2, 2, "jxro_inverse_overlap_2", 23),
    (inverse_overlap2x2, 4, "jxro_inverse_overlap_2x2", 23),
]


@pytest.mark.parametrize("fn,n,oracle,bits", OPS, ids=[o[2] for o in OPS])
def test_generic_operator_matches_oracle(fn, n, oracle, bits):
    rng = np.random.default_rng(n * 31 + bits)
    for _ in range(300):
        x = [int(v) for v in rng.integers(-(2**bits), 2**bits, n)]
        want, err = O.op(oracle, x)
        assert not err
        assert [int(v) for v in fn(x)] == want.tolist()


@pytest.mark.parametrize("fn,n,oracle,bits", OPS, ids=[o[2] for o in OPS])
def test_fast_path_bound_keeps_every_intermediate_in_i32(fn, n, oracle, bits):
    B = 2**bits
    Iv.peak = 0
    out = fn([Iv(-B, B - 1) for _ in range(n)])
    assert Iv.peak <= I32_MAX, f"{oracle}: intermediate reaches {Iv.peak} for inputs below 2^{bits}"
    assert all(isinstance(o, Iv) for o in out)


# ---- gains: output magnitude per input bound --------------------------------
def _peak_out(fn, n, B):
    out = fn([Iv(-B, B) for _ in range(n)])
    return max(max(abs(o.lo), abs(o.hi)) for o in out)


@pytest.mark.parametrize("B", [1, 2, 3, 7, 100, 1023, 2**15, 2**19 + 5, 2**20])
def test_second_level_gain_bound(B):
    """recon.hip stands 64 x the input bound in for the magnitudes of the
    second-level transform's outputs when inputs are below 2^20
    (kGainBoundInput, kPctGainShift): outputs stay below 64 x the bound at
    every scale (the gain approaches 35.3 for large inputs; rounding adds
    a little for tiny ones), so below 2^26."""
    peak = _peak_out(inverse_core, 16, B)
    assert peak < 64 * B
    if B >= 100:
        assert peak <= 36 * B


@pytest.mark.parametrize("fn,n", [(inverse_overlap4x4, 16), (inverse_overlap4, 4), (inverse_overlap2x2, 4),
                                  (inverse_overlap2, 2)], ids=["4x4", "4", "2x2", "2"])
@pytest.mark.parametrize("B", [1, 2, 5, 100, 2**12, 2**18])
def test_overlap_gain_bound(fn, n, B):
    """recon.hip stands 128 x (input OR + 1) in for the magnitudes of the
    sample-level overlap's outputs when the OR of the inputs' magnitudes is
    below 2^18 (kPotGainInput, kPotGainShift): |input| <= OR + 1 = B, and the
    outputs stay below 128 B."""
    assert _peak_out(fn, n, B) < 128 * B
    assert 128 * 2**18 < 2**26


def test_gain_bound_is_not_loose_by_construction():
    """The bound above is met with little room: a 16-input interval's output
    reaches more than 30 x its bound, so a transform with a larger gain (a
    dropped halving, say) would fail it."""
    assert _peak_out(inverse_core, 16, 2**20) > 30 * 2**20


# ---- output stage under plain arithmetic ------------------------------------
def _ceil_half(v):
    # (x >> 1) + (x & 1), monotone in x
    return Iv(-((-v.lo) >> 1), -((-v.hi) >> 1))


def _weighted_fast(a, b, wa):
    """b + (((a - b) * wa + 4) >> 3), wa in 0..8 (recon.hip `weighted`,
    Fast). The intermediates are propagated as intervals; the result is
    floor((a*wa + b*(8-wa) + 4) / 8), a rounded convex combination, so it
    lies between min(a, b) and max(a, b) (which interval arithmetic, losing
    the correlation of the two uses of b, cannot see; checked by brute force
    in test_weighted_stays_between_its_inputs)."""
    _ = ((a - b) * wa + 4) >> 3
    return Iv(min(a.lo, b.lo), max(a.hi, b.hi))


def test_weighted_stays_between_its_inputs():
    for a in range(-20, 21):
        for b in range(-20, 21):
            for wa in range(9):
                r = b + (((a - b) * wa + 4) >> 3)
                assert min(a, b) <= r <= max(a, b)


def _upsampled(S):
    """Every chroma value the plain output path computes from samples in S:
    two passes of weighted picks (centered) or the centering-0 averages."""
    outs = []
    for wa in range(9):
        once = _weighted_fast(S, S, wa)
        outs.append(_weighted_fast(once, once, wa))
    avg = (S + S + 1) >> 1
    outs.append((avg + avg + 1) >> 1)
    lo = min(o.lo for o in outs)
    hi = max(o.hi for o in outs)
    return Iv(lo, hi)


def test_plain_output_arithmetic_stays_in_i32():
    """The output stage runs plain i32 on a tile when every sample is below
    2^27 in magnitude and every channel's |bias| + rounding is below 2^29
    (kOutputSampleBound, out_fast): chroma interpolation, the YUV/YUVK
    inverses and the scale's additions then stay inside i32."""
    Iv.peak = 0
    S = Iv(-(2**27), 2**27 - 1)
    C = _upsampled(S)
    bias = Iv(-(2**29), 2**29)
    outs = []
    # YUV -> RGB (output_format/color.rs:100-110)
    y, u, v = S, C, C
    t0 = -u
    g = y - (t0 >> 1)
    rr = (t0 + g) - _ceil_half(v)
    outs += [rr, g, v + rr]
    # YUVK -> CMYK (color.rs:112-123)
    k = S
    black = k + (y >> 1)
    magenta = (black - y) - (u >> 1)
    cyan = (u + magenta) + (v >> 1)
    outs += [cyan, magenta, cyan - v, black]
    for o in outs + [S, C]:
        _ = o + bias  # bias and rounding, added before the shift
    assert Iv.peak <= I32_MAX


def test_plain_output_bound_is_tight_enough_to_matter():
    """Samples of 2^29 would not be safe under the same arithmetic, so the
    2^27 limit is doing work."""
    Iv.peak = 0
    S = Iv(-(2**29), 2**29 - 1)
    t0 = -S
    g = S - (t0 >> 1)
    rr = (t0 + g) - _ceil_half(S)
    _ = (S + rr) + Iv(-(2**29), 2**29)
    assert Iv.peak > I32_MAX
