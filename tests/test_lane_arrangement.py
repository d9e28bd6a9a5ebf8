"""The algebra behind the staged kR->kq kernel's column loop, checked on the CPU (no GPU, no kernel).

The kernel (csrc/gridopt.hip, gridopt_kq_rotate_kernel) forms, per kept q and row i of a 16-lane row
group, g_i = sum_c e3[c] sum_b e2[b] B1_i[c][b] and then the row sums of conj(w_i) g_i over i with DPP
lane exchanges. On boxes symmetric about 0 along r2 and r3 it relies on three identities, checked here
on random data so that a wrong partner lane, sign or mirror fails:

* the folded sum (kq_folded_sum after kq_fold) equals the plain phase sum, and the fold over c may be
  applied to the stage-1 sums ahead of stage 2;
* the lanes that keep the imaginary part hold i conj(B1) mirrored over b and c, which makes their stage
  sums i conj(g) = (g_i, g_r), and read conj(w): their w.x p1 + w.y p2 is then Im(conj(w) g) and
  w.x p2 - w.y p1 is Re(conj(w) g), the same formulas as the other lanes with (p1, p2) = (g_r, g_i);
* the DPP networks (row rotation by 8, half-row mirror, quad permutations) leave the row sums in every
  lane, where the kernel stores them from: lanes with bit 2 clear / set hold re / im (one column, parts by
  bit 2), the same by bit 3 (one column, parts by bit 3), lanes 4s .. 4s + 3 value s (two columns); the
  lanes holding one value store it together.
"""
import numpy as np
import pytest

RNG = np.random.default_rng(20261203)


def _phases(q, m, n):
    """e[x] = e^{2 pi i q (x - h) / m}, x = 0 .. n-1, on the box -h .. h (n = 2h + 1)."""
    h = n // 2
    return np.exp(2j * np.pi * q * (np.arange(n) - h) / m)


def _fold(v):
    """kq_fold along the last axis: v[h+j] <- v[h+j] + v[h-j], v[h-j] <- v[h+j] - v[h-j]."""
    v = v.copy()
    h = v.shape[-1] // 2
    for j in range(1, h + 1):
        p, q = v[..., h + j].copy(), v[..., h - j].copy()
        v[..., h + j], v[..., h - j] = p + q, p - q
    return v


def _folded_sum(e, v):
    """kq_folded_sum: v folded (_fold), e the phases of the positive offsets e[h + j] = cos_j + i sin_j."""
    h = v.shape[-1] // 2
    s = v[..., h].copy()
    for j in range(1, h + 1):
        c, sn = e[h + j].real, e[h + j].imag
        s = s + c * v[..., h + j] + 1j * sn * v[..., h - j]
    return s


def _cplx(shape):
    return RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)


@pytest.mark.parametrize("n", [3, 5, 7])
def test_folded_sums_and_fold_ahead_of_stage2(n):
    m2, m3 = 24, 17
    B1 = _cplx((n, n))                    # [c][b]
    for q2, q3 in [(0, 0), (5, 3), (23, 16), (12, 8)]:
        e2, e3 = _phases(q2, m2, n), _phases(q3, m3, n)
        g = np.sum(e3[:, None] * e2[None, :] * B1)
        # stage 2 per c on the b-folded sums, then stage 3 on the c-folded result
        b2 = _folded_sum(e2, _fold(B1))
        assert abs(_folded_sum(e3, _fold(b2)) - g) < 1e-12 * np.abs(B1).sum()
        # the fold over c applied to the stage-1 sums first (once per workgroup): same g
        b2_pre = _folded_sum(e2, _fold(_fold(B1.T).T))
        assert abs(_folded_sum(e3, b2_pre) - g) < 1e-12 * np.abs(B1).sum()


@pytest.mark.parametrize("n", [3, 5, 7])
def test_mirrored_conjugate_stage1_gives_i_conj_g(n):
    m2, m3 = 24, 24
    B1 = _cplx((n, n))
    B1s = 1j * np.conj(B1[::-1, ::-1])    # (im, re) of the sums at (N-1-c, N-1-b)
    assert np.allclose(B1s, B1[::-1, ::-1].imag + 1j * B1[::-1, ::-1].real)
    for q2, q3 in [(1, 2), (7, 19), (23, 0)]:
        e2, e3 = _phases(q2, m2, n), _phases(q3, m3, n)
        assert np.allclose(e2[::-1], np.conj(e2))
        g = np.sum(e3[:, None] * e2[None, :] * B1)
        gs = np.sum(e3[:, None] * e2[None, :] * B1s)
        assert abs(gs - 1j * np.conj(g)) < 1e-12 * np.abs(B1).sum()
        assert np.isclose(gs.real, g.imag) and np.isclose(gs.imag, g.real)
    # the identity needs the symmetric box: on a box 0 .. n-1 the mirror adds the phase
    # e^{2 pi i (q2 / m2 + q3 / m3)(n - 1)}, not 1 for these q
    e2 = np.exp(2j * np.pi * 5 * np.arange(n) / m2)
    e3 = np.exp(2j * np.pi * 2 * np.arange(n) / m3)
    g = np.sum(e3[:, None] * e2[None, :] * B1)
    gs = np.sum(e3[:, None] * e2[None, :] * B1s)
    assert abs(gs - 1j * np.conj(g)) > 1e-3


def _lane_parts(g, w, swap):
    """(x0, x1) of one lane: x0 = w.x p1 + w.y p2 (the part it keeps), x1 = w.x q1 + w.y q2 (the other),
    with (q1, q2) = (p2, -p1); swap lanes hold p = (g_i, g_r) and read conj(w)."""
    if swap:
        p1, p2, wx, wy = g.imag, g.real, w.real, -w.imag
    else:
        p1, p2, wx, wy = g.real, g.imag, w.real, w.imag
    return wx * p1 + wy * p2, wx * p2 - wy * p1


def test_lane_parts_give_re_and_im():
    for _ in range(20):
        g, w = _cplx(()), _cplx(())
        a = np.conj(w) * g
        x0, x1 = _lane_parts(g, w, swap=False)
        assert np.isclose(x0, a.real) and np.isclose(x1, a.imag)
        x0, x1 = _lane_parts(g, w, swap=True)
        assert np.isclose(x0, a.imag) and np.isclose(x1, a.real)


# DPP lane exchanges inside a 16-lane row: lane l receives the value of lane src(l)
def _take(x, src):
    return np.array([x[src(l)] for l in range(16)])


ROR8 = lambda l: (l + 8) & 15                       # noqa: E731
HALF_MIRROR = lambda l: (l & 8) | (7 - (l & 7))      # noqa: E731
QUAD_2301 = lambda l: l ^ 2                           # noqa: E731
QUAD_1032 = lambda l: l ^ 1                           # noqa: E731


def test_one_column_row_sum_parts_by_bit2():
    """dpp_row_sum2_bit2 with the lane arrangement by bit 2 (width 2, one band at k+q): the lanes with
    bit 2 clear end with re, those with it set with im of sum_i conj(w_i) g_i."""
    g, w = _cplx(16), _cplx(16)
    target = np.sum(np.conj(w) * g)
    swap = (np.arange(16) & 4) != 0
    x0, x1 = np.array([_lane_parts(g[l], w[l], swap[l]) for l in range(16)]).T
    x = x0 + _take(x1, HALF_MIRROR)
    x = x + _take(x, QUAD_2301)
    x = x + _take(x, QUAD_1032)
    x = x + _take(x, ROR8)
    assert np.allclose(x, np.where(np.arange(16) & 4, target.imag, target.real))
    # a partner that keeps the same part (row rotation first, as for bit 3) would not sum the row
    y = x0 + _take(x1, ROR8)
    assert not np.isclose(y[0] + y[1] + y[2] + y[3], target.real)


def test_one_column_row_sum_parts_by_bit3():
    """dpp_row_sum2 with the lane arrangement by bit 3 (odd widths, one band at k+q): lanes 0 .. 7 end
    with re, lanes 8 .. 15 with im."""
    g, w = _cplx(16), _cplx(16)
    target = np.sum(np.conj(w) * g)
    swap = (np.arange(16) & 8) != 0
    x0, x1 = np.array([_lane_parts(g[l], w[l], swap[l]) for l in range(16)]).T
    x = x0 + _take(x1, ROR8)
    x = x + _take(x, HALF_MIRROR)
    x = x + _take(x, QUAD_2301)
    x = x + _take(x, QUAD_1032)
    assert np.allclose(x, np.where(np.arange(16) & 8, target.imag, target.real))


def test_pair_row_sums():
    """dpp_row_sum4_pairs: two columns a, b; lanes with bit 3 set take b as their own column (staged in
    that order), lanes with bit 2 set keep the imaginary part (arranged as above). Lanes 4s end with
    value s of (re a, im a, re b, im b) in lanes 4s .. 4s + 3."""
    g, wa, wb = _cplx(16), _cplx(16), _cplx(16)
    want = [np.sum(np.conj(wa) * g).real, np.sum(np.conj(wa) * g).imag,
            np.sum(np.conj(wb) * g).real, np.sum(np.conj(wb) * g).imag]
    x0 = np.zeros(16)
    x1 = np.zeros(16)
    y0 = np.zeros(16)
    y1 = np.zeros(16)
    for l in range(16):
        own, other = (wb[l], wa[l]) if l & 8 else (wa[l], wb[l])
        x0[l], x1[l] = _lane_parts(g[l], own, bool(l & 4))
        y0[l], y1[l] = _lane_parts(g[l], other, bool(l & 4))
    x0 = x0 + _take(y0, ROR8)
    x1 = x1 + _take(y1, ROR8)
    x = x0 + _take(x1, HALF_MIRROR)
    x = x + _take(x, QUAD_2301)
    x = x + _take(x, QUAD_1032)
    for s in range(4):
        assert np.allclose(x[4 * s:4 * s + 4], want[s]), s
