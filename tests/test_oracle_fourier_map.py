"""The oracle's computeFourierTransformMap (src/reconstruct.cpp:106-223) pinned by closed forms:
a centred delta gives a constant grid, a shifted delta a pure phase ramp (sign, axis order, centring
and normalisation), a direct DFT on a tiny map agrees, and griddingCorrect divides by sinc^2."""
import math

import numpy as np
import pytest

import oracle.mstep as OM


@pytest.mark.parametrize("dim", [2, 3])
def test_centred_delta_is_flat(dim):
    ori, cur = 16, 16
    vol = np.zeros((ori,) * dim)
    vol[(ori // 2,) * dim] = 1.0
    grid, power, corr = OM.fourier_map(vol, cur)
    pad = 2 * (2 * (cur // 2) + 1) + 1
    assert grid.shape == ((pad, pad, pad // 2 + 1) if dim == 3 else (pad, pad // 2 + 1))
    normfft = 8.0 * ori if dim == 3 else 4.0
    val = normfft / (2 * ori) ** dim  # delta at the padded centre -> 1 everywhere, / N, * normfft
    inside = np.abs(grid) > 0
    np.testing.assert_allclose(grid[inside], val, rtol=1e-13, atol=1e-15)
    # the window keeps |k|^2 <= (2 r_max)^2 exactly; an FFT of 2*ori points has no -ori frequency
    # (it is +ori), so that row / plane stays empty
    c = pad // 2
    idx = np.indices(grid.shape)
    r2 = sum((idx[a] - c) ** 2 for a in range(dim - 1)) + idx[-1] ** 2
    exists = np.all([idx[a] - c > -ori for a in range(dim - 1)], axis=0)
    assert np.array_equal(inside, (r2 <= (2 * (cur // 2)) ** 2) & exists)
    np.testing.assert_allclose(power[: ori // 2 + 1], val * val / 2.0, rtol=1e-13)
    assert corr[(ori // 2,) * dim] == 1.0  # r = 0 is not corrected


def test_shifted_delta_is_a_phase_ramp():
    ori = 16
    vol = np.zeros((ori, ori, ori))
    dz, dy, dx = 1, -2, 3
    vol[ori // 2 + dz, ori // 2 + dy, ori // 2 + dx] = 1.0
    grid, _, corr = OM.fourier_map(vol, ori, do_gridding=False)
    N = 2 * ori
    pad = grid.shape[0]
    c = pad // 2
    for (z, y, x) in [(c, c, 1), (c + 1, c - 2, 0), (c - 3, c + 2, 4), (c, c + 5, 7)]:
        kz, ky, kx = z - c, y - c, x
        want = 8.0 * ori / N ** 3 * np.exp(-2j * math.pi * (kz * dz + ky * dy + kx * dx) / N)
        assert abs(grid[z, y, x] - want) < 1e-15


def test_tiny_map_matches_direct_dft():
    rng = np.random.default_rng(3)
    ori = 4
    vol = rng.normal(size=(ori, ori, ori))
    grid, _, corr = OM.fourier_map(vol, ori)
    N = 2 * ori
    # direct DFT of the corrected map placed at the padded centre, origin at index N/2
    pos = np.arange(ori) - ori // 2
    pad = grid.shape[0]
    c = pad // 2
    zz, yy, xx = np.meshgrid(pos, pos, pos, indexing="ij")
    for z in range(pad):
        for y in range(pad):
            for x in range(pad // 2 + 1):
                kz, ky, kx = z - c, y - c, x
                if kz * kz + ky * ky + kx * kx > (2 * (ori // 2)) ** 2 or kz == -ori or ky == -ori:
                    assert grid[z, y, x] == 0  # outside the window, or the frequency -N/2 (stored as +N/2)
                    continue
                ph = np.exp(-2j * math.pi * (kz * zz + ky * yy + kx * xx) / N)
                want = 8.0 * ori / N ** 3 * np.sum(corr * ph)
                assert abs(grid[z, y, x] - want) < 1e-13 * (1 + abs(want))


def test_gridding_correction_is_sinc_squared():
    ori = 8
    vol = np.ones((ori, ori, ori))
    corr = OM.gridding_correct(vol)
    for (k, i, j) in [(0, 0, 0), (4, 4, 5), (4, 7, 4), (1, 2, 3)]:
        r = math.sqrt((k - 4) ** 2 + (i - 4) ** 2 + (j - 4) ** 2)
        s = math.sin(math.pi * r / 16) / (math.pi * r / 16)
        assert abs(corr[k, i, j] - 1.0 / (s * s)) < 1e-14
    assert corr[4, 4, 4] == 1.0
