"""Pins the oracle's image, projector and backprojector arithmetic (src/image.h, src/reconstruct.cpp).

Independent checks: the identity shift is a copy (src/image.h:1517-1524); tabulated phases stay
within the 2*pi/5000 table step of the exact phase with the reference's sign convention; down-
windowing keeps logical frequencies; trilinear / bilinear interpolation reproduces linear functions
exactly and samples the grid exactly at integer coordinates; the backprojector is the exact adjoint
of the projector on the pixels both visit; trilinear weights conserve the total weight.
"""
import numpy as np
import pytest

import oracle as O
from rome_amd import synth


def test_zero_shift_is_copy():
    rng = np.random.default_rng(0)
    n = 16 * 9
    re, im = rng.normal(size=n), rng.normal(size=n)
    ore, oim = O.shift_image(re, im, 16, 0.0, 0.0, 32)
    assert np.array_equal(ore, re) and np.array_equal(oim, im)


@pytest.mark.parametrize("sx,sy", [(1.0, 0.0), (-2.3, 0.7), (0.25, -4.75)])
def test_shift_phase_within_table_step(sx, sy):
    size, ori = 24, 64
    x, y = synth.halfplane_xy(size)
    re, im = np.ones(len(x)), np.zeros(len(x))
    ore, oim = O.shift_image(re, im, size, sx, sy, ori)
    exact = np.exp(-2j * np.pi * (x * sx + y * sy) / ori)
    got = ore + 1j * oim
    assert np.max(np.abs(np.abs(got) - 1)) < 1e-12
    assert np.max(np.abs(np.angle(got * np.conj(exact)))) <= 2 * np.pi / 5000 + 1e-12


def test_window_keeps_frequencies():
    rng = np.random.default_rng(1)
    big, small = 32, 12
    xb, yb = synth.halfplane_xy(big)
    re = rng.normal(size=len(xb))
    im = rng.normal(size=len(xb))
    ore, oim = O.window_fourier(re, im, big, small)
    xs, ys = synth.halfplane_xy(small)
    lut = {(a, b): q for q, (a, b) in enumerate(zip(xb, yb))}
    for q, (a, b) in enumerate(zip(xs, ys)):
        assert ore[q] == re[lut[(a, b)]] and oim[q] == im[lut[(a, b)]]


def test_mresol_definition():
    for size in (16, 24, 64):
        x, y = synth.halfplane_xy(size)
        r = np.sqrt(x * x + y * y)
        want = np.where((np.round(r) < size // 2 + 1) & ~((x == 0) & (y < 0)), np.round(r), -1)
        assert np.array_equal(O.mresol(size), want.astype(np.int32))


def _linear_grid(pad, coef, is3d=True):
    c = pad // 2
    padF = pad // 2 + 1
    if is3d:
        z, y, x = np.meshgrid(np.arange(pad) - c, np.arange(pad) - c, np.arange(padF), indexing="ij")
        v = coef[0] + coef[1] * x + coef[2] * y + coef[3] * z
    else:
        y, x = np.meshgrid(np.arange(pad) - c, np.arange(padF), indexing="ij")
        v = coef[0] + coef[1] * x + coef[2] * y
    g = np.zeros(v.shape + (2,))
    g[..., 0] = v
    g[..., 1] = 0.5 * v
    return g.reshape(-1)


@pytest.mark.parametrize("is3d", [True, False])
def test_projector_reproduces_linear_functions(is3d):
    r_max, size = 8, 16
    pad = 2 * (2 * r_max + 1) + 1
    coef = (0.3, 0.7, -1.1, 0.45)
    g = _linear_grid(pad, coef, is3d)
    rng = np.random.default_rng(2)
    for _ in range(5):
        A = synth.euler_matrix(*(rng.uniform(-180, 180, 3) if is3d else (0, 0, rng.uniform(0, 360))))
        fn = O.project3d if is3d else O.rotate2d
        re, im = fn(g, pad, r_max, A, size)
        x, y = synth.halfplane_xy(size)
        F = size // 2 + 1
        i = np.arange(size * F) // F
        ok = ((i <= r_max) | (i >= size - r_max)) & (x * x + y * y <= r_max * r_max)
        p = 2.0 * (A.T @ np.vstack([x, y, np.zeros_like(x)]))
        neg = p[0] < 0
        p[:, neg] *= -1
        v = coef[0] + coef[1] * p[0] + coef[2] * p[1] + (coef[3] * p[2] if is3d else 0)
        np.testing.assert_allclose(re[ok], v[ok], atol=1e-10)
        np.testing.assert_allclose(im[ok], np.where(neg, -0.5, 0.5)[ok] * v[ok], atol=1e-10)
        assert np.all(re[~ok] == 0) and np.all(im[~ok] == 0)


def test_projector_identity_samples_grid():
    r_max, size = 8, 16
    pad = 2 * (2 * r_max + 1) + 1
    padF, c = pad // 2 + 1, pad // 2
    rng = np.random.default_rng(3)
    g = rng.normal(size=(pad, pad, padF, 2))
    re, im = O.project3d(g.reshape(-1), pad, r_max, np.eye(3), size)
    x, y = synth.halfplane_xy(size)
    for q in range(len(x)):
        if x[q] * x[q] + y[q] * y[q] <= r_max * r_max and (q // (size // 2 + 1) <= r_max or
                                                            q // (size // 2 + 1) >= size - r_max):
            assert re[q] == g[c, c + 2 * y[q], 2 * x[q], 0] and im[q] == g[c, c + 2 * y[q], 2 * x[q], 1]


@pytest.mark.parametrize("is3d", [True, False])
def test_backprojector_is_adjoint_of_projector(is3d):
    r_max, size = 8, 16
    pad = 2 * (2 * r_max + 1) + 1
    padF = pad // 2 + 1
    gridn = pad * pad * padF if is3d else pad * padF
    rng = np.random.default_rng(4)
    g = rng.normal(size=(gridn, 2))
    x, y = synth.halfplane_xy(size)
    F = size // 2 + 1
    u = rng.normal(size=(len(x), 2))
    u[(x == 0) & (np.arange(len(x)) // F > r_max)] = 0  # the backprojector skips the x=0, y<0 half column
    for _ in range(4):
        A = synth.euler_matrix(*(rng.uniform(-180, 180, 3) if is3d else (0, 0, rng.uniform(0, 360))))
        fp = O.project3d if is3d else O.rotate2d
        bp = O.backproject3d if is3d else O.backrotate2d
        re, im = fp(g.reshape(-1), pad, r_max, A, size)
        lhs = np.sum(re * u[:, 0] + im * u[:, 1])
        data = np.zeros(2 * gridn)
        wgt = np.zeros(gridn)
        mw = np.ones(len(x))
        bp(data, wgt, pad, r_max, A, size, u[:, 0].copy(), u[:, 1].copy(), mw)
        rhs = np.sum(g.reshape(-1) * data)
        assert abs(lhs - rhs) <= 1e-9 * (abs(lhs) + 1)
        # trilinear / bilinear weights sum to one per visited pixel
        visited = (x * x + y * y <= r_max * r_max) & ~((x == 0) & (np.arange(len(x)) // F > r_max))
        assert abs(wgt.sum() - visited.sum()) < 1e-9
