"""The oracle's M-step reconstruction (MyBackProjector::reconstruct, src/reconstruct.cpp:1929-2522)
pinned piecewise: the blob table against the closed form of the Kaiser-Bessel transform, the shell
spectra and the MAP term by closed forms on constant weights, the Hermitian fold of the x = 0 plane,
and the end-to-end invariant that a map sampled on the grid (unit weights, no regularisation, no
gridding correction) comes back after reconstruction."""
import math

import numpy as np
import pytest

import oracle.mstep as OM


def test_blob_table_is_the_kaiser_bessel_transform():
    table, sampling = OM.tab_ftblob(3.8, 15.0, 0)
    assert len(table) == 10000 and sampling == 0.5 / 10000
    # order-0 KB blob transform ~ I_{3/2}(sig) / sig^{3/2}, sig = sqrt(alpha^2 - (2 pi a w)^2)
    for i in (0, 17, 2500, 9999):
        w = i * sampling
        sig = math.sqrt(15.0 ** 2 - (2 * math.pi * 3.8 * w) ** 2)
        i15 = math.sqrt(2 / (math.pi * sig)) * (math.cosh(sig) - math.sinh(sig) / sig)
        s0 = 15.0
        i15_0 = math.sqrt(2 / (math.pi * s0)) * (math.cosh(s0) - math.sinh(s0) / s0)
        assert abs(table[i] / table[0] - (i15 / sig ** 1.5) / (i15_0 / s0 ** 1.5)) < 1e-12
    assert np.all(np.diff(table[:5000]) < 0)  # decreasing away from 0


def _grids(P, PF, dim, rng, w=1.0):
    shape = (P, P, PF) if dim == 3 else (P, PF)
    data = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    return data, np.full(shape, w)


@pytest.mark.parametrize("dim", [2, 3])
def test_constant_weight_spectra(dim):
    ori, cur = 16, 16
    P = 2 * (2 * (cur // 2) + 1) + 1
    rng = np.random.default_rng(1)
    data, weight = _grids(P, P // 2 + 1, dim, rng, w=2.5)
    tau2 = np.full(ori // 2 + 1, 4.0)
    _, t2, sigma2, dvp = OM.reconstruct(data, weight, ori, cur, max_iter_preweight=1, do_map=True, tau2_fudge=1.5,
                                        tau2=tau2, normalise=3.0, minres_map=2)
    osc = 8.0 if dim == 3 else 4.0
    # shells by direct enumeration of the centred half grid: after the Hermitian fold every x = 0 point
    # but the origin carries the weight of its Friedel pair as well (2w)
    c, max_r2 = P // 2, (2 * (cur // 2)) ** 2
    n = np.zeros(ori // 2 + 1)
    S = np.zeros(ori // 2 + 1)
    zr = range(-c, c + 1) if dim == 3 else [0]
    for z in zr:
        for y in range(-c, c + 1):
            for x in range(0, c + 1):
                r2 = z * z + y * y + x * x
                if r2 < max_r2:
                    sh = int(math.floor(math.sqrt(r2) / 2 + 0.5))
                    n[sh] += 1
                    S[sh] += 2.5 * (2 if (x == 0 and (z, y) != (0, 0)) else 1)
    reached = n > 0
    # sigma2 = count / sum(osc * w);  data_vs_prior = mean(w / invtau2) = osc * fudge * tau2 * mean(w)
    np.testing.assert_allclose(sigma2[reached], n[reached] / (osc * S[reached]), rtol=1e-13)
    np.testing.assert_allclose(dvp[reached], osc * 1.5 * 4.0 * S[reached] / n[reached], rtol=1e-13)
    np.testing.assert_array_equal(t2, tau2)  # not updated without FSC


def test_hermitian_fold_of_the_x0_plane():
    ori = cur = 8
    P = 2 * (2 * (cur // 2) + 1) + 1
    rng = np.random.default_rng(2)
    data, weight = _grids(P, P // 2 + 1, 3, rng)
    weight[:] = rng.uniform(1, 2, weight.shape)
    c = P // 2
    # the reconstruction starts from the folded grids: check the fold by reconstructing twice, once
    # from grids whose x = 0 plane is pre-folded by hand (and halved, since the fold sums pairs)
    d2, w2 = data.copy(), weight.copy()
    for z in range(P):
        for y in range(P):
            zm, ym = 2 * c - z, 2 * c - y
            if (z, y) == (c, c):
                continue
            d2[z, y, 0] = 0.5 * (data[z, y, 0] + np.conj(data[zm, ym, 0]))
            w2[z, y, 0] = 0.5 * (weight[z, y, 0] + weight[zm, ym, 0])
    a = OM.reconstruct(data, weight, ori, cur, max_iter_preweight=2, do_map=False)[0]
    b = OM.reconstruct(d2, w2, ori, cur, max_iter_preweight=2, do_map=False)[0]
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * np.abs(a).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_round_trip_recovers_the_map(dim):
    ori = 32 if dim == 3 else 64
    cur = ori
    rng = np.random.default_rng(4)
    ax = np.arange(ori) - ori // 2
    g = np.meshgrid(*([ax] * dim), indexing="ij")
    vol = np.zeros([ori] * dim)
    for _ in range(6):
        c = rng.uniform(-ori / 6, ori / 6, dim)
        s = rng.uniform(2.0, 3.5)
        vol += np.exp(-sum((gi - ci) ** 2 for gi, ci in zip(g, c)) / (2 * s * s))
    grid, _, _ = OM.fourier_map(vol, cur, do_gridding=False)
    weight = (np.abs(grid) > 0).astype(float)
    out = OM.reconstruct(grid, weight, ori, cur, max_iter_preweight=10, do_map=False)[0]
    # reconstruct ends with the trilinear gridding correction: undo it for the comparison
    out_raw = out / OM.gridding_correct(np.ones_like(vol))
    mask = np.sqrt(sum(gi.astype(float) ** 2 for gi in g)) < ori / 2 - 4
    err = np.linalg.norm((out_raw - vol)[mask]) / np.linalg.norm(vol[mask])
    assert err < 0.01, err  # 0.1 % here; a one-voxel shift of the map gives ~20 %
