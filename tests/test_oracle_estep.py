"""Pins the oracle's E-step composition (src/map3d_optimizer.cpp:883-2057, src/map2d_optimizer.cpp:802-1873).

- diff2 of sampled entries recomputed from the separately pinned primitives (projector, tabulated
  shift, window, Minvsigma2) straight from the definition 1/2 (sum Minv |CTF*P - shift(X)|^2 + Xi2);
- weights: the reference's formula recomputed from the diff2 dump (pdf_offset on coarse
  translations, 700 cut), normalisation, PMAX, argmax in loop order;
- significance: threshold and count by an independent sort / cumulative sum;
- accumulated sums: sigma2 spectrum vs norm-correction sums, pdf sums vs captured mass,
  backprojector weight vs sum of w * Minv * CTF^2 over the backprojected pixels;
- error behaviour: an image with no positive weight raises like the reference's ERROR_REPORT.
"""
import numpy as np
import pytest

import oracle as O
from rome_amd import synth


@pytest.fixture(scope="module")
def p3d():
    prob, batch, truth = synth.make_problem(is3d=True, ori=24, current=24, coarse=12, K=2, order=1, n=3, snr=0.3,
                                            offset_range=1.5, seed=5)
    r = O.estep(prob, batch, dump_coarse=True, dump_fine=True, dump_fine_weight=True)
    return prob, batch, r


@pytest.fixture(scope="module")
def p2d():
    prob, batch, truth = synth.make_problem(is3d=False, ori=32, current=32, coarse=32, K=3, psi_step=20.0,
                                            oversampling=0, n=4, snr=0.3, offset_range=2, seed=6, nr_groups=2)
    r = O.estep(prob, batch, dump_coarse=True, dump_fine=True, dump_fine_weight=True)
    return prob, batch, r


def _diff2_direct(prob, batch, i, k, d, s, t, coarse, orot=0, ot=0):
    size = prob["coarse_size"] if coarse else prob["current_size"]
    cur = prob["current_size"]
    is3d = prob["is_3d"]
    F2c = cur * (cur // 2 + 1)
    fr, fi = O.window_fourier(batch["fmask_re"][i], batch["fmask_im"][i], cur, size)
    ctf = O.window_fourier(batch["fctf"][i], np.zeros(F2c), cur, size)[0]
    g = batch["group"][i]
    ires = O.mresol(size)
    minv = np.where(ires > 0, 1.0 / (prob["sigma2_fudge"] * prob["sigma2_noise"][g][np.maximum(ires, 0)]), 0.0)
    if coarse:
        sx, sy = prob["shift_coarse_x"][t], prob["shift_coarse_y"][t]
    else:
        sx, sy = prob["shift_fine_x"][t, ot], prob["shift_fine_y"][t, ot]
    xr, xi = O.shift_image(fr, fi, size, sx, sy, prob["ori_size"])
    if is3d:
        pose = prob["pose_coarse"][d, s] if coarse else prob["pose_fine"][d, s, orot]
        A = O.euler_matrix(*pose)
        rr, ri = O.project3d(prob["grids"][k], prob["pad"], cur // 2, A, size)
    else:
        psi = prob["pose_coarse"][s] if coarse else prob["pose_fine"][s, orot]
        A = O.euler_matrix(0, 0, psi)
        rr, ri = O.rotate2d(prob["grids"][k], prob["pad"], cur // 2, A, size)
    rr, ri = rr * ctf, ri * ctf
    return 0.5 * (np.sum(((rr - xr) ** 2 + (ri - xi) ** 2) * minv) + batch["highres_Xi2"][i])


def test_diff2_from_definition(p3d, p2d):
    rng = np.random.default_rng(0)
    for prob, batch, r in (p3d, p2d):
        cd = r["coarse_diff2"]
        for _ in range(12):
            i, k, d, s, t = [int(rng.integers(n)) for n in cd.shape]
            want = _diff2_direct(prob, batch, i, k, d, s, t, True)
            assert abs(cd[i, k, d, s, t] - want) <= 1e-10 * abs(want)
    prob, batch, r = p3d
    fd = r["fine_diff2"]
    idx = np.argwhere(~np.isnan(fd))
    for q in rng.choice(len(idx), 8, replace=False):
        i, k, d, s, t, o, ot = idx[q]
        want = _diff2_direct(prob, batch, i, k, d, s, t, False, o, ot)
        assert abs(fd[i, k, d, s, t, o, ot] - want) <= 1e-10 * abs(want)


def _weights_from_diff2(prob, batch, d2, i):
    K, D, PSI, T = d2.shape[:4]
    w = np.full(d2.shape, np.nan)
    mn = np.nanmin(d2)
    for k in range(K):
        for t in range(T):
            ox = prob["trans_x"][t] + batch["old_offx"][i] - prob["prior_offx"][k]
            oy = prob["trans_y"][t] + batch["old_offy"][i] - prob["prior_offy"][k]
            s2 = prob["sigma2_offset"]
            poff = np.exp((ox * ox + oy * oy) / (-2 * s2)) / (2 * np.pi * s2)
            for d in range(D):
                dd = d2[k, d, :, t] - mn
                w[k, d, :, t] = np.where(dd > 700, 0.0, prob["pdf_direction"][k, d] * poff * np.exp(-dd))
    return w, mn


def test_weights_and_significance(p3d, p2d):
    for prob, batch, r in (p3d, p2d):
        fd, fw = r["fine_diff2"], r["fine_weight"]
        for i in range(batch["n"]):
            w, mn = _weights_from_diff2(prob, batch, fd[i], i)
            m = ~np.isnan(fd[i])
            np.testing.assert_allclose(fw[i][m], w[m], rtol=1e-12, atol=1e-300)
            assert abs(r["min_diff2"][i] - mn) == 0
            s = np.nansum(w)
            np.testing.assert_allclose(r["sum_weight"][i], s, rtol=1e-12)
            flat = np.where(m, w, -1).reshape(-1)
            best = int(np.argmax(flat))  # first maximum in (k, dir, psi, t, over_rot, over_trans) order
            shp = fd[i].shape
            k, d, ps, t, o, ot = np.unravel_index(best, shp)
            assert list(r["argmax"][i]) == [k, d, ps, o, t, ot]
            assert abs(r["meta"][i][0] - flat[best] / s) < 1e-12
            vals = np.sort(flat[flat >= 0])[::-1]
            cum = np.cumsum(vals)
            j = int(np.argmax(cum > prob["adaptive_fraction"] * s))
            assert r["sig_weight"][i] == vals[j]
    # coarse pass significance count (NR_SIGN)
    prob, batch, r = p3d
    for i in range(batch["n"]):
        w, _ = _weights_from_diff2(prob, batch, r["coarse_diff2"][i], i)
        vals = np.sort(w[~np.isnan(w)].reshape(-1))[::-1]
        cum = np.cumsum(vals)
        j = int(np.argmax(cum > prob["adaptive_fraction"] * np.sum(vals)))
        assert r["nr_sign"][i] == j + 1
        assert r["sig_weight_coarse"][i] == vals[j]


def test_accumulated_sums_are_consistent(p3d, p2d):
    for prob, batch, r in (p3d, p2d):
        n = batch["n"]
        # storeWeightedSums adds the same power tail to both; the rest is the same sum of w |CF - XS|^2
        np.testing.assert_allclose(r["wsum_sigma2_noise"].sum(), r["wsum_norm_correction"].sum(), rtol=1e-12)
        captured = r["wsum_pdf_class"].sum()
        assert prob["adaptive_fraction"] * n <= captured <= n * (1 + 1e-12)
        np.testing.assert_allclose(r["wsum_pdf_direction"].sum(), captured, rtol=1e-12)
        np.testing.assert_allclose(r["wsum_sumw_group"].sum(), captured, rtol=1e-12)
        assert np.all(r["bp_weight"] >= -1e-12)
        assert np.isclose(r["wsum_scalars"][2], r["meta"][:, 0].sum())


def test_backprojected_weight_equals_captured_ctf2_minv(p2d):
    prob, batch, r = p2d
    cur = prob["current_size"]
    x, y = synth.halfplane_xy(cur)
    F = cur // 2 + 1
    rows = np.arange(len(x)) // F
    bp_set = (x * x + y * y <= (cur // 2) ** 2) & ~((x == 0) & (rows > cur // 2))
    ires = O.mresol(cur)
    total = 0.0
    for i in range(batch["n"]):
        g = batch["group"][i]
        minv = np.where(ires > 0, 1.0 / prob["sigma2_noise"][g][np.maximum(ires, 0)], 0.0)
        minv[0] = 1.0 / prob["sigma2_noise"][g][0]
        mc2 = minv * batch["fctf"][i] ** 2
        fw = r["fine_weight"][i]
        sig = np.nansum(np.where(fw >= r["sig_weight"][i], fw, 0.0)) / r["sum_weight"][i]
        total += sig * np.sum(np.where(bp_set & (mc2 > 0), mc2, 0.0))
    np.testing.assert_allclose(r["bp_weight"].sum(), total, rtol=1e-10)


def test_zero_prior_raises():
    prob, batch, _ = synth.make_problem(is3d=False, ori=16, current=16, coarse=16, K=1, psi_step=45.0,
                                        oversampling=0, n=1, seed=2)
    prob["pdf_class"] = np.zeros(1)
    with pytest.raises(RuntimeError, match="significant"):
        O.estep(prob, batch)
