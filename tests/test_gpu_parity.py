"""GPU parity: the HIP E-step (through the C-ABI) against the FP64 oracle on the same seeded inputs.

Tolerances of the plain FP32 arm: diff2 within 1e-5 relative (coarse and fine); identical max-posterior
class and pose except near-ties (the two candidates' weights within 1e-3 relative); PMAX and offsets
within 1e-4; weighted sums within 1e-4 relative; backprojector data and weight grids within 1e-5
relative L2; NR_SIGN identical for the single pass (exact recompute) and, for the FP32 coarse pass of two
passes, significant sets that differ only inside the band its measured diff2 error allows.  Zero priors
(directions, a class) get no weight.  Cases cover 3D two-pass, 2D single-pass (MAP2D default), 2D two-pass, current_size <
ori_size (highres_Xi2 and the sigma2 tail), several groups, old offsets and offset priors, CTF off, and
the full MAP3D 128^2 / K=4 / 4608 x 81 search on a few particles.
"""
import numpy as np
import pytest

import oracle as O
from rome_amd import synth

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))


def _case(name):
    if name == "3d":
        return synth.make_problem(is3d=True, ori=32, current=32, coarse=16, K=2, order=1, n=8, snr=0.5, seed=3)
    if name == "3d_window_groups_offsets":
        prob, batch, truth = synth.make_problem(is3d=True, ori=40, current=32, coarse=16, K=3, order=1, n=6,
                                                snr=0.4, seed=4, nr_groups=3, offset_range=2.0)
        rng = np.random.default_rng(9)
        batch["old_offx"] = rng.integers(-2, 3, batch["n"]).astype(float)
        batch["old_offy"] = rng.integers(-2, 3, batch["n"]).astype(float)
        batch["norm"] = rng.uniform(0.8, 1.2, batch["n"])
        prob["prior_offx"] = np.array([0.5, -0.25, 0.0])
        prob["prior_offy"] = np.array([0.0, 0.75, -0.5])
        prob["pdf_class"] = np.array([0.5, 0.3, 0.2])
        return prob, batch, truth
    if name == "3d_noctf":
        return synth.make_problem(is3d=True, ori=32, current=32, coarse=16, K=2, order=1, n=5, snr=0.5, seed=8,
                                  do_ctf=False)
    if name == "2d_single_pass":
        return synth.make_problem(is3d=False, ori=64, current=64, coarse=64, K=4, psi_step=10.0, oversampling=0,
                                  offset_range=3.0, n=10, snr=0.2, seed=12)
    if name == "2d_two_pass":
        return synth.make_problem(is3d=False, ori=48, current=48, coarse=24, K=3, psi_step=15.0, oversampling=1,
                                  offset_range=2.0, n=6, snr=0.3, seed=13, nr_groups=2)
    if name == "2d_config0":  # the CPU-runnable MAP2D workload: 64^2, K=10, 36 psi x 29 translations, one pass
        return synth.make_problem(is3d=False, ori=64, current=64, coarse=64, K=10, psi_step=10.0, oversampling=0,
                                  offset_range=3.0, n=12, snr=0.05, seed=41, angpix=2.0)
    if name == "3d_full_map3d":
        return synth.make_problem(is3d=True, ori=128, current=128, coarse=24, K=4, order=2, offset_range=5.0,
                                  oversampling=1, n=32, snr=0.05, seed=14, nblobs=200)
    if name == "3d_zero_prior":  # directions and a class without prior (iteration >= 2 after an M-step)
        prob, batch, truth = synth.make_problem(is3d=True, ori=32, current=32, coarse=16, K=3, order=1, n=8,
                                                snr=0.5, seed=31)
        zero_priors(prob)
        return prob, batch, truth
    if name == "2d_zero_prior":
        prob, batch, truth = synth.make_problem(is3d=False, ori=64, current=64, coarse=64, K=4, psi_step=10.0,
                                                oversampling=0, offset_range=3.0, n=10, snr=0.2, seed=32)
        zero_priors(prob)
        return prob, batch, truth
    if name == "3d_paired_k1":  # 4608 rotations, one class: the paired cross term over 18 row tiles (not a
        # multiple of the 8 XCDs: the XCD-grouped tile order)
        return synth.make_problem(is3d=True, ori=64, current=64, coarse=24, K=1, order=2, offset_range=3.0,
                                  oversampling=1, n=4, snr=1.0, seed=21, nblobs=120)
    raise KeyError(name)


def zero_priors(prob):
    """The prior state of iteration >= 2: M-step pdf_direction = wsum / sum leaves directions without weight
    at 0 and a class can empty (src/map3d_optimizer.cpp:2696-2705); the E-step then skips them
    (src/map3d_optimizer.cpp:1133, 1205).  Zeroes a block of directions of class 0, every other direction of
    class 1 and the whole last class (pdf_class = 0)."""
    K, D = prob["nr_classes"], prob["nr_dir"]
    pd = np.array(prob["pdf_direction"], dtype=float).reshape(K, D)
    pd[0, : max(1, D // 3)] = 0.0
    if K > 2:
        pd[1, ::2] = 0.0
    pc = np.array(prob["pdf_class"], dtype=float)
    if K > 1:
        pc[K - 1] = 0.0
    prob["pdf_direction"] = pd
    prob["pdf_class"] = pc / pc.sum()


CASES = ["3d", "3d_window_groups_offsets", "3d_noctf", "2d_single_pass", "2d_two_pass", "2d_config0", "3d_full_map3d",
         "3d_zero_prior", "2d_zero_prior"]


@pytest.fixture(scope="module", params=CASES)
def case(request):
    prob, batch, truth = _case(request.param)
    dense = request.param != "3d_full_map3d"  # dense fine dumps of the full search would take GBs
    ref = O.estep(prob, batch, dump_coarse=True, dump_fine=dense, dump_fine_weight=dense)
    print(f"[parity case {request.param}] n={batch['n']} K={prob['nr_classes']} D={prob['nr_dir']} "
          f"PSI={prob['nr_psi']} T={prob['nr_trans']}", flush=True)
    return request.param, prob, batch, ref


def _run(engine, prob, batch, precision="fp32"):
    synth.load_engine(engine, prob, precision=precision)
    r = engine.expectation(batch)
    return r, engine.get_sums()


def _run_both(engine, prob, batch, precision="fp32"):
    """The batch with the whole coarse pose term in HBM (ROME_COARSE_FULL: its coarse diff2 is readable),
    then through the product path, the coarse list (the default); the list must give the same E-step:
    it holds every entry whose weight reaches the sums."""
    n = batch["n"]
    engine.set_coarse_mode("full")
    try:
        rf, sf = _run(engine, prob, batch, precision)
        d = engine.get_coarse_diff2(n)
        assert engine.coarse_stats()["mode"] == 1
    finally:
        engine.set_coarse_mode("list")
    r, s = _run(engine, prob, batch, precision)
    st = engine.coarse_stats()
    if prob["adaptive_oversampling"]:
        assert st["mode"] == 0 and 0 < st["list_max"], st
    else:  # the single pass takes its diff2 from the full pose term whatever the mode asked for
        assert st["mode"] == 1, st
    _check_same(r, s, rf, sf)
    return r, s, d


def _check_same(r, s, rf, sf):
    """Coarse list vs full pose term: the same E values and the same window, so the same poses and
    significant sets; the weight sums differ only in FP64 summation order."""
    assert np.array_equal(r["argmax"], rf["argmax"])
    assert np.array_equal(r["nr_sign"], rf["nr_sign"])
    np.testing.assert_allclose(r["min_diff2_coarse"], rf["min_diff2_coarse"], rtol=0, atol=0)
    np.testing.assert_allclose(r["meta"], rf["meta"], rtol=1e-10, atol=1e-12)
    for k in sf:
        np.testing.assert_allclose(s[k], sf[k], rtol=1e-10, atol=1e-12, err_msg=k)


def test_projector_matches_oracle(engine):
    prob = _case("3d")[0]
    synth.load_engine(engine, prob)
    rng = np.random.default_rng(0)
    A = np.stack([synth.euler_matrix(*rng.uniform([-180, 0, 0], [180, 180, 360])) for _ in range(6)])
    re, im = engine.project(1, 32, A)
    for i in range(6):
        ore, oim = O.project3d(prob["grids"][1], prob["pad"], 16, A[i], 32)
        scale = np.max(np.abs(ore)) + np.max(np.abs(oim))
        assert np.max(np.abs(re[i] - ore)) <= 1e-5 * scale
        assert np.max(np.abs(im[i] - oim)) <= 1e-5 * scale


def test_estep_parity(engine, case):
    name, prob, batch, ref = case
    r, s, d = _run_both(engine, prob, batch)
    n = batch["n"]
    # coarse pass
    o = ref["coarse_diff2"]
    assert np.array_equal(np.isnan(d), np.isnan(o))
    m = ~np.isnan(o)
    assert _rel(d[m], o[m]) < 1e-5
    assert _rel(r["min_diff2_coarse"], ref["min_diff2_coarse"]) < 1e-5
    if prob["adaptive_oversampling"]:
        _check_significant_set(engine, prob, batch, ref, d, r)
    else:  # the single pass recomputes its marked samples exactly: the reference's own count
        assert np.array_equal(r["nr_sign"], ref["nr_sign"]), (r["nr_sign"], ref["nr_sign"])
    _check_zero_prior(engine, prob, r)
    if "fine_diff2" not in ref:  # full-size case: per-image results, sums and grids only
        check_argmax_near_ties(engine, prob, batch, r, ref, name)
        np.testing.assert_allclose(r["min_diff2"], ref["min_diff2"], rtol=1e-5)
        ref["fine_weight"] = None
    else:
        _check_lists(engine, r, ref, batch["n"], name)
    _check_sums(r, s, ref, prob, engine, name)


def _ln_prior(prob, batch):
    """ln(pdf_direction * pdf_offset) per (image, class, dir, psi, t) as the reference forms it
    (src/map3d_optimizer.cpp:1704-1726, pdf_offset of src/ml_model.h:231-242 on the coarse translations)."""
    K, D, PSI = prob["nr_classes"], prob["nr_dir"], prob["nr_psi"]
    pd = np.asarray(prob["pdf_direction"], float).reshape(K, D)
    pc = np.asarray(prob["pdf_class"], float)
    tx, ty = np.asarray(prob["trans_x"]), np.asarray(prob["trans_y"])
    ox = tx[None, None, :] + np.asarray(batch["old_offx"])[:, None, None] - np.asarray(prob["prior_offx"])[None, :, None]
    oy = ty[None, None, :] + np.asarray(batch["old_offy"])[:, None, None] - np.asarray(prob["prior_offy"])[None, :, None]
    s2 = prob["sigma2_offset"]
    with np.errstate(divide="ignore"):
        lpo = -(ox * ox + oy * oy) / (2 * s2) - np.log(2 * np.pi * s2)  # [n][K][T]
        lpd = np.where((pd > 0) & (pc[:, None] > 0), np.log(np.where(pd > 0, pd, 1.0)), -np.inf)  # [K][D]
    return lpd[None, :, :, None, None] + lpo[:, :, None, None, :] + np.zeros((1, 1, 1, PSI, 1))


def _check_significant_set(engine, prob, batch, ref, d, r):
    """Coarse significant samples of the FP32 pass against the oracle's.  With eps the largest coarse diff2
    error of the image (measured), each GPU weight is the oracle's times exp(delta), |delta| <= 2 eps, once
    both are anchored at their minima, and the sums agree to the same factor.  Both significant sets are
    top-mass prefixes (cumulative mass from the largest weight first exceeding 0.999 of the sum), so they
    can differ only in samples (a) whose oracle ln-weight is within 4 eps of the oracle's threshold, or
    (b) whose oracle cumulative mass lies within 4 eps of the sum (plus the sample's own weight) of the
    crossing 0.999 * sum.  The test holds the symmetric difference of the two sets to that band, which also
    bounds the NR_SIGN difference by the band's size."""
    n = batch["n"]
    K, D, PSI, T = prob["nr_classes"], prob["nr_dir"], prob["nr_psi"], prob["nr_trans"]
    keys, _ = engine.get_fine_list()
    ne = (8 if prob["is_3d"] else 2) * 4
    gsig = np.zeros(n * K * D * PSI * T, bool)
    gsig[np.unique(keys // ne)] = True
    gsig = gsig.reshape(n, -1)
    osig = ref["coarse_sig"].astype(bool).reshape(n, -1)
    o = ref["coarse_diff2"].reshape(n, -1)
    lnw = (_ln_prior(prob, batch) - (ref["coarse_diff2"] - ref["min_diff2_coarse"][:, None, None, None, None]))
    lnw = lnw.reshape(n, -1)
    f = prob["adaptive_fraction"]
    for i in range(n):
        m = ~np.isnan(o[i]) & np.isfinite(lnw[i])
        eps = float(np.max(np.abs(d.reshape(n, -1)[i][m] - o[i][m])))
        w = np.where(m, np.exp(np.where(m, lnw[i], -np.inf)), 0.0)
        order = np.argsort(-w, kind="stable")
        cum = np.empty_like(w)
        cum[order] = np.cumsum(w[order])
        tot = cum[order[-1]]
        lthr = np.log(ref["sig_weight_coarse"][i])
        band = m & ((np.abs(lnw[i] - lthr) <= 4 * eps + 1e-9) | (np.abs(cum - f * tot) <= 4 * eps * tot + w))
        diff = gsig[i] ^ osig[i]
        assert not np.any(diff & ~band), (i, int(diff.sum()), int((diff & ~band).sum()), eps)
        assert abs(int(r["nr_sign"][i]) - int(ref["nr_sign"][i])) <= int(band.sum()), i


def check_argmax_near_ties(engine, prob, batch, r, ref, name, tol=1e-3):
    """Max-posterior poses without the oracle's dense dumps: where the GPU's pose differs from the oracle's,
    both must be in the GPU's last-pass list with weights (diff2 minus ln prior) within tol relative of each
    other: a near-tie."""
    n = batch["n"]
    K, D, PSI, T = prob["nr_classes"], prob["nr_dir"], prob["nr_psi"], prob["nr_trans"]
    nor, nt = ((8 if prob["is_3d"] else 2), 4) if prob["adaptive_oversampling"] else (1, 1)
    bad = [i for i in range(n) if not np.array_equal(r["argmax"][i], ref["argmax"][i])]
    if not bad:
        return
    keys, d2 = engine.get_fine_list()
    pos = {int(k): j for j, k in enumerate(keys)}
    lp = _ln_prior(prob, batch)

    def u_of(i, a):
        k, dd, ps, o_, t, ot = (int(v) for v in a)
        key = ((((i * K + k) * D + dd) * PSI + ps) * T + t) * nor * nt + o_ * nt + ot
        assert key in pos, f"{name} image {i}: pose {a} not in the GPU list"
        return d2[pos[key]] - lp[i, k, dd, ps, t]

    for i in bad:
        du = abs(u_of(i, r["argmax"][i]) - u_of(i, ref["argmax"][i]))
        assert du <= -np.log1p(-tol), f"{name} image {i}: argmax {r['argmax'][i]} vs {ref['argmax'][i]} ({du:.2e} nat)"


def _check_zero_prior(engine, prob, r):
    """Rotations without prior get no weight: the max-posterior class and direction always have one, and
    the fine list holds no sample of them."""
    K, D, PSI, T = prob["nr_classes"], prob["nr_dir"], prob["nr_psi"], prob["nr_trans"]
    pd = np.asarray(prob["pdf_direction"], float).reshape(K, D)
    pc = np.asarray(prob["pdf_class"], float)
    ok = (pd > 0) & (pc[:, None] > 0)
    if ok.all():
        return
    assert np.all(ok[r["argmax"][:, 0], r["argmax"][:, 1]])
    keys, _ = engine.get_fine_list()
    ne = ((8 if prob["is_3d"] else 2) * 4) if prob["adaptive_oversampling"] else 1
    c = keys // ne // T // PSI
    kd = c % (K * D)
    assert np.all(ok.reshape(-1)[kd])


def _check_lists(engine, r, ref, n, name):
    # last pass list
    keys, d2 = engine.get_fine_list()
    fo = ref["fine_diff2"].reshape(-1)
    ov = fo[keys]
    assert not np.any(np.isnan(ov)), "GPU holds samples the oracle did not compute"
    assert _rel(d2, ov) < 1e-5
    np.testing.assert_allclose(r["min_diff2"], ref["min_diff2"], rtol=1e-5)
    np.testing.assert_allclose(r["sum_weight"], ref["sum_weight"], rtol=1e-3)
    # max-posterior class and pose (near-ties excepted)
    fw = ref["fine_weight"].reshape(n, -1)
    for i in range(n):
        if np.array_equal(r["argmax"][i], ref["argmax"][i]):
            continue
        shp = ref["fine_weight"].shape[1:]
        k, dd, ps, o_, t, ot = r["argmax"][i]
        wg = ref["fine_weight"][i][k, dd, ps, t, o_, ot]
        wo = np.nanmax(fw[i])
        assert wg >= wo * (1 - 1e-3), f"{name} image {i}: argmax {r['argmax'][i]} vs {ref['argmax'][i]}"


def _check_sums(r, s, ref, prob, engine, name):
    np.testing.assert_allclose(r["meta"][:, 0], ref["meta"][:, 0], rtol=1e-3)
    same = np.all(r["argmax"] == ref["argmax"], axis=1)
    np.testing.assert_allclose(r["meta"][same, 1:7], ref["meta"][same, 1:7], rtol=1e-9, atol=1e-9)
    # sums
    np.testing.assert_allclose(s["wsum_sigma2_noise"], ref["wsum_sigma2_noise"], rtol=1e-4, atol=1e-12)
    np.testing.assert_allclose(s["wsum_pdf_class"], ref["wsum_pdf_class"], rtol=1e-4, atol=1e-12)
    np.testing.assert_allclose(s["wsum_pdf_direction"], ref["wsum_pdf_direction"], rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(s["wsum_sumw_group"], ref["wsum_sumw_group"], rtol=1e-4, atol=1e-12)
    # signed sums (weights x offsets of both signs): a class whose offsets cancel is held to 1e-4 of the
    # largest class's sum rather than of its own
    for key in ("wsum_prior_offx", "wsum_prior_offy"):
        np.testing.assert_allclose(s[key], ref[key], rtol=1e-4, atol=1e-9 + 1e-4 * np.max(np.abs(ref[key])), err_msg=key)
    np.testing.assert_allclose(s["wsum_scalars"], ref["wsum_scalars"], rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(r["meta"][:, 7:], ref["meta"][:, 7:], rtol=1e-4)
    np.testing.assert_allclose(r["wsum_norm_correction"], ref["wsum_norm_correction"], rtol=1e-4)
    # backprojection
    for k in range(prob["nr_classes"]):
        dg, wg = engine.get_backprojector(k)
        od, ow = ref["bp_data"][k], ref["bp_weight"][k]
        if np.linalg.norm(ow) == 0:
            assert np.linalg.norm(wg) == 0
            continue
        assert np.linalg.norm(dg - od) <= 1e-5 * np.linalg.norm(od), name
        assert np.linalg.norm(wg - ow) <= 1e-5 * np.linalg.norm(ow), name


def test_single_pass_in_a_fresh_context():
    """A new context asked for the coarse list runs the single pass (2D, no oversampling) on the full
    pose term, with the same results as the session's engine."""
    import rome_amd
    prob, batch, _ = _case("2d_single_pass")
    e = rome_amd.Engine(0)
    try:
        r, s = _run(e, prob, batch)
        assert e.coarse_stats()["mode"] == 1
    finally:
        e.close()
    ref = O.estep(prob, batch)
    np.testing.assert_allclose(r["meta"][:, 0], ref["meta"][:, 0], rtol=1e-3)
    np.testing.assert_allclose(s["wsum_sigma2_noise"], ref["wsum_sigma2_noise"], rtol=1e-4, atol=1e-12)


def test_list_overflow_reruns_full(engine):
    """A coarse list past its cap sends the pool through the full pose term: same E-step."""
    prob, batch, _ = _case("3d_window_groups_offsets")
    engine.set_coarse_mode("full")
    try:
        rf, sf = _run(engine, prob, batch)
    finally:
        engine.set_coarse_mode("list")
    engine.set_coarse_mode("list", 4)
    try:
        r, s = _run(engine, prob, batch)
        st = engine.coarse_stats()
    finally:
        engine.set_coarse_mode("list")
    assert st["mode"] == 2 and st["list_max"] > 4, st
    assert np.array_equal(r["argmax"], rf["argmax"]) and np.array_equal(r["nr_sign"], rf["nr_sign"])
    np.testing.assert_array_equal(r["meta"], rf["meta"])
    for k in sf:  # FP64 atomic accumulation: order-dependent last bits only
        np.testing.assert_allclose(s[k], sf[k], rtol=1e-12, atol=1e-15, err_msg=k)


@pytest.mark.parametrize("name", ["3d_window_groups_offsets", "2d_single_pass"])
def test_significance_keys_and_mask_agree(engine, name):
    """The fine pass built from the per-image keys of the significant samples, from the dense mask
    (cap 0), and from keys that overflow their cap (the pool then takes the mask): the same tasks in
    the same order, so the same E-step."""
    prob, batch, _ = _case(name)
    out = {}
    try:
        for cap in (-1, 0, 1):
            engine.set_significance_cap(cap)
            r, s = _run(engine, prob, batch)
            keys, d2 = engine.get_fine_list()
            out[cap] = (r, s, keys, d2, engine.significance_stats())
    finally:
        engine.set_significance_cap(-1)
    assert out[-1][4]["sig_form"] == 0 and out[-1][4]["sig_max"] > 1, out[-1][4]
    assert out[0][4]["sig_form"] == 1, out[0][4]
    assert out[1][4]["sig_form"] == 2, out[1][4]
    r0, s0, k0, d0, _ = out[-1]
    for cap in (0, 1):
        r, s, k, d, _ = out[cap]
        assert np.array_equal(r["argmax"], r0["argmax"]) and np.array_equal(r["nr_sign"], r0["nr_sign"])
        np.testing.assert_array_equal(r["meta"], r0["meta"])
        np.testing.assert_array_equal(k, k0)
        np.testing.assert_array_equal(d, d0)
        for key in s0:  # FP64 atomic accumulation: order-dependent last bits only
            np.testing.assert_allclose(s[key], s0[key], rtol=1e-12, atol=1e-15, err_msg=key)


def _check_model(got, want):
    for k in want:
        w = np.reshape(want[k], got[k].shape)
        # prior offsets are ratios of signed sums: held to 1e-4 of the largest class's value
        atol = 1e-12 + (1e-4 * np.max(np.abs(w)) if "prior" in k else 0.0)
        np.testing.assert_allclose(got[k], w, rtol=1e-4, atol=atol, err_msg=k)


def test_update_model_matches_oracle(engine, case):
    """M-step model parameters from the engine's own sums vs the oracle's update of the oracle's sums."""
    import oracle.mstep as OM
    name, prob, batch, ref = case
    _run(engine, prob, batch)
    got = engine.update_model(install=False)
    ref_sums = {k: ref[k] for k in ("wsum_pdf_class", "wsum_pdf_direction", "wsum_sigma2_noise", "wsum_sumw_group",
                                    "wsum_prior_offx", "wsum_prior_offy", "wsum_scalars")}
    want = OM.maximization_other(3 if prob["is_3d"] else 2, prob["ori_size"], ref_sums, prob["nr_dir"],
                                 prob["do_norm_correction"], prob["avg_norm_correction"], prob["prior_offx"],
                                 prob["prior_offy"])
    _check_model(got, want)


def test_config0_map2d_full_iteration(engine):
    """One E+M iteration of the MAP2D workload (64^2, K=10, 36 psi x 29 translations): the engine's E-step
    then its reconstruction of every class from its own backprojector, against the oracle's E-step and the
    oracle's reconstruction from the oracle's grids (maps within 1e-5 relative L2), identical max-posterior
    class and pose except near-ties, and the model update from the engine's sums against the oracle's."""
    import oracle.mstep as OM
    prob, batch, _ = _case("2d_config0")
    ref = O.estep(prob, batch, dump_fine=True, dump_fine_weight=True)
    r, s = _run(engine, prob, batch)
    _check_lists(engine, r, ref, batch["n"], "config0 iteration")
    ori, cur = prob["ori_size"], prob["current_size"]
    tau2 = np.full(ori // 2 + 1, 0.5)
    shape = (prob["pad"], prob["pad"] // 2 + 1)
    for k in range(prob["nr_classes"]):
        if np.linalg.norm(ref["bp_weight"][k]) == 0:
            continue
        vol = engine.reconstruct(k, tau2)[0]
        od = np.asarray(ref["bp_data"][k])  # interleaved re, im
        od = (od[0::2] + 1j * od[1::2]).reshape(shape)
        ow = np.asarray(ref["bp_weight"][k]).reshape(shape)
        want = OM.reconstruct(od, ow, ori, cur, tau2=tau2)[0]
        assert np.linalg.norm(vol - want) <= 1e-5 * np.linalg.norm(want), k
    got = engine.update_model(install=False)
    ref_sums = {k: ref[k] for k in ("wsum_pdf_class", "wsum_pdf_direction", "wsum_sigma2_noise", "wsum_sumw_group",
                                    "wsum_prior_offx", "wsum_prior_offy", "wsum_scalars")}
    want = OM.maximization_other(2, ori, ref_sums, prob["nr_dir"], prob["do_norm_correction"],
                                 prob["avg_norm_correction"], prob["prior_offx"], prob["prior_offy"])
    _check_model(got, want)


def test_config3_shape_192_hp3(engine):
    """The batched fine-pass workload's shape on two particles: 192^2 (14615 fine pixels), HEALPix order 3
    x 48 psi (36864 rotations per class), 49 translations, oversampling 1 (fine: order 4, 8 rotations x 4
    translations per significant sample), against the oracle: poses, diff2, sums and grids."""
    prob, batch, _ = synth.make_problem(is3d=True, ori=192, current=192, coarse=32, K=2, order=3, offset_range=4.0,
                                        oversampling=1, n=2, snr=0.05, seed=51, nblobs=100)
    print(f"[config3 shape] D={prob['nr_dir']} PSI={prob['nr_psi']} T={prob['nr_trans']} pad={prob['pad']}",
          flush=True)
    ref = O.estep(prob, batch)
    r, s = _run(engine, prob, batch)
    assert np.array_equal(r["argmax"], ref["argmax"])
    np.testing.assert_allclose(r["min_diff2"], ref["min_diff2"], rtol=1e-5)
    np.testing.assert_allclose(r["min_diff2_coarse"], ref["min_diff2_coarse"], rtol=1e-5)
    _check_sums(r, s, ref, prob, engine, "config3 shape")


def test_class_range_single_class_pool(engine):
    """The seeding iteration's pools run one class each (exp_iclass_min = exp_iclass_max,
    src/map3d_optimizer.cpp:823-840): rome_set_class_range(k, k) against the oracle run with every other
    class left out (the oracle skips classes without prior, as the reference's loop skips classes outside
    the range); then the full range again gives the all-class E-step."""
    prob, batch, _ = _case("3d_window_groups_offsets")
    K = prob["nr_classes"]
    print(f"[class range] K={K}", flush=True)
    for k in range(K):
        only = dict(prob, pdf_class=np.where(np.arange(K) == k, prob["pdf_class"], 0.0))
        ref = O.estep(only, batch)
        synth.load_engine(engine, prob)
        engine.set_class_range(k, k)
        try:
            r = engine.expectation(batch)
            s = engine.get_sums()
        finally:
            engine.set_class_range()
        assert np.all(r["argmax"][:, 0] == k)
        assert np.array_equal(r["argmax"], ref["argmax"]), k
        np.testing.assert_allclose(r["min_diff2"], ref["min_diff2"], rtol=1e-5)
        _check_sums(r, s, ref, only, engine, f"class range {k}")
    synth.load_engine(engine, prob)
    r = engine.expectation(batch)
    np.testing.assert_allclose(r["min_diff2"], O.estep(prob, batch)["min_diff2"], rtol=1e-5)
