"""M-step model parameters (maximizationOtherParameters, src/map3d_optimizer.cpp:2620-2798 and
src/map2d_optimizer.cpp:1985-2045): the C-ABI host function against the FP64 oracle, and the oracle
pinned by hand-counted shells and closed-form identities."""
import numpy as np
import pytest

import oracle.mstep as OM
import rome_amd


def test_npix_per_shell_by_hand():
    # ori 4: half-plane rows ip = 0, 1, 2, -1; columns 0..2.  Shell 0: (0,0).  Shell 1: (0,1) (1,0) (1,1)
    # (-1,1).  Shell 2: (0,2) (1,2) (2,0) (2,1) (-1,2).  (2,2) rounds to 3 (dropped), (-1,0) is the
    # duplicated x = 0 column.
    assert list(OM.npix_per_shell(4)) == [1, 4, 5]


def test_npix_per_shell_totals():
    for ori in (8, 16, 33, 64):
        n = OM.npix_per_shell(ori)
        F = ori // 2 + 1
        # every stored half-plane pixel whose rounded radius is below F, x = 0 column counted once
        ip = np.array([i if i < F else i - ori for i in range(ori)])
        r = np.sqrt(ip[:, None] ** 2 + np.arange(F)[None, :] ** 2)
        keep = (np.floor(r + 0.5) < F) & ~((np.arange(F)[None, :] == 0) & (ip[:, None] < 0))
        assert n.sum() == keep.sum()
        assert n[0] == 1 and n[1] == 4  # the origin; (0,1) (1,0) (1,1) (-1,1)


def _sums(rng, K, G, D, F, ref_dim):
    return {"wsum_pdf_class": rng.uniform(0.5, 2.0, K), "wsum_pdf_direction": rng.uniform(0, 1, (K, D)),
            "wsum_sigma2_noise": rng.uniform(1, 10, (G, F)), "wsum_sumw_group": rng.uniform(10, 20, G),
            "wsum_prior_offx": rng.normal(size=K), "wsum_prior_offy": rng.normal(size=K),
            "wsum_scalars": np.array([3.0, -1234.5, 0.7, 1.9])}


@pytest.mark.parametrize("ref_dim", [2, 3])
def test_capi_matches_oracle(ref_dim):
    rng = np.random.default_rng(ref_dim)
    K, G, ori = 3, 2, 32
    D = 1 if ref_dim == 2 else 48
    s = _sums(rng, K, G, D, ori // 2 + 1, ref_dim)
    px, py = rng.normal(size=K), rng.normal(size=K)
    for norm in (0, 1):
        got = rome_amd.maximization_other(ref_dim, ori, s, D, norm, 1.25, px, py)
        ref = OM.maximization_other(ref_dim, ori, s, D, norm, 1.25, px, py)
        for k in ref:
            np.testing.assert_allclose(got[k], np.reshape(ref[k], got[k].shape), rtol=1e-15, atol=0, err_msg=k)


def test_oracle_closed_forms():
    rng = np.random.default_rng(5)
    K, G, ori, D = 4, 3, 24, 12
    F = ori // 2 + 1
    s = _sums(rng, K, G, D, F, 3)
    sig_true = rng.uniform(0.1, 5.0, (G, F))
    s["wsum_sigma2_noise"] = 2.0 * s["wsum_sumw_group"][:, None] * OM.npix_per_shell(ori)[None, :] * sig_true
    m = OM.maximization_other(3, ori, s, D, 1, 1.0, np.zeros(K), np.zeros(K))
    np.testing.assert_allclose(m["sigma2_noise"], sig_true, rtol=1e-14)
    assert abs(m["pdf_class"].sum() - 1.0) < 1e-15
    np.testing.assert_allclose(m["pdf_direction"].sum(), s["wsum_pdf_direction"].sum() / s["wsum_pdf_class"].sum())
    sw = s["wsum_pdf_class"].sum()
    np.testing.assert_allclose(m["scalars"], [3.0 / (2 * sw), 1.9 / sw, -1234.5, 0.7 / sw], rtol=1e-15)
    # 2D priors are the class-weighted mean offsets; an empty class gets 0
    s["wsum_pdf_class"][1] = 0.0
    m2 = OM.maximization_other(2, ori, {**s, "wsum_pdf_direction": s["wsum_pdf_direction"][:, :1]}, 1, 0, 1.0,
                               np.ones(K), np.ones(K))
    assert m2["prior_offx"][1] == 0.0 and m2["prior_offy"][1] == 0.0
    np.testing.assert_allclose(m2["prior_offx"][0], s["wsum_prior_offx"][0] / s["wsum_pdf_class"][0])
    assert m2["scalars"][1] == 1.0  # no norm correction: the incoming average is kept


def test_capi_rejects_empty_weights():
    rng = np.random.default_rng(1)
    s = _sums(rng, 2, 1, 1, 9, 2)
    s["wsum_pdf_class"][:] = 0.0
    with pytest.raises(rome_amd.RomeError):
        rome_amd.maximization_other(2, 16, s, 1, 0, 1.0, np.zeros(2), np.zeros(2))
