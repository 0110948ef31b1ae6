"""GPU parity of the mixed-precision arm: the coarse cross term in BF16 or FP16 MFMA (FP32 accumulation).

The coarse pass only selects the significant samples; the fine pass that sets the poses, weights, sums
and maps stays FP32/FP64.  Bars:

* coarse diff2: each operand is rounded once to the low format after an exact power-of-two scaling, so
  |d_mixed - d_oracle| <= (2u + u^2) sum_k |a_k b_k| (+ FP32 accumulation), u = 2^-8 (BF16), 2^-11 (FP16).
  sum_k |a_k b_k| <= At + C2 is bounded by the image's largest diff2 over all poses (there the cross
  term is near zero), so the test allows 4 u max_pose diff2 per image.  It also requires an error above
  u/256 of that scale, which an FP32 cross term (error ~2^-23) never reaches: the low-precision
  kernel, not the FP32 one, produced the numbers.
* identical max-posterior class and pose, except near-ties (weights within 1e-3 relative, as in the
  FP32 parity test);
* PMAX within 1e-3, sums and backprojector grids within 5e-3 relative L2: the significant set can
  differ from the FP64 one only in its tail, which holds 1 - adaptive_fraction = 0.1 % of the weight.
"""
import numpy as np
import pytest

import oracle as O
from rome_amd import synth

from test_gpu_parity import _case, _run_both, check_argmax_near_ties, zero_priors

pytestmark = pytest.mark.gpu

U = {"bf16": 2.0 ** -8, "fp16": 2.0 ** -11}
# nats below the low-precision significance threshold within which coarse samples are recomputed exactly
MARGIN = {"bf16": 12.0, "fp16": 2.0}
CASES = ["3d", "3d_window_groups_offsets", "2d_two_pass", "3d_full_map3d", "3d_paired_k1", "3d_paired_k1_zero_prior",
         "3d_zero_prior"]
FULL_SIZE = ("3d_full_map3d", "3d_paired_k1", "3d_paired_k1_zero_prior")  # no dense fine dumps


@pytest.fixture(scope="module", params=CASES)
def case(request):
    if request.param == "3d_paired_k1_zero_prior":  # zero-prior rotations through the paired list epilogue
        prob, batch, truth = _case("3d_paired_k1")
        zero_priors(prob)
    else:
        prob, batch, truth = _case(request.param)
    dense = request.param not in FULL_SIZE
    ref = O.estep(prob, batch, dump_coarse=True, dump_fine=dense, dump_fine_weight=dense)
    return request.param, prob, batch, ref


def _l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_mixed_precision_estep(engine, case, prec):
    name, prob, batch, ref = case
    print(f"[mixed case {name} {prec}] n={batch['n']} D={prob['nr_dir']} PSI={prob['nr_psi']} "
          f"T={prob['nr_trans']}", flush=True)
    r, s, d = _run_both(engine, prob, batch, prec)
    n = batch["n"]
    u = U[prec]
    # coarse pass
    d = d.reshape(n, -1)
    o = ref["coarse_diff2"].reshape(n, -1)
    assert np.array_equal(np.isnan(d), np.isnan(o))
    worst = 0.0
    for i in range(n):
        m = ~np.isnan(o[i])
        scale = np.max(o[i][m])
        err = np.max(np.abs(d[i][m] - o[i][m]))
        assert err <= 4 * u * scale, f"{name} image {i}: coarse error {err:.3e} vs bound {4 * u * scale:.3e}"
        worst = max(worst, err / scale)
    assert worst > u / 256, f"coarse error {worst:.3e} too small for {prec}: FP32 path ran?"
    # The recompute margin has to cover the error of the samples near the weights: the threshold's and a
    # sample's own error, each at most the largest error among the samples within reach of the minimum.
    near = [np.max(np.abs(d[i] - o[i])[~np.isnan(o[i]) & (o[i] < np.nanmin(o[i]) + MARGIN[prec] + 30)])
            for i in range(n)]
    print(f"[mixed case {name} {prec}] coarse error near the weights: max {max(near):.3f} nat", flush=True)
    assert 2 * max(near) < MARGIN[prec], (name, prec, near)
    # poses
    if ref.get("fine_weight") is not None:
        fw = ref["fine_weight"]
        for i in range(n):
            if np.array_equal(r["argmax"][i], ref["argmax"][i]):
                continue
            k, dd, ps, o_, t, ot = r["argmax"][i]
            assert fw[i][k, dd, ps, t, o_, ot] >= np.nanmax(fw[i]) * (1 - 1e-3), f"{name} image {i}"
    else:
        check_argmax_near_ties(engine, prob, batch, r, ref, name)
    # PMAX = max / sum of the fine weights.  With the oracle's significant set (same NR_SIGN) the fine sums
    # agree to FP32 rounding: 1e-3.  Where the 0.999-mass cut fell one side or the other of a coarse sample
    # whose recomputed weight sits at the threshold, the fine sum gains or loses that sample's fine
    # weights, a share of order 1 - adaptive_fraction of the sum: 5e-3, on at most a tenth of the images
    # (at least one).
    same_set = r["nr_sign"] == ref["nr_sign"]
    assert np.sum(~same_set) <= max(1, n // 10), (r["nr_sign"], ref["nr_sign"])
    np.testing.assert_allclose(r["meta"][same_set, 0], ref["meta"][same_set, 0], rtol=1e-3)
    np.testing.assert_allclose(r["meta"][:, 0], ref["meta"][:, 0], rtol=5e-3)
    same = np.all(r["argmax"] == ref["argmax"], axis=1)
    np.testing.assert_allclose(r["meta"][same, 1:7], ref["meta"][same, 1:7], rtol=1e-9, atol=1e-9)
    # sums and maps
    for key in ("wsum_sigma2_noise", "wsum_pdf_class", "wsum_pdf_direction", "wsum_sumw_group", "wsum_scalars"):
        assert _l2(s[key], ref[key]) < 5e-3, key
    for k in range(prob["nr_classes"]):
        dg, wg = engine.get_backprojector(k)
        assert _l2(dg, ref["bp_data"][k]) < 5e-3, f"bp_data class {k}"
        assert _l2(wg, ref["bp_weight"][k]) < 5e-3, f"bp_weight class {k}"


def test_mixed_precision_needs_two_pass(engine):
    prob = _case("2d_single_pass")[0]
    with pytest.raises(RuntimeError, match="adaptive_oversampling"):
        synth.load_engine(engine, prob, precision="bf16")
