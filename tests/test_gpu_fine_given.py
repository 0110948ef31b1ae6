"""The fine pass of a pool whose coarse significance is given (rome_expectation_fine_device) against the FP64
oracle run with the same significance (its coarse pass skipped), in every execution mode.

Significance is clustered around each particle's true pose with per-image counts drawn lognormal, the way
the batched adaptive workload draws them.  Bars as the FP32 parity test: fine diff2 within 1e-5 relative,
max-posterior pose identical except near-ties, PMAX within 1e-3, sums within 1e-4, backprojector grids
within 1e-5 relative L2, NR_SIGN = the given counts.  Malformed significance (a key out of range, a
repeated key, an empty image) is an argument error."""
import numpy as np
import pytest

import oracle as O
from rome_amd import synth

from test_gpu_parity import _case, _check_lists, _check_sums

pytestmark = pytest.mark.gpu


def _device_batch(batch):
    import torch
    dev = torch.device("cuda:0")
    t = {}
    t["fmask"] = torch.tensor(batch["fmask_re"] + 1j * batch["fmask_im"], dtype=torch.complex64, device=dev)
    t["fnomask"] = torch.tensor(batch["fnomask_re"] + 1j * batch["fnomask_im"], dtype=torch.complex64, device=dev)
    t["fctf"] = torch.tensor(batch["fctf"], dtype=torch.float32, device=dev)
    for k in ("highres_Xi2", "power_img", "old_offx", "old_offy", "norm"):
        t[k] = torch.tensor(np.ascontiguousarray(batch[k], dtype=np.float64), device=dev)
    t["group"] = torch.tensor(np.asarray(batch["group"], np.int32), device=dev)
    return t


def _run_given(engine, t, n, keys_t, cnt_t, cap):
    return engine.expectation_fine_device(
        n, t["fmask"].data_ptr(), t["fnomask"].data_ptr(), t["fctf"].data_ptr(), t["highres_Xi2"].data_ptr(),
        t["power_img"].data_ptr(), t["old_offx"].data_ptr(), t["old_offy"].data_ptr(), t["norm"].data_ptr(),
        t["group"].data_ptr(), keys_t.data_ptr(), cnt_t.data_ptr(), cap)


@pytest.mark.parametrize("name", ["3d", "3d_window_groups_offsets", "2d_two_pass"])
def test_given_significance_matches_oracle(engine, name):
    import torch
    prob, batch, truth = _case(name)
    n = batch["n"]
    rng = np.random.default_rng(77)
    counts = synth.lognormal_counts(n, 12, 1.0, 1, 80, rng)
    keys, counts = synth.clustered_significance(prob, np.asarray(truth, float), counts)
    mask = synth.significance_mask(prob, keys, counts)
    print(f"[given significance {name}] n={n} counts={counts.tolist()}", flush=True)
    ref = O.estep(prob, dict(batch, coarse_sig_in=mask), dump_fine=True, dump_fine_weight=True)
    assert np.array_equal(ref["nr_sign"], counts)
    t = _device_batch(batch)
    cap = keys.shape[1]
    kt = torch.tensor(keys, device="cuda")
    ct = torch.tensor(counts, device="cuda")
    try:
        for mode in ("host", "device", "graph", "graph"):
            synth.load_engine(engine, prob)
            engine.set_execution(mode)
            r = _run_given(engine, t, n, kt, ct, cap)
            s = engine.get_sums()
            assert np.array_equal(r["nr_sign"], counts), mode
            assert np.all(np.isnan(r["min_diff2_coarse"])), mode
            _check_lists(engine, r, ref, n, f"{name} {mode}")
            _check_sums(r, s, ref, prob, engine, f"{name} {mode}")
    finally:
        engine.set_execution("graph")


def test_given_significance_rejects_malformed(engine):
    import torch
    prob, batch, truth = _case("3d")
    n = batch["n"]
    counts = np.full(n, 3, np.int32)
    keys, counts = synth.clustered_significance(prob, np.asarray(truth, float), counts, cap=4)
    t = _device_batch(batch)
    MT = prob["nr_classes"] * prob["nr_dir"] * prob["nr_psi"] * prob["nr_trans"]
    synth.load_engine(engine, prob)
    print("[given significance malformed]", flush=True)
    for what, kk, cc in [("key out of range", np.where(np.arange(4) == 0, MT, keys), counts),
                         ("repeated key", np.column_stack([keys[:, :2], keys[:, 1:2], keys[:, 3:]]), counts),
                         ("empty image", keys, np.where(np.arange(n) == 1, 0, counts))]:
        for mode in ("host", "graph"):
            engine.set_execution(mode)
            engine.reset_sums()
            kt = torch.tensor(np.ascontiguousarray(kk, np.int32), device="cuda")
            ct = torch.tensor(np.ascontiguousarray(cc, np.int32), device="cuda")
            with pytest.raises(Exception) as e:
                _run_given(engine, t, n, kt, ct, 4)
            assert "given coarse significance" in str(e.value), (what, mode, str(e.value))
            s = engine.get_sums()
            assert np.all(s["wsum_pdf_class"] == 0), (what, mode)  # nothing accumulated
    engine.set_execution("graph")


def test_given_significance_config3_shape(engine):
    """The batched adaptive workload's shape on three particles: 192^2 (14615 fine pixels, two pixel ranges
    per backprojected slice), HEALPix order 3 x 48 psi (36864 rotations per class) x 49 translations,
    oversampling 1, lognormal significant counts around the true poses; against the oracle run with the
    same significance, host-synchronised and as a graph."""
    import torch
    from test_gpu_parity import check_argmax_near_ties
    prob, batch, truth = synth.make_problem(is3d=True, ori=192, current=192, coarse=32, K=2, order=3,
                                            offset_range=4.0, oversampling=1, n=3, snr=0.05, seed=52, nblobs=100)
    n = batch["n"]
    rng = np.random.default_rng(53)
    counts = synth.lognormal_counts(n, 60, 1.0, 1, 400, rng)
    keys, counts = synth.clustered_significance(prob, np.asarray(truth, float), counts)
    print(f"[given significance config3 shape] D={prob['nr_dir']} PSI={prob['nr_psi']} T={prob['nr_trans']} "
          f"counts={counts.tolist()}", flush=True)
    ref = O.estep(prob, dict(batch, coarse_sig_in=synth.significance_mask(prob, keys, counts)))
    t = _device_batch(batch)
    kt = torch.tensor(keys, device="cuda")
    ct = torch.tensor(counts, device="cuda")
    try:
        for mode in ("host", "graph", "graph"):
            synth.load_engine(engine, prob)
            engine.set_execution(mode)
            r = _run_given(engine, t, n, kt, ct, keys.shape[1])
            s = engine.get_sums()
            assert np.array_equal(r["nr_sign"], counts), mode
            check_argmax_near_ties(engine, prob, batch, r, ref, f"config3 shape {mode}")
            np.testing.assert_allclose(r["min_diff2"], ref["min_diff2"], rtol=1e-5)
            _check_sums(r, s, ref, prob, engine, f"config3 shape {mode}")
    finally:
        engine.set_execution("graph")
