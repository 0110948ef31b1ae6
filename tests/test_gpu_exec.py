"""Execution modes of a pool: host-synchronised launches, the device-driven fine pass (capacities, counts read
on the device, the fine diff2 on a work queue) and its captured hipGraph, replayed.

The three must give the same E-step: metadata, argmax, NR_SIGN and the per-image sums bit for bit (their
kernels are deterministic), the weighted sums, the norm correction (and the NORM it gives) and the
backprojector grids to the order of their atomic additions (1e-12 relative for the FP64 sums, 1e-6 for
the FP32 grids).  A graph is replayed several times
with the program's own eager work and synchronisations between the replays, and every replay is checked.
Pools the device-driven path cannot finish (a coarse list or the significant-sample keys past their caps)
are rerun host-synchronised and still give the same results."""
import numpy as np
import pytest

from rome_amd import synth

from test_gpu_parity import _case

pytestmark = pytest.mark.gpu

CASES = [("3d", "fp32"), ("3d_window_groups_offsets", "fp32"), ("2d_two_pass", "fp32"), ("3d_zero_prior", "fp32"),
         ("3d", "bf16"), ("3d_window_groups_offsets", "fp16"), ("3d_paired_k1", "fp16")]
EXACT = ("argmax", "nr_sign", "min_diff2", "sum_weight", "sig_weight", "min_diff2_coarse", "sig_weight_coarse")
META_EXACT = [0, 1, 2, 3, 4, 5, 6, 8]  # all but NORM (column 7), which follows the norm-correction sum
# the coarse weight sum adds the coarse list's entries in the order the cross-term epilogue appended them
# (atomic reservations), the norm-correction sum a pool's significant samples in the order of their atomic
# additions: equal to the last bits only
CLOSE = ("sum_weight_coarse", "wsum_norm_correction")


def _pool(engine, prob, batch, prec, mode):
    synth.load_engine(engine, prob, precision=prec)
    engine.set_execution(mode)
    r = engine.expectation(batch)
    return r, engine.get_sums(), [engine.get_backprojector(k) for k in range(prob["nr_classes"])], \
        engine.execution_stats()


def _same(a, b, what):
    ra, sa, ga, _ = a
    rb, sb, gb, _ = b
    for k in EXACT:
        assert np.array_equal(ra[k], rb[k]), f"{what}: {k}"
    assert np.array_equal(ra["meta"][:, META_EXACT], rb["meta"][:, META_EXACT]), f"{what}: meta"
    np.testing.assert_allclose(ra["meta"][:, 7], rb["meta"][:, 7], rtol=1e-12, err_msg=f"{what}: NORM")
    for k in CLOSE:
        np.testing.assert_allclose(ra[k], rb[k], rtol=1e-12, err_msg=f"{what}: {k}")
    for k in sa:
        np.testing.assert_allclose(sa[k], sb[k], rtol=1e-12, atol=1e-300, err_msg=f"{what}: {k}")
    for (da, wa), (db, wb) in zip(ga, gb):
        assert np.linalg.norm(da - db) <= 1e-6 * np.linalg.norm(db), what
        assert np.linalg.norm(wa - wb) <= 1e-6 * np.linalg.norm(wb), what


@pytest.mark.parametrize("name,prec", CASES)
def test_execution_modes_agree(engine, name, prec):
    prob, batch, _ = _case(name)
    print(f"[exec {name} {prec}] n={batch['n']}", flush=True)
    try:
        host = _pool(engine, prob, batch, prec, "host")
        assert host[3]["last"] == 0
        dev = _pool(engine, prob, batch, prec, "device")
        assert dev[3]["last"] == 1, dev[3]
        _same(dev, host, "device-driven")
        first = _pool(engine, prob, batch, prec, "graph")  # runs device-driven and captures
        assert first[3]["last"] == 1 and first[3]["graph_captures"] >= 1, first[3]
        _same(first, host, "graph mode, first pool")
        for rep in range(3):
            # eager work between replays: a host-synchronised pool of other images
            engine.set_execution("host")
            engine.reset_sums()
            other = dict(batch, fmask_re=batch["fmask_re"][::-1].copy(), fmask_im=batch["fmask_im"][::-1].copy())
            engine.expectation(other)
            engine.set_execution("graph")
            engine.reset_sums()
            r = engine.expectation(batch)
            st = engine.execution_stats()
            assert st["last"] == 2, (rep, st)
            got = (r, engine.get_sums(), [engine.get_backprojector(k) for k in range(prob["nr_classes"])], st)
            _same(got, host, f"graph replay {rep}")
    finally:
        engine.set_execution("graph")


def test_device_driven_overflow_reruns_host(engine):
    """A coarse list cap below the lists' sizes, then a key cap below the significant counts: the device-driven
    pool aborts on the device, accumulates nothing, and is rerun host-synchronised with the same results."""
    prob, batch, _ = _case("3d")
    print("[exec overflow reruns]", flush=True)
    try:
        ref = _pool(engine, prob, batch, "fp32", "host")
        engine.set_coarse_mode("list", 64)
        r = _pool(engine, prob, batch, "fp32", "device")
        assert r[3]["last"] == 1 | 8, r[3]
        _same(r, ref, "list overflow rerun")
        engine.set_coarse_mode("list", 0)
        small = int(np.min(ref[0]["nr_sign"]))
        assert small >= 1
        engine.set_significance_cap(max(1, small - 1))
        r = _pool(engine, prob, batch, "fp32", "device")
        assert r[3]["last"] == 1 | 8, r[3]
        _same(r, ref, "key overflow rerun")
    finally:
        engine.set_coarse_mode("list", 0)
        engine.set_significance_cap(-1)
        engine.set_execution("graph")


def test_work_past_the_first_launches(engine):
    """Grids of 8 blocks and no history of earlier pools: the fine diff2's first launch covers 4 tasks per
    image and the rest goes through its work queue, the projection and the post slices run most of their
    items in their tails; device-driven and replayed graphs still give the host-synchronised results."""
    # SNR 0.02: broad posteriors, 347 .. 4248 significant coarse samples per image
    prob, batch, _ = synth.make_problem(is3d=True, ori=32, current=32, coarse=16, K=2, order=1, n=8, snr=0.02,
                                        seed=5)
    n = batch["n"]
    print(f"[exec past first launches] n={n}", flush=True)
    try:
        engine.set_significance_cap(8192)
        host = _pool(engine, prob, batch, "fp32", "host")
        st = engine.batch_stats()
        assert st["tasks"] > 4 * n and st["unique_rot"] > 4 * n, st  # the queue and the tails take part
        engine.set_persistent_grids([8, 8, 8, 8])
        dev = _pool(engine, prob, batch, "fp32", "device")
        assert dev[3]["last"] == 1, dev[3]
        _same(dev, host, "device-driven, 8-block grids")
        engine.set_persistent_grids([8, 8, 8, 8])
        first = _pool(engine, prob, batch, "fp32", "graph")
        _same(first, host, "graph, first pool")
        engine.reset_sums()
        r = engine.expectation(batch)
        st = engine.execution_stats()
        assert st["last"] == 2, st
        _same((r, engine.get_sums(), [engine.get_backprojector(k) for k in range(prob["nr_classes"])], st), host,
              "graph replay, 8-block grids")
    finally:
        engine.set_persistent_grids([0, 0, 0, 0])
        engine.set_significance_cap(-1)
        engine.set_execution("graph")
