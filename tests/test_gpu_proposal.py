"""GPU tests of the learned pose proposal (hand-written training and inference kernels, BF16 MFMA).

- gradient check: the network's outputs, loss and every parameter gradient from the HIP kernels against
  the FP64 CPU oracle (oracle/mlp.py) on the same BF16-representable parameters and inputs, with the
  oracle rounding the product operands to BF16 as the GPU does (relative L2 within 1e-2: FP32
  accumulation and BF16 re-rounding of the activations), and against the plain FP64 oracle within what
  BF16 operands allow;
- training targets: the cells and shifts the E-step's own posteriors give (COLLECT mode) against the
  oracle's fine weights of the same particles, aggregated the same way;
- the learned E-step: trained on the exhaustive E-step's posteriors, then run on held-out particles, it
  must agree with the exhaustive pass on the max-posterior poses while computing diff2 for a fraction of
  the rotations; an untrained network at the same settings gains nothing (its predicted mass is below
  the bar, so every image takes every rotation).
"""
import numpy as np
import pytest

import oracle as O
from oracle import mlp
from rome_amd import synth

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


def _small_problem(n=8, seed=61):
    return synth.make_problem(is3d=True, ori=32, current=32, coarse=16, K=2, order=1, n=n, snr=0.5, seed=seed,
                              oversampling=1, offset_range=2.0)


def test_gradients_match_oracle(engine):
    prob, _, _ = _small_problem()
    synth.load_engine(engine, prob, precision="bf16")
    engine.proposal_configure(hidden=128, nr_coef=100, target_cells=16, max_train=64, seed=3)
    rng = np.random.default_rng(5)
    shapes = engine._prop_shapes()
    p = {}
    for k, shp in shapes.items():
        fan = shp[1] if len(shp) == 2 else 1
        v = rng.normal(0, 1.0 / np.sqrt(fan), shp) if len(shp) == 2 else rng.normal(0, 0.1, shp)
        p[k] = mlp.to_bf16(v)  # BF16-representable: the GPU's BF16 copies are the masters exactly
    engine.proposal_set_params(p)
    B = 64
    ncell = engine.prop["ncell"]
    x = mlp.to_bf16(rng.normal(0, 1, (B, shapes["W1"][1])))
    t = np.zeros((B, ncell))
    for i in range(B):
        c = rng.choice(ncell, 5, replace=False)
        t[i, c] = rng.random(5)
        t[i] /= t[i].sum()
    t = t.astype(np.float32).astype(np.float64)
    sh = rng.normal(0, 0.5, (B, 2)).astype(np.float32).astype(np.float64)
    z_gpu = engine.proposal_forward(x[:10])
    z_ref, _ = mlp.forward(p, x[:10], bf16=True)
    assert _rel(z_gpu, z_ref) < 1e-2
    loss, g = engine.proposal_gradients(x, t, sh)
    ref_loss, ref_g, _ = mlp.loss_and_grads(p, x, t, sh, ncell, 1.0, bf16=True)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    _, g64, _ = mlp.loss_and_grads(p, x, t, sh, ncell, 1.0)
    for k in shapes:
        e16, e64 = _rel(g[k], ref_g[k]), _rel(g[k], g64[k])
        print(f"[grad {k}] vs bf16 oracle {e16:.2e}, vs fp64 oracle {e64:.2e}", flush=True)
        assert e16 < 1e-2, k
        assert e64 < 6e-2, k


def test_training_targets_match_oracle_posteriors(engine):
    prob, batch, _ = _small_problem(n=6, seed=62)
    synth.load_engine(engine, prob, precision="bf16")
    engine.proposal_configure(hidden=128, nr_coef=100, target_cells=64, max_train=64, seed=3)
    engine.proposal_mode("collect")
    try:
        engine.expectation(batch)
    finally:
        engine.proposal_mode("off")
    n = batch["n"]
    x, t, sh = engine.proposal_training_rows(0, n)
    ref = O.estep(prob, batch, dump_fine=True, dump_fine_weight=True)
    K, D, PSI = prob["nr_classes"], prob["nr_dir"], prob["nr_psi"]
    nb = 4
    fw = ref["fine_weight"]  # [n][K][D][PSI][T][8][4]
    for i in range(n):
        w = np.nan_to_num(fw[i]) / ref["sum_weight"][i]
        cells = np.zeros((K, 4 * D, nb))
        k, d, s, tt, o, ot = np.nonzero(w)
        np.add.at(cells, (k, 4 * d + o // 2, s * nb // PSI), w[k, d, s, tt, o, ot])
        want = cells.reshape(-1)
        keep = np.argsort(-want)[:64]
        wk = np.zeros_like(want)
        wk[keep] = want[keep]
        wk /= wk.sum()
        assert _rel(t[i], wk) < 2e-3, i  # FP32 diff2 and near-threshold coarse samples
        sx = ref["meta"][i, 5]  # the posterior-mean shift is near the max-posterior one for sharp posteriors
        assert abs(sh[i, 0] * 2.0 - sx) < 1.0
    # features: whitened low-resolution coefficients at unit RMS
    assert np.allclose(np.sqrt(np.mean(x ** 2, axis=1)), 1.0, rtol=2e-2)


def _learned_setup(engine, n_train, n_test, seed):
    import torch
    prob, _, _ = synth.make_problem(is3d=True, ori=64, current=64, coarse=24, K=2, order=2, n=1, snr=1.0,
                                    seed=seed, oversampling=1, offset_range=2.0, nblobs=80)
    synth.load_engine_tables(engine, prob)
    F, ctf, truth, sigma2 = synth.make_particles_device(engine, prob["nr_classes"], n_train + n_test, 64, 64, 2.0,
                                                        2.0, 1.0, seed + 1)
    prob["sigma2_noise"] = np.full((1, 33), sigma2)
    return prob, F, ctf, truth


def _run_device(engine, F, ctf, o, n):
    import torch
    dev = F.device
    z = torch.zeros(n, dtype=torch.float64, device=dev)
    power = torch.zeros((n, 33), dtype=torch.float64, device=dev)
    one = torch.ones(n, dtype=torch.float64, device=dev)
    grp = torch.zeros(n, dtype=torch.int32, device=dev)
    F2 = F.shape[1]
    r = engine.expectation_device(n, F.data_ptr() + o * F2 * 8, F.data_ptr() + o * F2 * 8,
                                  ctf.data_ptr() + o * F2 * 4, z.data_ptr(), power.data_ptr(), z.data_ptr(),
                                  z.data_ptr(), one.data_ptr(), grp.data_ptr(), results=True)
    torch.cuda.synchronize()
    return r


def test_learned_proposal_estep(engine):
    n_train, n_test = 6144, 512
    prob, F, ctf, truth = _learned_setup(engine, n_train, n_test, 71)
    synth.load_engine(engine, prob, precision="bf16")
    M = prob["nr_classes"] * prob["nr_dir"] * prob["nr_psi"]
    # untrained network: its proposals are not trusted, every image takes every rotation
    engine.proposal_configure(max_train=n_train, seed=7)
    engine.proposal_mode("use")
    try:
        _run_device(engine, F, ctf, n_train, n_test)
        st0 = engine.proposal_stats()
    finally:
        engine.proposal_mode("off")
    # trained on the exhaustive E-step's posteriors of the training particles
    engine.proposal_configure(max_train=n_train, seed=7)
    engine.proposal_mode("collect")
    try:
        for o in range(0, n_train, 1024):
            _run_device(engine, F, ctf, o, 1024)
    finally:
        engine.proposal_mode("off")
    losses = engine.proposal_train(epochs=12, batch=256, lr=1e-3)
    print("[learned] losses", np.round(losses, 3), flush=True)
    assert losses[-1] < losses[0]
    ref = _run_device(engine, F, ctf, n_train, n_test)  # exhaustive mixed-precision pass, held-out particles
    # the per-image predicted-mass gate: below it an image takes every rotation (the audit judges the
    # captured posterior mass of the pool); statistics of each pool from the differences of the totals
    res = {}
    prev = engine.proposal_stats()
    for gate in (0.99, 0.9, 0.5):
        engine.proposal_set_gate(gate)
        engine.proposal_mode("use")
        try:
            r = _run_device(engine, F, ctf, n_train, n_test)
            st = engine.proposal_stats()
        finally:
            engine.proposal_mode("off")
        agree = float(np.mean(np.all(r["argmax"] == ref["argmax"], axis=1)))
        cand = (st["cand_rot_per_image"] * st["images"] - prev["cand_rot_per_image"] * prev["images"]) / n_test
        frac = cand / M
        every = st["images_all_rotations"] - prev["images_all_rotations"]
        res[gate] = (agree, frac, every, st["min_audit_mass"], st["fallback_pools"] - prev["fallback_pools"])
        print(f"[learned] gate {gate}: agreement {agree:.4f}, rotations evaluated {frac:.3f}, images on every "
              f"rotation {every}, lowest audited captured mass {st['min_audit_mass']:.4f}", flush=True)
        prev = st
    agree, frac, _, _, _ = res[0.9]
    print(f"[learned] untrained {st0}", flush=True)
    assert st0["cand_rot_per_image"] >= 0.99 * M  # no gain without the trained network
    assert st["train_images"] == n_train
    # at the default gate (0.9) the trained network restricts the coarse pass to under half the rotations
    # and agrees with the exhaustive pass on the max-posterior poses of 99 % of the held-out particles
    assert agree >= 0.99, res
    assert frac < 0.5, res
