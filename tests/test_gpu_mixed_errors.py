"""Error statement of the precision arms against the FP64 oracle (the mixed-precision configuration of
the MAP3D workload): coarse diff2 (max and median relative error), the normalised fine posterior
(L1 distance per image, both posteriors formed with the same priors from each side's own diff2 and
significant set) and the backprojector maps (FSC per shell, relative L2).  The FP32 arm is the floor.

Bounds asserted: FP32 arm L1 <= 1e-4 (its diff2 carries FP32 rounding: 2e-6 relative, 2-20e-6 of L1
measured) and FSC >= 1 - 1e-9; mixed arms L1 <= 5e-3 (the significant set
can change only inside its 0.1 % tail) and FSC >= 1 - 1e-5 on every shell that holds signal; coarse
diff2 relative error within the operand-rounding bound of the format (median far below it)."""
import math

import numpy as np
import pytest

import oracle as O
from rome_amd import synth

from test_gpu_parity import _case

pytestmark = pytest.mark.gpu

U = {"fp32": 2.0 ** -24, "bf16": 2.0 ** -8, "fp16": 2.0 ** -11}
CASES = ["3d", "3d_window_groups_offsets", "2d_two_pass"]


@pytest.fixture(scope="module", params=CASES)
def case(request):
    prob, batch, _ = _case(request.param)
    ref = O.estep(prob, batch, dump_coarse=True, dump_fine=True, dump_fine_weight=True)
    return request.param, prob, batch, ref


def _prior(prob, batch, i, k, d, t):
    """pdf_direction * pdf_offset of a fine sample (calculatePdfOffset on the coarse translation)."""
    ox = prob["trans_x"][t] + batch["old_offx"][i]
    oy = prob["trans_y"][t] + batch["old_offy"][i]
    td2 = (ox - prob["prior_offx"][k]) ** 2 + (oy - prob["prior_offy"][k]) ** 2
    s2 = prob["sigma2_offset"]
    po = (0.0 if td2 > 0 else 1.0) if s2 < 1e-4 else math.exp(td2 / (-2 * s2)) / (2 * math.pi * s2)
    return prob["pdf_direction"][k][d] * po


def _posteriors(prob, batch, keys, d2, shape):
    """{image: {key: normalised weight}} from a fine list (keys in the oracle's dense order)."""
    n, K, D, PSI, T, NOR, NOT = shape
    out = {}
    per_img = keys // (K * D * PSI * T * NOR * NOT)
    for i in np.unique(per_img):
        sel = per_img == i
        kk, dd = keys[sel], d2[sel]
        w = {}
        dmin = dd.min()
        for key, v in zip(kk, dd):
            r = key // (NOR * NOT)
            t = r % T
            r //= T
            r //= PSI
            d = r % D
            k = (r // D) % K
            w[int(key)] = _prior(prob, batch, i, k, d, t) * (0.0 if v - dmin > 700 else math.exp(-(v - dmin)))
        s = sum(w.values())
        out[int(i)] = {q: x / s for q, x in w.items()}
    return out


def _fsc(a, b, shape, pad):
    """FSC per integer shell of two complex grids in the projector layout (origin at pad // 2)."""
    a, b = a.reshape(shape), b.reshape(shape)
    c = pad // 2
    idx = np.indices(shape)
    r = np.sqrt(sum((idx[q] - c) ** 2 for q in range(len(shape) - 1)) + idx[-1] ** 2)
    sh = np.floor(r + 0.5).astype(int).ravel()
    num = np.bincount(sh, (a * np.conj(b)).real.ravel())
    den = np.sqrt(np.bincount(sh, (np.abs(a) ** 2).ravel()) * np.bincount(sh, (np.abs(b) ** 2).ravel()))
    ok = den > 1e-30 * den.max()
    return num[ok] / den[ok]


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_error_statement(engine, case, prec):
    name, prob, batch, ref = case
    n = batch["n"]
    synth.load_engine(engine, prob, precision=prec)
    engine.set_coarse_mode("full")
    try:
        engine.expectation(batch)
        d = engine.get_coarse_diff2(n).reshape(n, -1)
    finally:
        engine.set_coarse_mode("list")
    # coarse diff2
    o = ref["coarse_diff2"].reshape(n, -1)
    m = ~np.isnan(o)
    rel = np.abs(d[m] - o[m]) / np.abs(o[m])
    # fine posterior
    shape = ref["fine_diff2"].shape
    keys, d2 = engine.get_fine_list()
    fo = ref["fine_diff2"].reshape(-1)
    okeys = np.flatnonzero(~np.isnan(fo))
    pg = _posteriors(prob, batch, keys, d2, shape)
    po = _posteriors(prob, batch, okeys, fo[okeys], shape)
    # the posterior helper reproduces the oracle's own normalised weights
    fw = ref["fine_weight"].reshape(n, -1)
    for i in range(n):
        ow = np.nan_to_num(fw[i]) / np.nansum(fw[i])
        for key, p in list(po[i].items())[:50]:
            assert abs(ow[key - i * fw.shape[1]] - p) <= 1e-12 + 1e-9 * p
    l1 = [sum(abs(pg[i].get(q, 0.0) - po[i].get(q, 0.0)) for q in set(pg[i]) | set(po[i])) for i in range(n)]
    # maps
    fsc_min, rl2 = 1.0, 0.0
    for k in range(prob["nr_classes"]):
        dg, _ = engine.get_backprojector(k)
        od = ref["bp_data"][k]
        if np.linalg.norm(od) == 0:
            continue
        a = dg[0::2] + 1j * dg[1::2]
        b = od[0::2] + 1j * od[1::2]
        gshape = (prob["pad"], prob["pad"], prob["pad"] // 2 + 1) if prob["is_3d"] else (prob["pad"], prob["pad"] // 2 + 1)
        fsc_min = min(fsc_min, float(_fsc(a, b, gshape, prob["pad"]).min()))
        rl2 = max(rl2, float(np.linalg.norm(a - b) / np.linalg.norm(b)))
    print(f"[{name} {prec}] diff2 rel max {rel.max():.2e} median {np.median(rel):.2e}; posterior L1 max "
          f"{max(l1):.2e}; map FSC min {fsc_min:.9f}, rel L2 {rl2:.2e}", flush=True)
    scale_bound = 4 * U[prec] * np.max(np.abs(o[m]))  # per-entry absolute bound, as a relative bound below
    assert np.max(np.abs(d[m] - o[m])) <= max(scale_bound, 1e-5 * np.max(np.abs(o[m])))
    assert np.median(rel) < 0.05 * max(4 * U[prec], 1e-5)
    if prec == "fp32":
        assert max(l1) <= 1e-4 and fsc_min >= 1 - 1e-9
    else:
        assert max(l1) <= 5e-3 and fsc_min >= 1 - 1e-5
