"""GPU parity of rome_reconstruct (MyBackProjector::reconstruct, src/reconstruct.cpp:1929-2522) against
the FP64 oracle.  Both sides are FP64 end to end (hipFFT vs pocketfft), so maps agree to 1e-9 of their
largest value and the spectra (sigma2, data_vs_prior, tau2) to 1e-10 relative.  Cases: 3D and 2D,
current_size = ori (the 2 ori window shrinks the padded grid) and < ori (it grows it), MAP term on and
off, tau2 from the FSC, and the context's own backprojector after an E-step."""
import numpy as np
import pytest

import oracle as O
import oracle.mstep as OM
from rome_amd import synth

pytestmark = pytest.mark.gpu


def _engine_for(engine, dim, ori, cur):
    prob, batch, _ = synth.make_problem(is3d=dim == 3, ori=ori, current=cur, coarse=min(cur, 16), K=2, n=2, seed=31)
    synth.load_engine(engine, prob)
    return prob, batch


def _grids(rng, dim, cur):
    P = 2 * (2 * (cur // 2) + 1) + 1
    shape = (P, P, P // 2 + 1) if dim == 3 else (P, P // 2 + 1)
    c = P // 2
    idx = np.indices(shape)
    r2 = sum((idx[a] - c) ** 2 for a in range(dim - 1)) + idx[-1] ** 2
    inside = r2 <= (2 * (cur // 2)) ** 2
    data = (rng.normal(size=shape) + 1j * rng.normal(size=shape)) * inside / (1.0 + np.sqrt(r2))
    weight = rng.uniform(0.5, 2.0, shape) * inside
    return data, weight


def _compare(got, want, name):
    vol, t2, sig, dvp = got
    v2, t2r, sr, dr = want
    scale = np.max(np.abs(v2))
    assert np.max(np.abs(vol - v2)) <= 1e-9 * scale, name
    np.testing.assert_allclose(sig, sr, rtol=1e-10, atol=0, err_msg=name)
    np.testing.assert_allclose(dvp, dr, rtol=1e-10, atol=0, err_msg=name)
    np.testing.assert_allclose(t2, t2r, rtol=1e-10, atol=0, err_msg=name)


@pytest.mark.parametrize("dim,ori,cur,do_map", [(3, 32, 32, True), (3, 32, 20, True), (3, 32, 32, False),
                                                (2, 64, 64, True), (2, 64, 40, True)])
def test_reconstruct_matches_oracle(engine, dim, ori, cur, do_map):
    print(f"[reconstruct case dim={dim} ori={ori} cur={cur} do_map={do_map}]", flush=True)
    _engine_for(engine, dim, ori, cur)
    rng = np.random.default_rng(ori + cur + dim)
    data, weight = _grids(rng, dim, cur)
    tau2 = rng.uniform(0.5, 2.0, ori // 2 + 1)
    kw = dict(max_iter_preweight=10, do_map=do_map, minres_map=5, tau2_fudge=1.3, normalise=3.7)
    got = engine.reconstruct(0, tau2, data=data, weight=weight, **kw)
    want = OM.reconstruct(data, weight, ori, cur, tau2=tau2, **kw)
    _compare(got, want, f"dim{dim} cur{cur} map{do_map}")


def test_reconstruct_tau2_from_fsc(engine):
    dim, ori, cur = 3, 32, 32
    _engine_for(engine, dim, ori, cur)
    rng = np.random.default_rng(77)
    data, weight = _grids(rng, dim, cur)
    fsc = np.clip(np.linspace(1.0, -0.2, ori // 2 + 1), -1, 1)
    kw = dict(max_iter_preweight=4, do_map=True, minres_map=5, tau2_fudge=2.0, normalise=1.0,
              update_tau2_with_fsc=True)
    got = engine.reconstruct(1, np.zeros(ori // 2 + 1), data=data, weight=weight, fsc=fsc, **kw)
    want = OM.reconstruct(data, weight, ori, cur, tau2=np.zeros(ori // 2 + 1), fsc=fsc, **kw)
    _compare(got, want, "fsc")


def test_reconstruct_from_the_engine_backprojector(engine):
    """The context's accumulated grids after an E-step, against the oracle fed the same grids."""
    prob, batch = _engine_for(engine, 3, 32, 32)
    engine.expectation(batch)
    d, w = engine.get_backprojector(1)
    data = d[0::2] + 1j * d[1::2]
    shape = (prob["pad"], prob["pad"], prob["pad"] // 2 + 1)
    data, w = data.reshape(shape), w.reshape(shape)
    tau2 = np.full(17, 0.8)
    got = engine.reconstruct(1, tau2, normalise=2.0)
    want = OM.reconstruct(data, w, 32, 32, tau2=tau2, normalise=2.0)
    _compare(got, want, "engine grids")
