"""GPU parity of rome_fourier_map (computeFourierTransformMap, src/reconstruct.cpp:106-223, with
griddingCorrect :61-103) against the FP64 oracle: both FP64 end to end (hipFFT D2Z vs pocketfft), so
the grid agrees to 1e-12 of its largest coefficient, the power spectrum to 1e-10 relative and the
gridding-corrected map to 1e-14; the installed FP32 grid projects like the oracle's grid (1e-5)."""
import numpy as np
import pytest

import oracle as O
import oracle.mstep as OM
from rome_amd import synth

pytestmark = pytest.mark.gpu

CASES = [(3, 32, 32), (3, 32, 20), (2, 64, 48)]


@pytest.mark.parametrize("dim,ori,cur", CASES)
def test_fourier_map_matches_oracle(engine, dim, ori, cur):
    print(f"[fourier map case dim={dim} ori={ori} cur={cur}]", flush=True)
    prob, batch, truth = synth.make_problem(is3d=dim == 3, ori=ori, current=cur, coarse=min(cur, 16), K=2, n=2,
                                            seed=21)
    synth.load_engine(engine, prob)
    rng = np.random.default_rng(7)
    vol = synth.blob_volume(ori, 12, (1.5, 3.0), 0.35 * ori, rng, ndim=dim)
    ps, corr, grid = engine.fourier_map(1, vol)
    g_ref, p_ref, c_ref = OM.fourier_map(vol, cur)
    assert grid.shape == g_ref.shape
    scale = np.max(np.abs(g_ref))
    assert np.max(np.abs(grid - g_ref)) <= 1e-12 * scale
    np.testing.assert_allclose(ps, p_ref, rtol=1e-10, atol=1e-14 * p_ref.max())
    np.testing.assert_allclose(corr, c_ref, rtol=1e-14, atol=0)
    # the installed grid feeds the projector: slices agree with the oracle projector on the oracle grid
    A = np.stack([synth.euler_matrix(*rng.uniform([-180, 0, 0], [180, 180, 360])) for _ in range(4)])
    if dim == 3:
        re, im = engine.project(1, cur, A)
        gi = np.stack([g_ref.real, g_ref.imag], axis=-1).reshape(-1)
        for i in range(len(A)):
            ore, oim = O.project3d(gi, prob["pad"], cur // 2, A[i], cur)
            s = np.max(np.abs(ore)) + np.max(np.abs(oim))
            assert np.max(np.abs(re[i] - ore)) <= 1e-5 * s
            assert np.max(np.abs(im[i] - oim)) <= 1e-5 * s
