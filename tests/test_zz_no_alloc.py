"""The allocation-free projection entry point against rome_project_device, through the C-ABI.

Two calls in a row (the second with fewer rotations, so the shared staging array still holds the first
call's tail) must give the same planes, bit for bit, as the allocating entry point; more rotations than
the staging array holds are refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

import rome_amd
from rome_amd import synth

pytestmark = pytest.mark.gpu

MAX_ROT = 16384


def _rotations(n, seed):
    rng = np.random.default_rng(seed)
    ang = np.column_stack([rng.uniform(-180, 180, n), np.degrees(np.arccos(rng.uniform(-1, 1, n))),
                           rng.uniform(-180, 180, n)])
    return np.stack([synth.euler_matrix(*a) for a in ang])


def test_noalloc_projection_matches_allocating_entry_point():
    import torch
    prob, _, _ = synth.make_problem(is3d=True, ori=32, current=32, coarse=16, K=2, order=1, n=2, snr=1.0, seed=5)
    eng = rome_amd.Engine(0)
    try:
        synth.load_engine(eng, prob)
        size = 32
        npx = size * (size // 2 + 1)
        for call, (k, nrot, seed) in enumerate([(1, 3000, 1), (0, 700, 2)]):
            print(f"[no-alloc projection] call {call}: class {k}, {nrot} rotations", flush=True)
            A = _rotations(nrot, seed)
            want = torch.full((nrot, npx, 2), float("nan"), dtype=torch.float32, device="cuda")
            got = torch.full((nrot, npx, 2), float("nan"), dtype=torch.float32, device="cuda")
            eng.project_device(k, size, A, want.data_ptr())
            eng.project_device_noalloc(k, size, A, got.data_ptr())
            torch.cuda.synchronize()
            w, g = want.cpu().numpy(), got.cpu().numpy()
            assert np.all(np.isfinite(w))
            assert np.array_equal(w.view(np.uint32), g.view(np.uint32)), call
            assert np.any(w != 0)
        # past the staging array's size: an argument error, nothing launched
        A = np.tile(np.eye(3), (MAX_ROT + 1, 1, 1))
        out = torch.zeros(1, dtype=torch.float32, device="cuda")
        rc = rome_amd.lib().rome_project_device_noalloc(eng._h, 0, size, MAX_ROT + 1,
                                                         A.ctypes.data_as(C.POINTER(C.c_double)),
                                                         C.c_void_p(out.data_ptr()))
        assert rc == -1
        assert b"exceed" in rome_amd.lib().rome_last_error(eng._h)
    finally:
        eng.close()
