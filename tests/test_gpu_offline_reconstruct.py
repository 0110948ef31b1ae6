"""The offline reconstruction (rome_amd.reconstruct, the reference's rome_reconstruct) on the GPU against the
oracle: particles written as STAR + MRCS, read back, transformed, shifted by their origins, multiplied by
their CTF, back-projected at their angles and reconstructed (10 gridding iterations, no MAP term, whole-map
FSC); the oracle does the same with its pinned primitives (image preparation, tabulated shift, CTF,
backprojector, reconstruct).  Maps within 1e-5 relative L2 (FP32 grids on the GPU)."""
import numpy as np
import pytest

import oracle as O
import oracle.mstep as OM
from oracle import imageprep as OI
from rome_amd import reconstruct, starfile, synth

pytestmark = pytest.mark.gpu


def _oracle_map(star, base, ori, angpix):
    imgs = starfile.stack_images(star, base)
    n = len(imgs)
    r_max = ori // 2
    pad = O.pad_size(2 * r_max)
    padF = pad // 2 + 1
    data = np.zeros(2 * pad * pad * padF)
    weight = np.zeros(pad * pad * padF)
    for i in range(n):
        F = OI.fft_half(OI.center_fft(imgs[i], True)).reshape(-1)
        sr, si = O.shift_image(F.real, F.imag, ori, star["rlnOriginX"][i], star["rlnOriginY"][i], ori)
        c = OI.ctf_image(ori, ori, angpix, star["rlnDefocusU"][i], star["rlnDefocusV"][i], star["rlnDefocusAngle"][i],
                         star["rlnVoltage"][i], star["rlnSphericalAberration"][i], star["rlnAmplitudeContrast"][i],
                         star["rlnBfactor"][i]).reshape(-1)
        A = O.euler_matrix(star["rlnAngleRot"][i], star["rlnAngleTilt"][i], star["rlnAnglePsi"][i])
        O.backproject3d(data, weight, pad, r_max, A, ori, sr * c, si * c, c * c)
    d = (data[0::2] + 1j * data[1::2]).reshape(pad, pad, padF)
    w = weight.reshape(pad, pad, padF)
    return OM.reconstruct(d, w, ori, 2 * r_max, max_iter_preweight=10, do_map=False, tau2=np.zeros(ori // 2 + 1),
                          is_whole_instead_of_half=True, minres_map=-1, normalise=1.0)[0]


def test_offline_reconstruction_matches_oracle(engine, tmp_path):
    ori, angpix = 32, 2.0
    prob, batch, truth = synth.make_problem(is3d=True, ori=ori, current=ori, coarse=16, K=1, order=1, n=160, snr=4.0,
                                            seed=81, angpix=angpix, offset_range=2.0)
    path = synth.write_dataset(str(tmp_path), "parts", batch, truth, ori, angpix)
    print(f"[offline reconstruct] {path}", flush=True)
    vol = reconstruct.reconstruct_star(path, str(tmp_path / "map"), angpix, do_ctf=True, engine=engine)
    star = starfile.read_star(path)
    want = _oracle_map(star, str(tmp_path), ori, angpix)
    rel = np.linalg.norm(vol - want) / np.linalg.norm(want)
    print(f"[offline reconstruct] map rel L2 {rel:.2e}", flush=True)
    assert rel < 1e-5
    written = starfile.read_mrc(str(tmp_path / "map.mrc"))
    np.testing.assert_allclose(written, vol.astype(np.float32), rtol=0, atol=0)
