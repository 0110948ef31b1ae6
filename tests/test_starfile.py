"""STAR / MRC / MRCS files in the reference's layouts (src/metadata.cpp:343-395, src/mrcs.cpp:127-236) and the
synthetic-dataset writer: round trips, the 1024-byte header words the reference reads, and the real-space
images reproducing the E-step's Fourier inputs through the (pinned) image-preparation oracle."""
import os

import numpy as np

from oracle import imageprep as OI
from rome_amd import starfile, synth


def test_mrcs_round_trip_and_header(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.normal(0, 1, (5, 16, 16)).astype(np.float32)
    p = str(tmp_path / "x.mrcs")
    starfile.write_mrcs(p, a, pixel_size=1.5)
    assert os.path.getsize(p) == 1024 + a.nbytes
    h = np.fromfile(p, dtype="<i4", count=256)
    assert list(h[:4]) == [16, 16, 5, 2] and list(h[16:19]) == [1, 2, 3]
    assert h.view(np.float32)[10] == 16 * 1.5
    hd = starfile.read_mrc_header(p)
    assert (hd["nx"], hd["ny"], hd["nz"]) == (16, 16, 5) and hd["pixel_size"] == 1.5
    np.testing.assert_array_equal(starfile.read_mrcs(p), a)
    np.testing.assert_array_equal(starfile.read_mrcs(p, [3, 1]), a[[2, 0]])


def test_star_round_trip(tmp_path):
    p = str(tmp_path / "d.star")
    cols = {"rlnImageName": [starfile.image_name(i + 1, "s.mrcs") for i in range(3)],
            "rlnAngleRot": [10.5, -20.25, 0.0], "rlnClassNumber": [1, 2, 1]}
    starfile.write_star(p, cols)
    txt = open(p).read()
    assert txt.startswith("\ndata_\n\nloop_\n_rlnImageName #1\n_rlnAngleRot #2\n_rlnClassNumber #3\n")
    r = starfile.read_star(p)
    assert r["rlnImageName"] == cols["rlnImageName"]
    np.testing.assert_allclose(r["rlnAngleRot"], cols["rlnAngleRot"])
    assert list(r["rlnClassNumber"]) == [1, 2, 1]
    assert starfile.parse_image_name(r["rlnImageName"][2]) == (3, "s.mrcs")


def test_dataset_images_reproduce_the_fourier_inputs(tmp_path):
    prob, batch, truth = synth.make_problem(is3d=True, ori=32, current=32, coarse=16, K=2, order=1, n=4, snr=2.0,
                                            seed=5)
    path = synth.write_dataset(str(tmp_path), "parts", batch, truth, 32)
    star = starfile.read_star(path)
    assert len(star["rlnImageName"]) == 4
    np.testing.assert_allclose(star["rlnAngleRot"], [t[1] for t in truth], atol=1e-5)
    imgs = starfile.stack_images(star, str(tmp_path))
    F = np.asarray(batch["fnomask_re"]) + 1j * np.asarray(batch["fnomask_im"])
    for i in range(4):
        G = OI.fft_half(OI.center_fft(imgs[i], True)).reshape(-1)
        # float32 storage; the x = 0 and x = ori/2 columns of the input are not Hermitian (their y < 0 halves
        # hold independent noise), so a real image keeps only their symmetric parts
        x = np.tile(np.arange(17), 32)
        m = (x > 0) & (x < 16)
        np.testing.assert_allclose(G[m], F[i][m], rtol=0, atol=2e-6 * np.max(np.abs(F[i])))
