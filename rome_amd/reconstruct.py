"""Offline reconstruction of a map from particles with known poses, on the engine (rome_reconstruct).

The counterpart of the reference's rome_reconstruct (apps/rome_reconstruct.cpp:104-330 and its flags,
apps/rome_reconstruct.cpp:30-100): read a STAR file (image names, angles, origins, CTF parameters) and
its MRCS stacks, then per image the centred FFT, the shift by (OriginX, OriginY) and the CTF, back-projected
at (AngleRot, AngleTilt, AnglePsi); finally MyBackProjector::reconstruct without the MAP term (10 gridding
iterations by default, the whole map's FSC) and the map written as MRC.  The image transforms, the
back-projection and the reconstruction run on the GPU (rome_prepare_images, rome_backproject_images,
rome_reconstruct).

    python -m rome_amd.reconstruct --i particles.star --o map --angpix 1.5 [--ctf] [--maxres A] [--iter 10]
"""
import argparse
import math
import os
import sys

import numpy as np

from . import Engine, starfile


def reconstruct_star(star_path, out_path, angpix, maxres=-1.0, do_ctf=False, ctf_phase_flipped=False,
                     only_flip_phases=False, ctf_intact_first_peak=False, nr_iters=10, angular_error=0.0,
                     shift_error=0.0, chunk=1024, engine=None, device=0, seed=33):
    """Reconstruct the map of the particles a STAR file lists; returns the map (and writes out_path.mrc when
    out_path is given)."""
    star = starfile.read_star(star_path)
    base = os.path.dirname(os.path.abspath(star_path))
    for lab in ("rlnImageName", "rlnAngleRot", "rlnAngleTilt", "rlnAnglePsi", "rlnOriginX", "rlnOriginY"):
        if lab not in star:
            raise ValueError(f"{star_path}: label {lab} missing")
    n = len(star["rlnImageName"])
    _, stack0 = starfile.parse_image_name(star["rlnImageName"][0])
    hd = starfile.read_mrc_header(stack0 if os.path.isabs(stack0) else os.path.join(base, stack0))
    ori = hd["nx"]
    r_max = ori // 2 if maxres < 0 else min(ori // 2, int(math.ceil(ori * angpix / maxres)))
    cur = 2 * r_max
    eng = engine or Engine(device)
    try:
        eng.configure(ref_dim=3, ori_size=ori, current_size=cur, coarse_size=cur, nr_classes=1, nr_groups=1,
                      adaptive_oversampling=0, nr_dir=12, nr_psi=6, nr_trans=1, adaptive_fraction=0.999,
                      sigma2_fudge=1.0, sigma2_offset=1.0, avg_norm_correction=1.0, do_ctf_correction=int(do_ctf),
                      refs_are_ctf_corrected=1, do_norm_correction=0, do_map=0, precision=0)
        eng.reset_sums()
        rng = np.random.default_rng(seed)
        euler = np.column_stack([star["rlnAngleRot"], star["rlnAngleTilt"], star["rlnAnglePsi"]]).astype(float)
        shift = np.column_stack([star["rlnOriginX"], star["rlnOriginY"]]).astype(float)
        if angular_error > 0:
            euler = euler + rng.normal(0, angular_error, euler.shape)
        if shift_error > 0:
            shift = shift + rng.normal(0, shift_error, shift.shape)
        for i0 in range(0, n, chunk):
            i1 = min(n, i0 + chunk)
            sub = {"rlnImageName": star["rlnImageName"][i0:i1]}
            imgs = starfile.stack_images(sub, base)
            ctf = None
            if do_ctf:
                ctf = np.column_stack([star["rlnDefocusU"][i0:i1], star["rlnDefocusV"][i0:i1],
                                       star["rlnDefocusAngle"][i0:i1], star["rlnVoltage"][i0:i1],
                                       star["rlnSphericalAberration"][i0:i1], star["rlnAmplitudeContrast"][i0:i1],
                                       star.get("rlnBfactor", np.zeros(n))[i0:i1]]).astype(float)
            prep = eng.prepare_images(imgs, np.zeros(i1 - i0), np.zeros(i1 - i0), ctf=ctf, pixel_size=angpix,
                                      ctf_phase_flipped=ctf_phase_flipped, only_flip_phases=only_flip_phases,
                                      intact_ctf_first_peak=ctf_intact_first_peak)
            eng.backproject_images(0, prep, euler[i0:i1], shift[i0:i1])
        vol = eng.reconstruct(0, np.zeros(ori // 2 + 1), max_iter_preweight=nr_iters, do_map=False, minres_map=-1,
                              is_whole_instead_of_half=True, normalise=1.0)[0]
    finally:
        if engine is None:
            eng.close()
    if out_path:
        starfile.write_mrc(out_path if out_path.endswith(".mrc") else out_path + ".mrc", vol, angpix)
    return vol


def main(argv=None):
    ap = argparse.ArgumentParser(description="reconstruct a map from particles with known poses (GPU)")
    ap.add_argument("--i", required=True, help="input STAR file")
    ap.add_argument("--o", required=True, help="output map (.mrc)")
    ap.add_argument("--angpix", type=float, required=True)
    ap.add_argument("--maxres", type=float, default=-1.0)
    ap.add_argument("--ctf", action="store_true")
    ap.add_argument("--ctf_phase_flipped", action="store_true")
    ap.add_argument("--only_flip_phases", action="store_true")
    ap.add_argument("--ctf_intact_first_peak", action="store_true")
    ap.add_argument("--iter", type=int, default=10)
    ap.add_argument("--angular_error", type=float, default=0.0)
    ap.add_argument("--shift_error", type=float, default=0.0)
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    reconstruct_star(a.i, a.o, a.angpix, a.maxres, a.ctf, a.ctf_phase_flipped, a.only_flip_phases,
                     a.ctf_intact_first_peak, a.iter, a.angular_error, a.shift_error, device=a.device)
    print(f"wrote {a.o}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
