#!/usr/bin/env python3
"""E-step throughput of the MI355X engine on the plain MAP3D workload.

Workload (one GPU): N synthetic particles 128x128 (pixel 1.5 A), K=4 reference volumes 128^3
(200 Gaussian blobs, sigma 2-6 px, inside a 50 px sphere), uniform SO(3) poses, shifts +-5 px,
CTF 300 kV / Cs 2.7 / Q0 0.1 / defocus 1-3 um, SNR 0.05.  Search: HEALPix order 2 x 24 psi
(4608 rotations) x 81 translations (disc of radius 5, step 1, as the reference enumerates them),
adaptive oversampling 1 (fine pass: 8 rotations x 4 translations per significant coarse sample),
current_size 128, coarse_size 24, backprojection into 4 grids of 259x259x130.

A step = one E-step (coarse pass, fine pass, weights, significance, sigma2/norm sums,
backprojection, per-image sums) over one batch of particles already resident in HBM.

    python bench.py [--steps K] [--warmup W] [--batch B] [--n-images N]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_MFMA_PEAK_TFLOPS = 157.3   # MI355X dense FP32 matrix
HBM_PEAK_GBS = 8000.0


def build_problem(args):
    import rome_amd
    from rome_amd import synth
    rng = np.random.default_rng(args.seed)
    ori, cur, coarse, K, order = 128, 128, 24, 4, 2
    t0 = time.time()
    grids = []
    pad = None
    for _ in range(K):
        vol = synth.blob_volume_fast(ori, 200, (2.0, 6.0), 50.0, rng)
        g, pad = synth.fourier_grid(vol, cur)
        grids.append(g)
    grids = np.stack(grids)
    pose_c = rome_amd.pose_table3d(order, 0, args.perturbation)
    pose_f = rome_amd.pose_table3d(order, 1, args.perturbation)
    tx, ty = rome_amd.translations(5.0, 1.0)
    scx, scy = rome_amd.shift_table(5.0, 1.0, 0, args.perturbation)
    sfx, sfy = rome_amd.shift_table(5.0, 1.0, 1, args.perturbation)
    D, PSI, T = pose_c.shape[0], pose_c.shape[1], len(tx)
    prob = dict(is_3d=1, ori_size=ori, current_size=cur, coarse_size=coarse, nr_classes=K, nr_groups=1,
                adaptive_oversampling=1, nr_dir=D, nr_psi=PSI, nr_trans=T, adaptive_fraction=0.999,
                sigma2_fudge=1.0, sigma2_offset=10.0, avg_norm_correction=1.0, do_ctf_correction=1,
                refs_are_ctf_corrected=1, do_norm_correction=1, do_map=1, pad=pad,
                pdf_class=np.full(K, 1.0 / K), pdf_direction=np.full((K, D), 1.0 / D), prior_offx=np.zeros(K),
                prior_offy=np.zeros(K), grids=grids, pose_coarse=pose_c, pose_fine=pose_f,
                shift_coarse_x=scx[:, 0], shift_coarse_y=scy[:, 0], shift_fine_x=sfx, shift_fine_y=sfy,
                trans_x=tx, trans_y=ty)
    return prob, time.time() - t0


def cpu_baseline(prob, F, ctf, idx, threads):
    """The FP64 oracle (CPU restatement of the reference loops) on a bounded sample of the same workload."""
    import oracle as O
    Fh = F[idx].cpu().numpy().astype(np.complex128)
    ch = ctf[idx].cpu().numpy().astype(np.float64)
    n = len(idx)
    oriF = prob["ori_size"] // 2 + 1
    batch = dict(n=n, fmask_re=Fh.real.copy(), fmask_im=Fh.imag.copy(), fnomask_re=Fh.real.copy(),
                 fnomask_im=Fh.imag.copy(), fctf=ch, highres_Xi2=np.zeros(n), power_img=np.zeros((n, oriF)),
                 old_offx=np.zeros(n), old_offy=np.zeros(n), group=np.zeros(n, np.int32), norm=np.ones(n))
    t0 = time.perf_counter()
    O.estep(prob, batch, nr_threads=threads)
    dt = time.perf_counter() - t0
    return n / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=2000)
    ap.add_argument("--n-images", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=20260101)
    ap.add_argument("--perturbation", type=float, default=0.25)
    ap.add_argument("--cpu-sample", type=int, default=0, help="images timed on the CPU oracle (0: 2 per thread)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (multi-GPU is design-only)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import rome_amd
    from rome_amd import synth

    def log(*a):
        if args.verbose:
            print(*a, file=sys.stderr, flush=True)

    prob, t_gen = build_problem(args)
    log(f"volumes + grids {t_gen:.1f}s")
    eng = rome_amd.Engine(0)
    synth.load_engine_tables(eng, prob)
    n_img = args.n_images
    F, ctf, truth, sigma2 = synth.make_particles_device(eng, prob["nr_classes"], n_img, 128, 128, 1.5, 5.0, 0.05,
                                                        args.seed + 1)
    prob["sigma2_noise"] = np.full((1, prob["ori_size"] // 2 + 1), sigma2)
    eng.set_model(prob["sigma2_noise"], prob["pdf_class"], prob["pdf_direction"], prob["prior_offx"],
                  prob["prior_offy"])
    eng.reset_sums()
    dev = F.device
    oriF = prob["ori_size"] // 2 + 1
    hxi2 = torch.zeros(n_img, dtype=torch.float64, device=dev)
    power = torch.zeros((n_img, oriF), dtype=torch.float64, device=dev)
    offx = torch.zeros(n_img, dtype=torch.float64, device=dev)
    offy = torch.zeros(n_img, dtype=torch.float64, device=dev)
    norm = torch.ones(n_img, dtype=torch.float64, device=dev)
    group = torch.zeros(n_img, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    B = min(args.batch, n_img)
    nb = max(1, n_img // B)
    F2 = F.shape[1]

    def step(i):
        b = i % nb
        o = b * B
        return eng.expectation_device(
            B, F.data_ptr() + o * F2 * 8, F.data_ptr() + o * F2 * 8, ctf.data_ptr() + o * F2 * 4,
            hxi2.data_ptr() + o * 8, power.data_ptr() + o * oriF * 8, offx.data_ptr() + o * 8,
            offy.data_ptr() + o * 8, norm.data_ptr() + o * 8, group.data_ptr() + o * 4, results=True)

    for i in range(args.warmup):
        step(i)
        log(f"warmup {i} done", eng.batch_stats())
    eng.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    agree = []
    for i in range(args.steps):
        r = step(args.warmup + i)
        b = (args.warmup + i) % nb
        tr = truth[b * B:(b + 1) * B]
        agree.append(float(np.mean(r["meta"][:, 1] - 1 == tr[:, 0])))
        log(f"step {i} done", eng.batch_stats())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prof = eng.profile()
    stats = eng.batch_stats()
    value = args.steps * B / dt
    # roofline of the dominant kernel
    dom = max(prof, key=lambda k: prof[k][0])
    M, T, Pc = stats["M"], stats["T"], stats["Pc"]
    gemm_ms, gemm_calls = prof["coarse_cross"]
    gemm_flops = 2.0 * M * (B * T) * (2 * Pc)  # algorithmic: Re<CF, XS> = 2 real MACs per complex pixel
    gemm_tflops = gemm_flops * gemm_calls / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    roofline = {"kernel": "k_coarse_gemm (coarse cross term, FP32 MFMA)", "bound": "mfma",
                "achieved": round(gemm_tflops, 2), "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(gemm_tflops / FP32_MFMA_PEAK_TFLOPS, 4), "traffic": None,
                "avg_launch_ms": round(gemm_ms / max(gemm_calls, 1), 3), "dominant_stage": dom}
    cpu = None
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        ns = args.cpu_sample or 2 * threads
        ips, cdt = cpu_baseline(prob, F, ctf, np.arange(ns), threads)
        cpu = {"value": round(ips, 3), "unit": "images/s", "cores": threads, "kind": "port",
               "sample": f"{ns} particles of the same workload through the FP64 oracle E-step ({cdt:.1f}s)"}
    out = {
        "metric": "E-step images/sec (3D 128^2, K=4) at >=99.5% pose agreement; mixed vs FP32",
        "value": round(value, 2), "unit": "images/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "plain MAP3D E-step: 128^2, K=4, hp2 x 24 psi (4608) x 81 shifts, oversampling 1",
                   "batch_images": B, "n_images": n_img, "current_size": 128, "coarse_size": 24,
                   "perturbation": args.perturbation, "precision_arm": "plain-fp32"},
        "roofline": roofline, "cpu_baseline": cpu,
        "stages_ms_per_step": {k: round(v[0] / max(v[1], 1), 3) for k, v in prof.items()},
        "batch_stats": stats, "class_agreement_vs_truth": round(float(np.mean(agree)), 4),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
