"""FP64 restatement of ROME's M-step stages (TEST INFRASTRUCTURE: only tests/, the smoke check and
bench.py's CPU-baseline leg may use it).

* maximization_other: maximizationOtherParameters, src/map3d_optimizer.cpp:2620-2798 (3D) and
  src/map2d_optimizer.cpp:1985-2045 (2D), outside CC mode and without scale correction.
* npix_per_shell: MAPModel::updateImageSizeAndResolutionPointers, src/map_model.cpp:522-538.
* fourier_map: MyProjector::computeFourierTransformMap, src/reconstruct.cpp:106-223, with
  griddingCorrect src/reconstruct.cpp:61-103 and centerFFT3D/2D src/reconstruct.h:103-180; the FFT is
  numpy's (pocketfft) in place of FFTW/MKL, normalised by the number of real samples as
  FourierTransformerBase::normalizeComplex does (src/fft_base.cpp:103-109).
"""
import math

import numpy as np


def npix_per_shell(ori_size):
    """Pixels of the ori_size half-plane per rounded radius (C round: half away from zero), the x = 0
    column counted for ip >= 0 only (src/map_model.cpp:527-537)."""
    F = ori_size // 2 + 1
    n = np.zeros(F)
    for i in range(ori_size):
        ip = i if i < F else i - ori_size
        for jp in range(F):
            r = math.sqrt(ip * ip + jp * jp)
            ires = int(math.floor(r + 0.5))  # r >= 0: round() half away from zero
            if ires < F and not (jp == 0 and ip < 0):
                n[ires] += 1
    return n


def maximization_other(ref_dim, ori_size, sums, nr_dir, do_norm_correction, avg_norm_in, prior_offx_in,
                       prior_offy_in):
    """sums: dict with wsum_pdf_class [K], wsum_pdf_direction [K][D], wsum_sigma2_noise [G][F],
    wsum_sumw_group [G], wsum_prior_offx/y [K], wsum_scalars [4] (sigma2_offset, LL, ave_Pmax,
    avg_norm_correction).  Returns the updated model (src/map3d_optimizer.cpp:2622-2735)."""
    wpc = np.asarray(sums["wsum_pdf_class"], float)
    K = len(wpc)
    sum_weight = 0.0
    for k in range(K):  # same summation order as the reference (:2623-2625)
        sum_weight += wpc[k]
    sc = np.asarray(sums["wsum_scalars"], float)
    avg_norm = sc[3] / sum_weight if do_norm_correction else avg_norm_in
    pdf_class = wpc / sum_weight
    pdf_direction = np.asarray(sums["wsum_pdf_direction"], float).reshape(K, nr_dir) / sum_weight
    px = np.array(prior_offx_in, float).copy()
    py = np.array(prior_offy_in, float).copy()
    if ref_dim == 2:  # src/map2d_optimizer.cpp:2004-2011
        for k in range(K):
            if wpc[k] > 0:
                px[k] = sums["wsum_prior_offx"][k] / wpc[k]
                py[k] = sums["wsum_prior_offy"][k] / wpc[k]
            else:
                px[k] = py[k] = 0.0
    npix = npix_per_shell(ori_size)
    ws = np.asarray(sums["wsum_sigma2_noise"], float)
    wg = np.asarray(sums["wsum_sumw_group"], float)
    sigma2 = ws / (2.0 * wg[:, None] * npix[None, :])
    return dict(pdf_class=pdf_class, pdf_direction=pdf_direction, prior_offx=px, prior_offy=py,
                sigma2_noise=sigma2,
                scalars=np.array([sc[0] / (2.0 * sum_weight), avg_norm, sc[1], sc[2] / sum_weight]))


def gridding_correct(vol, padding_factor=2):
    """Divide the real-space map by sinc^2 of the trilinear interpolator, r > 0 only
    (src/reconstruct.cpp:61-103); returns a corrected copy."""
    v = np.array(vol, dtype=np.float64, copy=True)
    dims = v.shape
    shifts = [-int(d / 2.0) for d in dims]
    grids = np.meshgrid(*[np.arange(d) + s for d, s in zip(dims, shifts)], indexing="ij")
    r = np.sqrt(sum(g.astype(np.float64) ** 2 for g in grids))
    nz = r > 0
    rval = r[nz] / (dims[-1] * padding_factor)
    sinc = np.sin(math.pi * rval) / (math.pi * rval)
    v[nz] /= sinc * sinc
    return v


def fourier_map(vol, current_size, do_gridding=True, padding_factor=2):
    """Projector grid and power spectrum of a real map (2D [ori][ori] or 3D [ori][ori][ori]).

    Returns (grid, power_spectrum, corrected_map): grid is complex128 [pad][pad][pad//2+1] (3D) or
    [pad][pad//2+1] (2D) with pad = 2*(padding_factor*r_max+1)+1, r_max = min(current, ori)//2."""
    vol = np.asarray(vol, dtype=np.float64)
    ref_dim = vol.ndim
    ori = vol.shape[-1]
    ori_F = ori // 2 + 1
    r_max = min(current_size // 2, ori // 2)
    pad = 2 * (padding_factor * r_max + 1) + 1
    padF = pad // 2 + 1
    padori = padding_factor * ori
    padoriF = padori // 2 + 1
    normfft = float(padding_factor ** 3 * ori) if ref_dim == 3 else float(padding_factor ** 2)
    ref = gridding_correct(vol, padding_factor) if do_gridding else vol.copy()
    # pad, centred: index + ori_shift - padori_shift  (src/reconstruct.cpp:153-164)
    off = -int(ori / 2.0) + int(padori / 2.0)
    Mpad = np.zeros((padori,) * ref_dim)
    sl = tuple(slice(off, off + ori) for _ in range(ref_dim))
    Mpad[sl] = ref
    # centerFFT forward: out[(i + dim/2) % dim] = in[i] along every axis
    Mc = np.roll(Mpad, shift=padori // 2, axis=tuple(range(ref_dim)))
    Faux = np.fft.rfftn(Mc) / Mc.size  # e^{-i}, normalised by the real size
    grid = np.zeros((pad, pad, padF) if ref_dim == 3 else (pad, padF), dtype=np.complex128)
    power = np.zeros(ori_F)
    counter = np.zeros(ori_F)
    max_r2 = r_max * r_max * padding_factor * padding_factor
    shift = -int(pad / 2.0)
    if ref_dim == 3:
        kk, ii, jj = np.meshgrid(np.arange(padori), np.arange(padori), np.arange(padoriF), indexing="ij")
        kp = np.where(kk < padoriF, kk, kk - padori)
    else:
        ii, jj = np.meshgrid(np.arange(padori), np.arange(padoriF), indexing="ij")
        kp = np.zeros_like(ii)
    ip = np.where(ii < padoriF, ii, ii - padori)
    jp = jj
    r2 = ip * ip + jp * jp + kp * kp
    sel = r2 <= max_r2
    vals = Faux[sel] * normfft
    if ref_dim == 3:
        grid[kp[sel] - shift, ip[sel] - shift, jp[sel]] = vals
    else:
        grid[ip[sel] - shift, jp[sel]] = vals
    ires = np.floor(np.sqrt(r2[sel].astype(np.float64)) / padding_factor + 0.5).astype(int)
    np.add.at(power, ires, (vals.real ** 2 + vals.imag ** 2) / 2.0)
    np.add.at(counter, ires, 1.0)
    power = np.where(counter < 1.0, 0.0, power / np.maximum(counter, 1.0))
    return grid, power, ref
