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


# ---------------------------------------------------------------------------- reconstruction (M-step)
# MyBackProjector::reconstruct, src/reconstruct.cpp:1929-2522, with enforceHermitianSymmetry
# (:1501-1537), decenter (src/reconstruct.h:281-372), convoluteBlobRealSpace (:1721-1788),
# windowToOridimRealSpace (:1790-1927), windowFourierTransform (src/reconstruct.h:471-533), window
# (:184-206), softMaskOutsideMap (:211-268), griddingCorrect (:61-103), TabFtBlob
# (src/reconstruct.h:604-636) and kaiser_Fourier_value (src/bessel.cpp:26-50).  C1 symmetry
# (symmetrise is the identity), do_map as given, update_tau2_with_fsc optional.

def xmipp_origin(n):
    return -int(n / 2.0)


def tab_ftblob(radius=3.8, alpha=15.0, order=0, nr_elem=10000):
    """Tabulated Kaiser-Bessel blob transform (TabFtBlob::init); returns (table, sampling)."""
    assert order == 0
    sampling = 0.5 / float(nr_elem)
    w = np.arange(nr_elem) * sampling
    x = 2.0 * math.pi * radius * w
    out = np.empty(nr_elem)
    bi0 = _bessi0(alpha)
    for i in range(nr_elem):
        sig = math.sqrt(abs(alpha * alpha - x[i] * x[i]))
        if x[i] > alpha:  # J_{3/2}
            jb = math.sqrt(2.0 / (math.pi * sig)) * (math.sin(sig) / sig - math.cos(sig))
            out[i] = (2.0 * math.pi) ** 1.5 * radius ** 3 * jb / (bi0 * sig ** 1.5)
        else:             # I_{3/2}
            ib = 0.0 if sig == 0 else math.sqrt(2.0 / (math.pi * sig)) * (math.cosh(sig) - math.sinh(sig) / sig)
            out[i] = (2.0 * math.pi) ** 1.5 * radius ** 3 * ib / (bi0 * sig ** 1.5)
    return out, sampling


def _bessi0(x):
    """Modified Bessel function I0 by its power series.  Its value only scales the table, and every use
    of the table divides by the table's value at 0, so any accurate I0 gives the same ratios."""
    t, k, s = 1.0, 0, 1.0
    while t > 1e-17 * s:
        k += 1
        t *= (x / 2.0) ** 2 / (k * k)
        s += t
    return s


def _tab_lookup(table, sampling, val):
    idx = (np.abs(val) / sampling).astype(np.int64)
    out = np.zeros_like(val, dtype=np.float64)
    ok = idx < len(table)
    out[ok] = table[idx[ok]]
    return out


def _fftw_signed(n_axis, dimx, dim):
    """The reference's FFTW-order signed index: k < dimx ? k : k - dim (dimx = the half axis length)."""
    k = np.arange(n_axis)
    return np.where(k < dimx, k, k - dim)


def reconstruct(data, weight, ori, current_size, max_iter_preweight=10, do_map=True, tau2_fudge=1.0, tau2=None,
                fsc=None, normalise=1.0, update_tau2_with_fsc=False, is_whole_instead_of_half=False,
                minres_map=5, padding_factor=2):
    """data: complex [pad][pad][pad//2+1] (3D) or [pad][pad//2+1] (2D), centred projector layout;
    weight: real, same shape.  Returns (vol_out, tau2, sigma2, data_vs_prior)."""
    data = np.array(data, dtype=np.complex128, copy=True)
    weight = np.array(weight, dtype=np.float64, copy=True)
    is3d = data.ndim == 3
    if not is3d:
        data = data[None]
        weight = weight[None]
    pf = padding_factor
    r_max = min(current_size // 2, ori // 2)
    P = 2 * (pf * r_max + 1) + 1
    PF = P // 2 + 1
    assert data.shape == ((P, P, PF) if is3d else (1, P, PF))
    oriF = ori // 2 + 1
    Dz = P if is3d else 1
    max_r2 = r_max * r_max * pf * pf
    tau2 = np.zeros(oriF) if tau2 is None else np.array(tau2, dtype=np.float64, copy=True)
    # enforceHermitianSymmetry: pairs in the x = 0 plane
    oz, oy = xmipp_origin(Dz), xmipp_origin(P)
    for iz in range(oz, Dz + oz):
        for iy in range(0 if iz < 0 else 1, P + oy):
            zr, yr = iz - oz, iy - oy
            zl, yl = -iz - oz, -iy - oy
            fr = data[zr, yr, 0].real + data[zl, yl, 0].real
            fi = data[zr, yr, 0].imag - data[zl, yl, 0].imag
            data[zr, yr, 0] = fr + 1j * fi
            data[zl, yl, 0] = fr - 1j * fi
            s = weight[zr, yr, 0] + weight[zl, yl, 0]
            weight[zr, yr, 0] = s
            weight[zl, yl, 0] = s
    # decenter: FFTW order, signed indices k < dimx ? k : k - dim; copy where r^2 <= max_r2
    kp = _fftw_signed(Dz, PF, Dz) if is3d else np.zeros(1, int)
    ip = _fftw_signed(P, PF, P)
    jp = np.arange(PF)
    KP, IP, JP = np.meshgrid(kp, ip, jp, indexing="ij")
    R2 = KP * KP + IP * IP + JP * JP

    def decenter(cen):
        out = np.zeros_like(cen)
        sel = R2 <= max_r2
        out[sel] = cen[(KP - oz)[sel], (IP - oy)[sel], JP[sel]]
        return out

    Fweight = decenter(weight)
    osc = float(pf ** 3) if is3d else float(pf ** 2)
    inside = R2 < max_r2
    ires = np.floor(np.sqrt(R2.astype(np.float64)) / pf + 0.5).astype(int)
    sigma2 = np.zeros(oriF)
    counter = np.zeros(oriF)
    np.add.at(sigma2, ires[inside], osc * Fweight[inside])
    np.add.at(counter, ires[inside], 1.0)
    for i in range(oriF):
        if sigma2[i] > 1e-10:
            sigma2[i] = counter[i] / sigma2[i]
        elif sigma2[i] == 0:
            sigma2[i] = 0.0
        else:
            raise ValueError("unexpectedly small, yet non-zero sigma2 value")
    data_vs_prior = np.zeros(oriF)
    if update_tau2_with_fsc:
        for i in range(oriF):
            myfsc = max(0.001, fsc[i])
            if is_whole_instead_of_half:
                myfsc = math.sqrt(2.0 * myfsc / (myfsc + 1.0))
            myfsc = min(0.999, myfsc)
            myssnr = myfsc / (1.0 - myfsc)
            tau2[i] = myssnr * sigma2[i]
            data_vs_prior[i] = myssnr
    if do_map:
        counter = np.zeros(oriF)
        dvp = np.zeros(oriF)
        for flat in np.flatnonzero(inside):
            r = ires.flat[flat]
            invw = Fweight.flat[flat]
            if tau2[r] > 0:
                invtau2 = 1.0 / (osc * tau2_fudge * tau2[r])
            elif tau2[r] == 0:
                with np.errstate(divide="ignore"):
                    invtau2 = np.float64(1.0) / (0.001 * invw)
            else:
                raise ValueError("negative tau2")
            dvp[r] += invw / invtau2
            counter[r] += 1.0
            if r >= minres_map:
                Fweight.flat[flat] = invw + invtau2
        if not update_tau2_with_fsc:
            for i in range(oriF):
                if i > r_max:
                    data_vs_prior[i] = 0.0
                elif counter[i] < 0.001:
                    data_vs_prior[i] = 999.0
                else:
                    data_vs_prior[i] = dvp[i] / counter[i]
    Fweight /= normalise
    data /= normalise
    # Fnewweight: 1 inside r^2 < max_r2 (centred), decentred
    ZC, YC, XC = np.meshgrid(np.arange(Dz) + oz, np.arange(P) + oy, np.arange(PF), indexing="ij")
    ones = ((ZC * ZC + YC * YC + XC * XC) < max_r2).astype(np.float64)
    Fnewweight = decenter(ones)
    table, sampling = tab_ftblob(1.9 * pf, 15.0, 0)
    normftblob = _tab_lookup(table, sampling, np.zeros(1))[0]
    padh = P // 2
    kr = np.arange(Dz)
    kr = np.where(kr < padh, kr, kr - P) if is3d else np.zeros(1, int)
    ir = np.arange(P)
    ir = np.where(ir < padh, ir, ir - P)
    KR, IR, JR = np.meshgrid(kr, ir, ir, indexing="ij")
    rval = np.sqrt((KR * KR + IR * IR + JR * JR).astype(np.float64)) / (ori * pf)
    blob = _tab_lookup(table, sampling, rval) / normftblob
    rshape = (P, P, P) if is3d else (1, P, P)
    axes = (0, 1, 2) if is3d else (1, 2)
    nreal = float(np.prod([rshape[a] for a in axes]))
    for _ in range(max_iter_preweight):
        Fconv = Fnewweight * Fweight + 0j
        M = np.fft.irfftn(Fconv, s=[rshape[a] for a in axes], axes=axes) * nreal  # FFTW backward
        M *= blob
        Fconv = np.fft.rfftn(M, axes=axes) / nreal  # normalised forward
        w = np.maximum(1e-6, np.abs(Fconv))
        Fnewweight = np.where(inside, Fnewweight / w, Fnewweight)
    Fconv = decenter(data) * Fnewweight
    # windowToOridimRealSpace
    N = pf * ori
    NF = N // 2 + 1
    NZ = N if is3d else 1
    kk = np.arange(NZ)
    kk = np.where(kk < NF, kk, kk - NZ) if is3d else np.zeros(1, int)
    ii = np.arange(N)
    ii = np.where(ii < NF, ii, ii - N)
    if NF < PF:  # shrinking branch (src/reconstruct.h:518-531)
        k2 = np.where(kk < 0, kk + Dz, kk)
        i2 = np.where(ii < 0, ii + P, ii)
        Ftmp = Fconv[np.ix_(k2, i2, np.arange(NF))]
    else:        # growing branch (:498-517): entries of Fconv with r^2 <= (PF - 1)^2
        Ftmp = np.zeros((NZ, N, NF), dtype=np.complex128)
        sel = (KP * KP + IP * IP + JP * JP) <= (PF - 1) ** 2
        kd = np.where(KP < 0, KP + NZ, KP)
        idd = np.where(IP < 0, IP + N, IP)
        Ftmp[kd[sel], idd[sel], JP[sel]] = Fconv[sel]
    oshape = (N, N, N) if is3d else (1, N, N)
    Mout = np.fft.irfftn(Ftmp, s=[oshape[a] for a in axes], axes=axes) * float(np.prod([oshape[a] for a in axes]))
    Mc = np.roll(Mout, shift=N // 2, axis=axes)  # centerFFT forward
    off = [xmipp_origin(o) - xmipp_origin(i) for o, i in zip((ori if is3d else 1, ori, ori), oshape)]
    vol = Mc[off[0]:off[0] + (ori if is3d else 1), off[1]:off[1] + ori, off[2]:off[2] + ori].copy()
    vol /= float(pf ** 3 * ori) if is3d else float(pf ** 2)
    vol = _soft_mask_outside(vol, is3d)
    vol = gridding_correct(vol if is3d else vol[0], pf)
    if update_tau2_with_fsc:
        F = np.fft.rfftn(vol) / vol.size
        shp = F.shape
        fk = [np.where(np.arange(n) < shp[-1], np.arange(n), np.arange(n) - n) for n in shp[:-1]] + [np.arange(shp[-1])]
        G = np.meshgrid(*fk, indexing="ij")
        idx = np.floor(np.sqrt(sum(g.astype(np.float64) ** 2 for g in G)) + 0.5).astype(int)
        spec = np.zeros(vol.shape[-1])
        cnt = np.zeros(vol.shape[-1])
        np.add.at(spec, idx.ravel(), (F.real ** 2 + F.imag ** 2).ravel())
        np.add.at(cnt, idx.ravel(), 1.0)
        spec = spec[:oriF] / cnt[:oriF]
        spec *= (ori * ori if is3d else 1.0) / 2.0
        tau2 = tau2_fudge * spec
    return vol, tau2, sigma2, data_vs_prior


def _soft_mask_outside(vol, is3d, cosine_width=3.0):
    """softMaskOutsideMap with the average background (src/reconstruct.h:211-268)."""
    dz, dy, dx = vol.shape
    radius = dy / 2.0
    radius_p = radius + cosine_width
    Z, Y, X = np.meshgrid(np.arange(dz) + xmipp_origin(dz), np.arange(dy) + xmipp_origin(dy),
                          np.arange(dx) + xmipp_origin(dx), indexing="ij")
    r = np.sqrt((Z * Z + Y * Y + X * X).astype(np.float64))
    outer = r > radius_p
    edge = (r >= radius) & ~outer
    rc = 0.5 + 0.5 * np.cos(math.pi * (radius_p - r[edge]) / cosine_width)
    s = float(outer.sum()) + rc.sum()
    bg = (vol[outer].sum() + (rc * vol[edge]).sum()) / s
    out = vol.copy()
    out[outer] = bg
    out[edge] = (1.0 - rc) * vol[edge] + rc * bg
    return out
