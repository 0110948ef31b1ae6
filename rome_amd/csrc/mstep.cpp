// M-step model parameters other than the maps: maximizationOtherParameters of ROME
// (src/map3d_optimizer.cpp:2620-2798, 2D src/map2d_optimizer.cpp:1985-2045), host FP64, for
// iterations outside CC mode and without scale correction.
#include <cmath>
#include <vector>

#include "../../include/rome_gpu.h"

namespace rome {

// Npix_per_shell of MAPModel::updateImageSizeAndResolutionPointers (src/map_model.cpp:522-538): pixels
// of the ori_size half-plane per rounded radius, the x = 0 column counted once (ip >= 0).
std::vector<double> npix_per_shell(int ori_size) {
    const int F = ori_size / 2 + 1;
    std::vector<double> n(F, 0.);
    for (int i = 0; i < ori_size; ++i)
        for (int j = 0; j < F; ++j) {
            const int ip = i < F ? i : i - ori_size, jp = j;
            const int ires = (int)std::round(std::sqrt((double)(ip * ip + jp * jp)));
            if (ires < F && !(jp == 0 && ip < 0)) n[ires] += 1.;
        }
    return n;
}

}  // namespace rome

extern "C" int rome_maximization_other(int ref_dim, int ori_size, int nr_classes, int nr_groups, int nr_dir,
                                       int do_norm_correction, double avg_norm_in, const double* prior_offx_in,
                                       const double* prior_offy_in, const rome_sums* s, rome_model* out) {
    if (!s || !out || (ref_dim != 2 && ref_dim != 3) || ori_size <= 0 || nr_classes <= 0 || nr_groups <= 0 ||
        nr_dir <= 0)
        return ROME_ERR_ARG;
    if (!s->wsum_pdf_class || !s->wsum_pdf_direction || !s->wsum_sigma2_noise || !s->wsum_sumw_group ||
        !s->wsum_scalars || (ref_dim == 2 && (!s->wsum_prior_offx || !s->wsum_prior_offy)))
        return ROME_ERR_ARG;
    const int K = nr_classes, G = nr_groups, D = nr_dir, F = ori_size / 2 + 1;
    double sum_weight = 0.;
    for (int k = 0; k < K; ++k) sum_weight += s->wsum_pdf_class[k];
    if (!(sum_weight > 0.)) return ROME_ERR_NUMERIC;
    // wsum_scalars: sigma2_offset, LL, ave_Pmax, avg_norm_correction
    const double w_off = s->wsum_scalars[0], w_ll = s->wsum_scalars[1], w_pmax = s->wsum_scalars[2],
                 w_norm = s->wsum_scalars[3];
    const double avg_norm = do_norm_correction ? w_norm / sum_weight : avg_norm_in;
    for (int k = 0; k < K; ++k) {
        if (out->pdf_class) out->pdf_class[k] = s->wsum_pdf_class[k] / sum_weight;
        double px = prior_offx_in ? prior_offx_in[k] : 0., py = prior_offy_in ? prior_offy_in[k] : 0.;
        if (ref_dim == 2) {
            if (s->wsum_pdf_class[k] > 0.) {
                px = s->wsum_prior_offx[k] / s->wsum_pdf_class[k];
                py = s->wsum_prior_offy[k] / s->wsum_pdf_class[k];
            } else {
                px = py = 0.;
            }
        }
        if (out->prior_offx) out->prior_offx[k] = px;
        if (out->prior_offy) out->prior_offy[k] = py;
        if (out->pdf_direction)
            for (int d = 0; d < D; ++d) out->pdf_direction[(size_t)k * D + d] = s->wsum_pdf_direction[(size_t)k * D + d] / sum_weight;
    }
    if (out->sigma2_noise) {
        const std::vector<double> npix = rome::npix_per_shell(ori_size);
        for (int g = 0; g < G; ++g)
            for (int n = 0; n < F; ++n)
                out->sigma2_noise[(size_t)g * F + n] =
                    s->wsum_sigma2_noise[(size_t)g * F + n] / (2. * s->wsum_sumw_group[g] * npix[n]);
    }
    if (out->scalars) {
        out->scalars[0] = w_off / (2. * sum_weight);
        out->scalars[1] = avg_norm;
        out->scalars[2] = w_ll;
        out->scalars[3] = w_pmax / sum_weight;
    }
    return ROME_OK;
}
