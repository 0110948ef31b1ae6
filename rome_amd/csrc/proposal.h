// Argument blocks and launchers of the learned pose proposal (proposal.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace rome {

constexpr int kMaxGroupWords = 1024;  // candidate groups (class x coarse direction x in-plane bin) / 32 in LDS

// Training targets from the last pool's fine pass (block lists of the fine pass, as the fine weights use).
struct ProposalTargetArgs {
    int K, C, nbins, R, PSI, NOR, NOT, tcap;
    const long long* img_blk;
    const int* blk_task;
    const int* blk_t;
    const int* task_m;
    const double* D2;
    const double* dmin;
    const double* sumw;
    const double* pdf_m;
    const double *trans_x, *trans_y, *old_offx, *old_offy, *prior_x, *prior_y;
    double sigma2_offset;
    const double *shift_x, *shift_y;  // fine shifts [T][NOT] (pixels)
    float shift_scale;                // shift targets are in units of this (the offset range)
    int row0;                         // training-set row of the pool's first image
    int* tcell;                       // [rows][tcap]
    float* tprob;                     // [rows][tcap]
    int* tcnt;                        // [rows]
    float* tshift;                    // [rows][2]
};

struct CandidateArgs {
    const float* z;  // [n][ldz] network outputs
    int ldz;
    int K, C, D, nbins, PSI, topk;
    float min_mass;
    const int* nbr;   // [D][8] neighbouring coarse directions (-1: none)
    unsigned* gmask;  // [n][gwords] marked groups
    int gwords;
    int* cand_m;  // [n][cap] candidate rotations, ascending
    int cap;
    int* cand_cnt;  // [n]
    float* pmass;   // [n] predicted posterior mass of the marked groups
    int* all;       // [n] 1: predicted mass below min_mass, every rotation marked
};

struct CandCoarseArgs {
    const int* cand_m;
    const int* cand_cnt;
    int cap_m;
    const void *Ahi, *Alo;  // coarse table, hi / lo parts (half table when psih > 0)
    int psih;
    const void *Whi, *Wlo;  // image operand [n*T][K16], hi / lo parts
    const float* inv_scale;
    const float* At;     // [n][Mpad]
    const float* reach;  // [n]
    int K16, Mpad, Mtot, T;
    ListArgs list;  // lpd, ent, val, count, cap (vbound, reach unused)
    int* emin_key;
};

struct AuditArgs {
    const float* E;  // [na][T][Mtot] exhaustive pose term of the audited images
    int img0, T, Mtot, nbins, gwords;
    const unsigned* gmask;
    WeightParams wp;
    double* mass;  // [na]
};

hipError_t mlp_gemm(hipStream_t s, const void* A, int lda, const void* B, int ldb, float* C, int ldc, const float* bias,
                    int M, int N, int K);
hipError_t mlp_transpose(hipStream_t s, const void* in, bool in_f32, int R, int Cn, int ldi, void* out, int ldo);
hipError_t mlp_gelu_fwd(hipStream_t s, const float* a, int lda, void* h, int ldh, int rows, int cols);
hipError_t mlp_gelu_bwd(hipStream_t s, float* dh, const float* a, void* daq, size_t n);
hipError_t mlp_colsum(hipStream_t s, const float* d, int rows, int cols, int ld, float* db);
hipError_t mlp_softmax_ce(hipStream_t s, const float* z, int ldz, int ncell, int B, const int* tcell, const float* tprob,
                          const int* tcnt, int tcap, const float* tshift, float lam, float* dz, void* dzq, int ldd,
                          float* loss);
hipError_t mlp_adam(hipStream_t s, float* w, float* m, float* v, const float* g, void* wq, size_t n, float lr, float b1,
                    float b2, float eps, float bc1, float bc2);
hipError_t mlp_to_bf16(hipStream_t s, const float* in, void* out, size_t n);
hipError_t mlp_gather_rows(hipStream_t s, const void* src, const int* idx, int rows, int ld, void* dst);
hipError_t mlp_gather_targets(hipStream_t s, const int* idx, int rows, int tcap, const int* tcell, const float* tprob,
                              const int* tcnt, const float* tshift, int* ocell, float* oprob, int* ocnt, float* oshift);
hipError_t mlp_features(hipStream_t s, int n, const float2* fmask, int F2f, const int* coef_n, const int* coef_ires,
                        int ncoef, const double* sigma2, const int* group, int oriF, void* x, int ldx);
hipError_t proposal_targets(hipStream_t s, int n, const ProposalTargetArgs& a);
hipError_t select_candidates(hipStream_t s, int n, const CandidateArgs& a);
hipError_t cand_coarse(hipStream_t s, int precision, int n, const CandCoarseArgs& a);
hipError_t audit_mass(hipStream_t s, int na, const AuditArgs& a);

}  // namespace rome
