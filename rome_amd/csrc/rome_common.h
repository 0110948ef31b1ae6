// Shared host-side declarations of the MI355X E-step engine.
#pragma once
#include <cstddef>
#include <cstdint>

namespace rome {

// ---- sampling.cpp (host pose / shift tables) ----
void healpix_nest2xyf(int order, int pix, int& ix, int& iy, int& face);
int healpix_xyf2nest(int order, int ix, int iy, int face);
void healpix_pix2ang(int order, int pix, double& z, double& phi);
void euler_matrix(double rot, double tilt, double psi, double A[9]);
void euler_angles(const double A[9], double& rot, double& tilt, double& psi);
int pose_table3d(int order, int oversampling, double perturbation, double* out);
int pose_table2d(double psi_step, int oversampling, double perturbation, double* out);
int translations(double range, double step, double* tx, double* ty);
int shift_table(double range, double step, int oversampling, double perturbation, double* ox, double* oy);

// ---- device-side parameter blocks (plain structs, passed by value to kernels) ----
struct PixList {          // one Fourier half-plane pixel set
    int count;            // number of pixels
    const int* n_fine;    // index in the current_size half-plane (input images)
    const int* xy;        // packed (x, y) logical frequency, x >= 0
    const int* ires;      // resolution shell
};

constexpr int kBinCount = 2048;       // significance histogram bins (u - umin in [0, 64) nats)
constexpr double kBinScale = 32.0;    // bins per nat
constexpr int kSortCap = 1024;        // candidates sorted exactly in LDS

}  // namespace rome
