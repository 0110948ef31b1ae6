// Thin C-ABI driver over the reference's own vendored HEALPix sources
// (Healpix_2.15a/healpix_base.cc), compiled in place from the reference tree
// by oracle/Makefile.  Used only by tests to pin the oracle's restatement of
// the NEST pixel geometry that src/sampling.cpp:207-221 and :552-579 rely on.
#include "healpix_base.h"

extern "C" {
int ref_healpix_npix(int order) { return Healpix_Base(order, NEST).Npix(); }
int ref_healpix_pix2ang_z_phi(int order, double* z, double* phi) {
    Healpix_Base b(order, NEST);
    for (int p = 0; p < b.Npix(); ++p) b.pix2ang_z_phi(p, z[p], phi[p]);
    return b.Npix();
}
int ref_healpix_nest2xyf(int order, int pix, int* ix, int* iy, int* face) {
    Healpix_Base b(order, NEST);
    b.nest2xyf(pix, *ix, *iy, *face);
    return 0;
}
int ref_healpix_xyf2nest(int order, int ix, int iy, int face) {
    Healpix_Base b(order, NEST);
    return b.xyf2nest(ix, iy, face);
}
double ref_healpix_max_pixrad(int order) { return Healpix_Base(order, NEST).max_pixrad(); }
}
