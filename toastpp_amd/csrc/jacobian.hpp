// Jacobian (nodal and raster basis) launchers.
#pragma once
#include "mesh.hpp"

namespace tmi {

// d_dphi [n][ldd], d_aphi [n][lda] complex (node-major, the first nq / nm
// entries of each row used); the links come from the mesh's link table
// m.d_lk (link of source q and detector m at q nm + m, or -1); d_proj: nqm
// complex linear projections; d_Pt: scratch of n*nqm*4 doubles (used only
// by the two-pass variant); d_J: (2 nqm) x (2 n) row-major.  dmax: longest
// CSR row of the mesh.
// cw: CW Jacobian of real fields given with zero imaginary parts; d_proj is
// then nqm real projections, d_J the nqm x n mua block and d_Jkap the
// nqm x n kappa block (either may be null).
// true when the environment sets TMI_JAC_MODE=two: node-major intermediate
// (d_Pt) plus tiled transpose, instead of the node kernel writing J
bool jacobian_two_pass();
void launch_jacobian_nodal(const Mesh &m, int nq, int nm, const double *d_dphi, int ldd, const double *d_aphi, int lda,
                           int nqm, const double *d_proj, int logscale, double *d_Pt, double *d_J, int dmax,
                           hipStream_t st, bool cw = false, double *d_Jkap = nullptr);
// Element-basis Jacobian: d_J (2 nqm) x (2 elen) row-major, or with cw the
// nqm x elen blocks d_J (mua) and d_Jkap (kappa, either may be null) from real
// fields with zero imaginary parts and nqm real projections.  mvec_n > 0: the
// mua columns n .. elen-1 repeat the kappa PMDF of elements 0 .. elen-n-1, as
// the reference's mvec overload lays it out.
void launch_jacobian_el(const Mesh &m, int nq, const double *d_dphi, int ldd, const double *d_aphi, int lda,
                        const int *d_lptr, const int *d_lm, int nqm, const double *d_proj, int logscale, int mvec_n,
                        double *d_J, hipStream_t st, bool cw = false, double *d_Jkap = nullptr);
// real [rows][ld] -> complex [rows][cols] (imaginary parts zero)
void launch_real_to_cplx(long rows, int cols, const double *src, int ld, double *dst, hipStream_t st);

}  // namespace tmi
