// Regular-grid basis of a mesh: Raster_Pixel with basis grid == intermediate
// grid (raster.cc:18-100, raster_px.cc:10-54), device resident.
#pragma once
#include "mesh.hpp"

namespace tmi {

struct Raster {
    Mesh *mesh = nullptr;
    int dim = 0;
    int gdim[3] = {1, 1, 1};
    double bbmin[3] = {0, 0, 0}, bbmax[3] = {0, 0, 0};
    double dW[3] = {0, 0, 0};          // grid spacing (bbmax - bbmin) / (g - 1)
    long glen = 0;
    int slen = 0;
    DBuf<int> elref;                   // glen: element holding each grid point, -1 outside
    DBuf<int> bcols;                   // glen x nn: mesh nodes of the grid map B (-1 outside)
    DBuf<double> bw;                   // glen x nn: their shape-function weights
    DBuf<int> sol2basis;               // slen: grid points inside the mesh, ascending
    DBuf<double> fgv;                  // Jacobian scratch: fields on the grid
    DBuf<double> fgs;                  // Jacobian scratch: fields + gradients on solution voxels
    DBuf<int> ncols;                   // nlen x 2^dim: grid points of NodeMapMatrix (-1 unused)
    DBuf<double> nw;                   // nlen x 2^dim: their interpolation weights
    int ncols_nodemap = 0;             // 2^dim once built

    void build(hipStream_t st);        // elref, B, sol2basis
    void build_nodemap(hipStream_t st);   // NodeMapMatrix (on first use)
};

// Map_GridToMesh: ncols complex grid columns [glen][ncols] -> [nlen][ncols]
void launch_grid_to_mesh(Raster &r, int ncols, const double *d_g, double *d_out, hipStream_t st);

// the grid map B applied to ncols node-major complex columns:
// out[g][c] = sum_j w[g][j] f[col[g][j]][c]   (f rows of ld entries)
void launch_raster_map(const Raster &r, int ncols, const double *d_f, int ld, double *d_out, hipStream_t st);
// GenerateJacobian_grid for fields already on the device (node-major
// [n][ldd] / [n][lda] complex); J: (2 nqm) x (2 slen) row-major
// AddDataGradient's grid part: d_dphi [n][ldd] direct and d_aphi [n][lda]
// adjoint complex fields of nq sources; d_grad (2 slen) is decremented by
// the mua and kappa PMDF sums of every source, in source order.
void launch_data_gradient(Raster &r, int nq, const double *d_dphi, int ldd, const double *d_aphi, int lda,
                          double *d_grad, hipStream_t st);
// cw: CW Jacobian of real fields given with zero imaginary parts; d_proj is
// then nqm real projections, d_J the nqm x slen mua block and d_Jkap the
// kappa block (either may be null).
void launch_jacobian_grid(Raster &r, int nq, int nm, const double *d_dphi, int ldd, const double *d_aphi, int lda,
                          const int *d_lptr, const int *d_lm, int nqm, const double *d_proj, int logscale,
                          double *d_J, hipStream_t st, bool cw = false, double *d_Jkap = nullptr);

}  // namespace tmi
