// Regular-grid basis of a mesh: Raster_Pixel (raster.cc:18-100,
// raster_px.cc:10-54), device resident.  The intermediate grid gdim carries
// the fields (map B from the mesh, NodeMapMatrix back); the basis grid bdim
// carries the solution.  With bdim == gdim the two coincide (G = NULL in the
// reference); else G = Grid2LinPixMatrix (mesh.cc:3138-3288) maps grid to
// basis and GI = LinPix2GridMatrix (mesh.cc:3540-3718) basis to grid.
#pragma once
#include "mesh.hpp"

namespace tmi {

struct Raster {
    Mesh *mesh = nullptr;
    int dim = 0;
    int gdim[3] = {1, 1, 1};
    int bdim[3] = {1, 1, 1};
    double bbmin[3] = {0, 0, 0}, bbmax[3] = {0, 0, 0};
    double dW[3] = {0, 0, 0};          // grid spacing (bbmax - bbmin) / (g - 1)
    long glen = 0, blen = 0;
    bool has_g = false;                // bdim != gdim
    int gslen = 0;                     // grid points inside the mesh
    int slen = 0;                      // solution voxels (basis voxels with support)
    DBuf<int> elref;                   // glen: element holding each grid point, -1 outside
    DBuf<int> bcols;                   // glen x nn: mesh nodes of the grid map B (-1 outside)
    DBuf<double> bw;                   // glen x nn: their shape-function weights
    DBuf<int> gsol;                    // gslen: grid points inside the mesh, ascending (grid-stage voxels)
    DBuf<int> sol2basis;               // slen: basis index of each solution voxel, ascending
    DBuf<int> g_rp, g_ci, g_cs;        // G (blen x glen CSR): columns as grid indices / positions in gsol
    DBuf<double> g_val;
    DBuf<int> gi_rp, gi_ci;            // GI (glen x blen CSR), built on first use
    DBuf<double> gi_val;
    bool gi_built = false;
    DBuf<double> fgv;                  // Jacobian scratch: fields on the grid
    DBuf<double> fgs;                  // Jacobian scratch: fields + gradients on the grid voxels inside the mesh
    DBuf<int> ncols;                   // nlen x 2^dim: grid points of NodeMapMatrix (-1 unused)
    DBuf<double> nw;                   // nlen x 2^dim: their interpolation weights
    int ncols_nodemap = 0;             // 2^dim once built

    void build(hipStream_t st);        // elref, B, grid voxels, G and the solution voxels
    void build_nodemap(hipStream_t st);   // NodeMapMatrix (on first use)
    void build_gimap(hipStream_t st);     // LinPix2GridMatrix (on first use)
    double voxel_size() const;            // prod gsize / (bdim - 1) (jacobian.cc:368-372)
};

// Map_GridToBasis / Map_BasisToGrid (raster_px.cc:102-130): G or GI applied
// to ncols complex columns, [glen][ncols] <-> [blen][ncols] (a copy when
// bdim == gdim)
void launch_grid_to_basis(Raster &r, int ncols, const double *d_g, double *d_b, hipStream_t st);
void launch_basis_to_grid(Raster &r, int ncols, const double *d_b, double *d_g, hipStream_t st);

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
// d_lk: nq x nm link index of each (source, detector) pair, -1 if not linked
void launch_jacobian_grid(Raster &r, int nq, int nm, const double *d_dphi, int ldd, const double *d_aphi, int lda,
                          const int *d_lk, int nqm, const double *d_proj, int logscale, double *d_J, hipStream_t st,
                          bool cw = false, double *d_Jkap = nullptr);

}  // namespace tmi
