// Raster basis (Raster_Pixel, basis grid == intermediate grid) and the grid
// Jacobian (GenerateJacobian_grid, jacobian.cc:293-373) on the GPU.
//
// Compiled without floating-point contraction: the grid coordinates, local
// coordinates and shape functions follow the reference operation by
// operation (mesh.cc:2486-2554, 2944-3015; tet4.cc:186-196, 268-279,
// 302-313; tri3.cc:134-144, 191-201, 231-241).
#include <hip/hip_runtime.h>

#include <climits>

#include "raster.hpp"

namespace tmi {

struct GridDesc {
    int nx, ny, nz;
    double bx, by, bz;     // bbmin
    double hx, hy, hz;     // spacing (bbmax - bbmin) / (g - 1)
    double tx, ty, tz;     // bbmax
};

__global__ void k_fill_int(int *p, long n, int v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// GenerateElementPixelRef: one thread per element visits the grid points of
// its bounding box (inflated by 1e-6) and claims those it contains; the
// reference keeps the first (lowest-numbered) containing element, which is
// the minimum over all containing elements.
template <int DIM>
__global__ void k_elref(int elen, const double *__restrict__ nodes, const int *__restrict__ el,
                        const double *__restrict__ geo, GridDesc gd, int *__restrict__ elref)
{
    constexpr int NN = DIM + 1;
    constexpr int GS = DIM == 3 ? GEO3 : GEO2;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= elen) return;
    const double EPS = 1e-6, EPSB = 1e-8, EPSL = 1e-8;
    double lo[3] = {1e10, 1e10, 1e10}, hi[3] = {-1e10, -1e10, -1e10};   // element bounding box
    for (int i = 0; i < NN; i++) {
        const double *x = nodes + (long)el[(long)e * NN + i] * DIM;
        for (int d = 0; d < DIM; d++) {
            if (x[d] < lo[d]) lo[d] = x[d];
            if (x[d] > hi[d]) hi[d] = x[d];
        }
    }
    const double bmin[3] = {gd.bx, gd.by, gd.bz}, bmax[3] = {gd.tx, gd.ty, gd.tz}, h[3] = {gd.hx, gd.hy, gd.hz};
    const int g[3] = {gd.nx, gd.ny, gd.nz};
    int rmin[3] = {0, 0, 0}, rmax[3] = {1, 1, 1};
    for (int d = 0; d < DIM; d++) {
        const double l = lo[d] - EPS, u = hi[d] + EPS;
        if (l > bmax[d] || u < bmin[d]) return;
        rmin[d] = max(0, (int)ceil((l - bmin[d]) / h[d]));
        rmax[d] = min(g[d], (int)floor((u - bmin[d]) / h[d]) + 1);
    }
    const double *gg = geo + (long)e * GS;
    const double *a = gg, *b = gg + NN, *c = gg + 2 * NN, *dd = gg + 3 * NN;
    double sum = a[0] + a[1] + a[2];
    if (DIM == 3) sum = sum + a[3];
    const double scale = 1.0 / sum;
    double p[3] = {0, 0, 0};
    for (int iz = rmin[2]; iz < rmax[2]; iz++) {
        if (DIM > 2) p[2] = iz * gd.hz + gd.bz;
        for (int iy = rmin[1]; iy < rmax[1]; iy++) {
            p[1] = iy * gd.hy + gd.by;
            for (int ix = rmin[0]; ix < rmax[0]; ix++) {
                p[0] = ix * gd.hx + gd.bx;
                bool in = true;
                for (int d = 0; d < DIM; d++)
                    if (p[d] < lo[d] - EPSB || p[d] > hi[d] + EPSB) in = false;
                if (!in) continue;
                double l0, l1, l2 = 0.0;
                if (DIM == 3) {
                    l0 = (a[1] + b[1] * p[0] + c[1] * p[1] + dd[1] * p[2]) * scale;
                    l1 = (a[2] + b[2] * p[0] + c[2] * p[1] + dd[2] * p[2]) * scale;
                    l2 = (a[3] + b[3] * p[0] + c[3] * p[1] + dd[3] * p[2]) * scale;
                    in = l0 + EPSL >= 0.0 && l1 + EPSL >= 0.0 && l2 + EPSL >= 0.0 && l0 + l1 + l2 - EPSL <= 1.0;
                } else {
                    l0 = (a[1] + b[1] * p[0] + c[1] * p[1]) * scale;
                    l1 = (a[2] + b[2] * p[0] + c[2] * p[1]) * scale;
                    in = l0 + EPSL >= 0.0 && l1 + EPSL >= 0.0 && l0 + l1 - EPSL <= 1.0;
                }
                if (in) atomicMin(&elref[ix + (long)gd.nx * (iy + (long)gd.ny * iz)], e);
            }
        }
    }
}

__global__ void k_elref_finish(int *elref, long n, int *flag)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = elref[i] == INT_MAX ? -1 : elref[i];
    elref[i] = v;
    flag[i] = v >= 0;
}

__global__ void k_compact(const int *__restrict__ flag, const int *__restrict__ pos, long n, int *__restrict__ out)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) out[pos[i]] = (int)i;
}

// GridMapMatrix rows: global shape functions of the holding element
template <int DIM>
__global__ void k_bmap(long glen, const int *__restrict__ elref, const int *__restrict__ el,
                       const double *__restrict__ geo, GridDesc gd, int *__restrict__ cols, double *__restrict__ w)
{
    constexpr int NN = DIM + 1;
    constexpr int GS = DIM == 3 ? GEO3 : GEO2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= glen) return;
    const int e = elref[i];
    if (e < 0) {
        for (int j = 0; j < NN; j++) { cols[i * NN + j] = -1; w[i * NN + j] = 0.0; }
        return;
    }
    const int ix = (int)(i % gd.nx), iy = (int)((i / gd.nx) % gd.ny), iz = (int)(i / ((long)gd.nx * gd.ny));
    double p[3];
    p[0] = ix * gd.hx + gd.bx;
    p[1] = iy * gd.hy + gd.by;
    p[2] = DIM > 2 ? iz * gd.hz + gd.bz : 0.0;
    const double *gg = geo + (long)e * GS;
    const double *a = gg, *b = gg + NN, *c = gg + 2 * NN, *dd = gg + 3 * NN;
    const double size = DIM == 3 ? gg[16] : gg[9];
    const double scale = DIM == 3 ? 1.0 / (6.0 * size) : 1.0 / (2.0 * size);
    for (int j = 0; j < NN; j++) {
        cols[i * NN + j] = el[(long)e * NN + j];
        w[i * NN + j] = DIM == 3 ? scale * (a[j] + b[j] * p[0] + c[j] * p[1] + dd[j] * p[2])
                                 : scale * (a[j] + b[j] * p[0] + c[j] * p[1]);
    }
}

// NodeMapMatrix (mesh.cc:2695-2853): for every mesh node the bi/trilinear
// weights of the 2^dim grid points around it that lie inside the mesh, in
// the reference's corner order (x fastest), renormalised when some of them
// are outside (equal weights when their sum is below 1e-8); no entries for a
// node outside the bounding box.  Rows of 2^dim slots, unused ones -1.
template <int DIM>
__global__ void k_nodemap(int n, const double *__restrict__ x, GridDesc gd, const int *__restrict__ elref,
                          int *__restrict__ cols, double *__restrict__ w)
{
    constexpr int NC = 1 << DIM;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int *oc = cols + (long)i * NC;
    double *ow = w + (long)i * NC;
    for (int k = 0; k < NC; k++) { oc[k] = -1; ow[k] = 0.0; }
    const double bmin[3] = {gd.bx, gd.by, gd.bz}, bmax[3] = {gd.tx, gd.ty, gd.tz}, dW[3] = {gd.hx, gd.hy, gd.hz};
    const int g[3] = {gd.nx, gd.ny, gd.nz};
    int i0[3] = {0, 0, 0};
    double d0[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < DIM; d++) {
        const double p = x[(long)i * DIM + d];
        if (!(p >= bmin[d] && p <= bmax[d])) return;
        i0[d] = (int)((p - bmin[d]) / dW[d]);
        if (i0[d] < 0) i0[d] = 0;
        else if (i0[d] >= g[d] - 1) i0[d] = g[d] - 2;
        d0[d] = (p - (i0[d] * dW[d] + bmin[d])) / dW[d];
        if (d0[d] < 0.0) d0[d] = 0.0;
        else if (d0[d] > 1.0) d0[d] = 1.0;
    }
    int cnt = 0;
    for (int k = 0; k < NC; k++) {
        const int cx = k & 1, cy = (k >> 1) & 1, cz = (k >> 2) & 1;
        const long pix = DIM == 3 ? (i0[0] + cx) + (long)g[0] * ((i0[1] + cy) + (long)g[1] * (i0[2] + cz))
                                  : (i0[0] + cx) + (long)g[0] * (i0[1] + cy);
        if (elref[pix] < 0) continue;
        double wt = (cx ? d0[0] : (1.0 - d0[0])) * (cy ? d0[1] : (1.0 - d0[1]));
        if (DIM == 3) wt = wt * (cz ? d0[2] : (1.0 - d0[2]));
        oc[cnt] = (int)pix;
        ow[cnt] = wt;
        cnt++;
    }
    if (cnt > 0 && cnt < NC) {
        double weight = 0.0;
        for (int k = 0; k < cnt; k++) weight += ow[k];
        if (weight > 1e-8)
            for (int k = 0; k < cnt; k++) ow[k] /= weight;
        else
            for (int k = 0; k < cnt; k++) ow[k] = 1.0 / (double)cnt;
    }
}

// Map_GridToMesh (raster.cc:212-224, Ax / Ax_cplx of NodeMapMatrix) for
// ncols complex columns: out[i][c] = sum_k w[i][k] g[col[i][k]][c]
template <int NC>
__global__ void k_grid_to_mesh(int n, int ncols, const int *__restrict__ cols, const double *__restrict__ w,
                               const double2 *__restrict__ g, double2 *__restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * ncols) return;
    const long i = t / ncols;
    const int c = (int)(t % ncols);
    double re = 0.0, im = 0.0;
    for (int k = 0; k < NC; k++) {
        const int pix = cols[i * NC + k];
        if (pix < 0) continue;
        const double2 v = g[(long)pix * ncols + c];
        re += w[i * NC + k] * v.x;
        im += w[i * NC + k] * v.y;
    }
    out[i * ncols + c] = make_double2(re, im);
}

void Raster::build_nodemap(hipStream_t st)
{
    if (ncols_nodemap) return;
    Mesh &m = *mesh;
    const int NC = 1 << dim;
    GridDesc gd{gdim[0], gdim[1], gdim[2], bbmin[0], bbmin[1], bbmin[2], dW[0], dW[1], dW[2],
                bbmax[0], bbmax[1], bbmax[2]};
    ncols.alloc((size_t)m.n * NC);
    nw.alloc((size_t)m.n * NC);
    const dim3 g((m.n + 255) / 256);
    if (dim == 3)
        hipLaunchKernelGGL(k_nodemap<3>, g, dim3(256), 0, st, m.n, m.nodes.p, gd, elref.p, ncols.p, nw.p);
    else
        hipLaunchKernelGGL(k_nodemap<2>, g, dim3(256), 0, st, m.n, m.nodes.p, gd, elref.p, ncols.p, nw.p);
    TMI_LAUNCH_CHECK();
    TMI_HIP(hipStreamSynchronize(st));
    ncols_nodemap = NC;
}

void launch_grid_to_mesh(Raster &r, int ncols, const double *d_g, double *d_out, hipStream_t st)
{
    r.build_nodemap(st);
    const long cnt = (long)r.mesh->n * ncols;
    const dim3 g((unsigned)((cnt + 255) / 256));
    if (r.dim == 3)
        hipLaunchKernelGGL(k_grid_to_mesh<8>, g, dim3(256), 0, st, r.mesh->n, ncols, r.ncols.p, r.nw.p,
                           (const double2 *)d_g, (double2 *)d_out);
    else
        hipLaunchKernelGGL(k_grid_to_mesh<4>, g, dim3(256), 0, st, r.mesh->n, ncols, r.ncols.p, r.nw.p,
                           (const double2 *)d_g, (double2 *)d_out);
    TMI_LAUNCH_CHECK();
}


// ------------------------------------------------------------ linear pixel basis
// Basis voxel ix of bx per axis sits at x = ix/(bx-1) of the unit box and
// covers the grid points xx of [floor(x0 (gx-1)), ceil(x1 (gx-1))] (x0, x1
// its neighbours, 0 and 1 at the ends) with d = |xx/(gx-1) - x| (bx-1) < 1;
// its tent weight is prod (1 - d).  Same loops and operation order as
// Grid2LinPixMatrix / LinPix2GridMatrix / SubsampleLinPixel
// (mesh.cc:3047-3718), so the weights are bit-identical to the reference's.

struct LpDims {
    int g[3], b[3];
};

struct LpAxis {
    double x, nb, dg;
    int g0, g1;
};

__device__ inline LpAxis lp_axis(int b, int nbv, int ng)
{
    LpAxis a;
    const double db = 1.0 / (nbv - 1.0);
    a.nb = nbv - 1.0;
    a.dg = 1.0 / (ng - 1.0);
    a.x = (double)b * db;
    const double x0 = (b ? (b - 1) * db : 0.0);
    const double x1 = (b < nbv - 1 ? (b + 1) * db : 1.0);
    a.g0 = (int)floor(x0 * (ng - 1));
    a.g1 = (int)ceil(x1 * (ng - 1));
    // the reference's windows lie inside [0, ng-1]; the clamp only guards memory
    a.g0 = max(a.g0, 0);
    a.g1 = min(a.g1, ng - 1);
    return a;
}

// visit the grid points of basis voxel b in the reference's order: f(g, t)
template <int DIM, class F>
__device__ inline void lp_row(const LpDims &L, const int *__restrict__ elref, long b, F &&f)
{
    const int ix = (int)(b % L.b[0]), iy = (int)((b / L.b[0]) % L.b[1]);
    const LpAxis ax = lp_axis(ix, L.b[0], L.g[0]), ay = lp_axis(iy, L.b[1], L.g[1]);
    if (DIM == 2) {
        for (int yy = ay.g0; yy <= ay.g1; yy++) {
            const double dy = fabs(yy * ay.dg - ay.x) * ay.nb;
            if (dy >= 1.0) continue;
            const double idy = 1.0 - dy;
            for (int xx = ax.g0; xx <= ax.g1; xx++) {
                const long g = xx + (long)yy * L.g[0];
                if (elref[g] < 0) continue;
                const double dx = fabs(xx * ax.dg - ax.x) * ax.nb;
                if (dx >= 1.0) continue;
                f(g, (1.0 - dx) * idy);
            }
        }
    } else {
        const int iz = (int)(b / ((long)L.b[0] * L.b[1]));
        const LpAxis az = lp_axis(iz, L.b[2], L.g[2]);
        for (int zz = az.g0; zz <= az.g1; zz++) {
            const double dz = fabs(zz * az.dg - az.x) * az.nb;
            if (dz >= 1.0) continue;
            const double idz = 1.0 - dz;
            for (int yy = ay.g0; yy <= ay.g1; yy++) {
                const double dy = fabs(yy * ay.dg - ay.x) * ay.nb;
                if (dy >= 1.0) continue;
                const double idyz = (1.0 - dy) * idz;
                for (int xx = ax.g0; xx <= ax.g1; xx++) {
                    const long g = xx + ((long)yy + (long)zz * L.g[1]) * L.g[0];
                    if (elref[g] < 0) continue;
                    const double dx = fabs(xx * ax.dg - ax.x) * ax.nb;
                    if (dx >= 1.0) continue;
                    f(g, (1.0 - dx) * idyz);
                }
            }
        }
    }
}

template <int DIM>
__global__ void k_g_count(long blen, LpDims L, const int *__restrict__ elref, int *__restrict__ cnt)
{
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= blen) return;
    int c = 0;
    lp_row<DIM>(L, elref, b, [&](long, double) { c++; });
    cnt[b] = c;
}

// G row b: tent weights divided by their sum (summed in visiting order),
// columns as grid indices and as positions among the grid points inside
// the mesh; flag[b] = 1 if the row is not empty (the basis support test of
// raster.cc:54-73: SubsampleLinPixel of the inside mask is > 0 exactly where
// an inside grid point has a positive weight)
template <int DIM>
__global__ void k_g_fill(long blen, LpDims L, const int *__restrict__ elref, const int *__restrict__ gpos,
                         const int *__restrict__ rp, int *__restrict__ ci, int *__restrict__ cs,
                         double *__restrict__ val, int *__restrict__ flag)
{
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= blen) return;
    int k = rp[b];
    double wgt = 0.0;
    lp_row<DIM>(L, elref, b, [&](long g, double t) {
        ci[k] = (int)g;
        cs[k] = gpos[g];
        val[k] = t;
        wgt += t;
        k++;
    });
    for (int i = rp[b]; i < k; i++) val[i] = wgt ? val[i] / wgt : 0.0;
    flag[b] = k > rp[b] ? 1 : 0;
}

// basis indices along one axis whose window holds grid index xx with
// d < 1, ascending, and their 1 - d
__device__ inline int lp_cover(int xx, int nbv, int ng, int *ib, double *w)
{
    const double u = xx * (1.0 / (ng - 1.0)) * (nbv - 1.0);
    const int lo = max((int)floor(u) - 1, 0), hi = min((int)floor(u) + 2, nbv - 1);
    int c = 0;
    for (int b = lo; b <= hi; b++) {
        const LpAxis a = lp_axis(b, nbv, ng);
        if (xx < a.g0 || xx > a.g1) continue;
        const double d = fabs(xx * a.dg - a.x) * a.nb;
        if (d >= 1.0) continue;
        ib[c] = b;
        w[c] = 1.0 - d;
        c++;
    }
    return c;
}

// GI row g (LinPix2GridMatrix): the unnormalised tent weights of the basis
// voxels covering grid point g, basis index ascending; with ci == nullptr
// only the count
template <int DIM>
__global__ void k_gi_row(long glen, LpDims L, const int *__restrict__ elref, const int *__restrict__ rp,
                         int *__restrict__ cnt, int *__restrict__ ci, double *__restrict__ val)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= glen) return;
    int n = 0;
    if (elref[g] >= 0) {
        const int xx = (int)(g % L.g[0]), yy = (int)((g / L.g[0]) % L.g[1]);
        const int zz = DIM > 2 ? (int)(g / ((long)L.g[0] * L.g[1])) : 0;
        int bx[4], by[4], bz[4] = {0, 0, 0, 0};
        double wx[4], wy[4], wz[4] = {1.0, 1.0, 1.0, 1.0};
        const int nx = lp_cover(xx, L.b[0], L.g[0], bx, wx);
        const int ny = lp_cover(yy, L.b[1], L.g[1], by, wy);
        const int nz = DIM > 2 ? lp_cover(zz, L.b[2], L.g[2], bz, wz) : 1;
        int k = ci ? rp[g] : 0;
        for (int a = 0; a < nz; a++)
            for (int j = 0; j < ny; j++) {
                const double wyz = DIM > 2 ? wy[j] * wz[a] : wy[j];
                for (int i = 0; i < nx; i++) {
                    if (ci) {
                        ci[k] = bx[i] + (by[j] + bz[a] * L.b[1]) * L.b[0];
                        val[k] = wx[i] * wyz;
                    }
                    k++;
                    n++;
                }
            }
    }
    if (!ci) cnt[g] = n;
}

static LpDims lp_dims(const Raster &r)
{
    LpDims L;
    for (int d = 0; d < 3; d++) {
        L.g[d] = r.gdim[d];
        L.b[d] = r.bdim[d];
    }
    return L;
}

// G and the solution voxels; gpos[g]: position of grid point g among the
// grid points inside the mesh
static void build_gmap(Raster &r, const int *gpos, hipStream_t st)
{
    const LpDims L = lp_dims(r);
    TMI_CHECK(r.blen < INT_MAX, TMI_ERR_UNSUPPORTED, "basis of %ld voxels is too large", r.blen);
    const int blen = (int)r.blen;
    const dim3 gb((unsigned)((blen + 255) / 256));
    DBuf<int> cnt(blen), flag(blen), spos(blen + 1);
    r.g_rp.alloc(blen + 1);
    if (r.dim == 3) hipLaunchKernelGGL(k_g_count<3>, gb, dim3(256), 0, st, r.blen, L, r.elref.p, cnt.p);
    else hipLaunchKernelGGL(k_g_count<2>, gb, dim3(256), 0, st, r.blen, L, r.elref.p, cnt.p);
    TMI_LAUNCH_CHECK();
    launch_exclusive_scan(cnt.p, r.g_rp.p, blen, st);
    int nnz = 0;
    TMI_HIP(hipMemcpyAsync(&nnz, r.g_rp.p + blen, sizeof(int), hipMemcpyDeviceToHost, st));
    TMI_HIP(hipStreamSynchronize(st));
    r.g_ci.alloc(nnz > 0 ? nnz : 1);
    r.g_cs.alloc(nnz > 0 ? nnz : 1);
    r.g_val.alloc(nnz > 0 ? nnz : 1);
    if (r.dim == 3)
        hipLaunchKernelGGL(k_g_fill<3>, gb, dim3(256), 0, st, r.blen, L, r.elref.p, gpos, r.g_rp.p, r.g_ci.p,
                           r.g_cs.p, r.g_val.p, flag.p);
    else
        hipLaunchKernelGGL(k_g_fill<2>, gb, dim3(256), 0, st, r.blen, L, r.elref.p, gpos, r.g_rp.p, r.g_ci.p,
                           r.g_cs.p, r.g_val.p, flag.p);
    TMI_LAUNCH_CHECK();
    launch_exclusive_scan(flag.p, spos.p, blen, st);
    TMI_HIP(hipMemcpyAsync(&r.slen, spos.p + blen, sizeof(int), hipMemcpyDeviceToHost, st));
    TMI_HIP(hipStreamSynchronize(st));
    r.sol2basis.alloc(r.slen > 0 ? r.slen : 1);
    hipLaunchKernelGGL(k_compact, gb, dim3(256), 0, st, flag.p, spos.p, r.blen, r.sol2basis.p);
    TMI_LAUNCH_CHECK();
    TMI_HIP(hipStreamSynchronize(st));
}

void Raster::build_gimap(hipStream_t st)
{
    if (gi_built || !has_g) return;
    const LpDims L = lp_dims(*this);
    const dim3 gg((unsigned)((glen + 255) / 256));
    DBuf<int> cnt(glen);
    gi_rp.alloc(glen + 1);
    if (dim == 3)
        hipLaunchKernelGGL(k_gi_row<3>, gg, dim3(256), 0, st, glen, L, elref.p, (const int *)nullptr, cnt.p,
                           (int *)nullptr, (double *)nullptr);
    else
        hipLaunchKernelGGL(k_gi_row<2>, gg, dim3(256), 0, st, glen, L, elref.p, (const int *)nullptr, cnt.p,
                           (int *)nullptr, (double *)nullptr);
    TMI_LAUNCH_CHECK();
    launch_exclusive_scan(cnt.p, gi_rp.p, (int)glen, st);
    int nnz = 0;
    TMI_HIP(hipMemcpyAsync(&nnz, gi_rp.p + glen, sizeof(int), hipMemcpyDeviceToHost, st));
    TMI_HIP(hipStreamSynchronize(st));
    gi_ci.alloc(nnz > 0 ? nnz : 1);
    gi_val.alloc(nnz > 0 ? nnz : 1);
    if (dim == 3)
        hipLaunchKernelGGL(k_gi_row<3>, gg, dim3(256), 0, st, glen, L, elref.p, gi_rp.p, cnt.p, gi_ci.p, gi_val.p);
    else
        hipLaunchKernelGGL(k_gi_row<2>, gg, dim3(256), 0, st, glen, L, elref.p, gi_rp.p, cnt.p, gi_ci.p, gi_val.p);
    TMI_LAUNCH_CHECK();
    TMI_HIP(hipStreamSynchronize(st));
    gi_built = true;
}

double Raster::voxel_size() const
{
    double sz = 1.0;
    for (int d = 0; d < dim; d++) sz *= (bbmax[d] - bbmin[d]) / (bdim[d] - 1.0);
    return sz;
}

// out[r][c] = sum_e val[e] in[col[e]][c] over CSR row r in order, complex
// columns (Ax_cplx, crmatrix_imp.hpp:1302-1327); thread per (r, c)
__global__ void k_csr_cols(long rows, int ncols, const int *__restrict__ rp, const int *__restrict__ ci,
                           const double *__restrict__ val, const double2 *__restrict__ in, double2 *__restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * ncols) return;
    const long r = t / ncols;
    const int c = (int)(t % ncols);
    double br = 0.0, bi = 0.0;
    for (int e = rp[r]; e < rp[r + 1]; e++) {
        const double v = val[e];
        const double2 z = in[(long)ci[e] * ncols + c];
        br += z.x * v;
        bi += z.y * v;
    }
    out[t] = make_double2(br, bi);
}

void launch_grid_to_basis(Raster &r, int ncols, const double *d_g, double *d_b, hipStream_t st)
{
    if (!r.has_g) {
        TMI_HIP(hipMemcpyAsync(d_b, d_g, sizeof(double) * 2 * (size_t)r.glen * ncols, hipMemcpyDeviceToDevice, st));
        return;
    }
    const long cnt = r.blen * ncols;
    hipLaunchKernelGGL(k_csr_cols, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, r.blen, ncols, r.g_rp.p,
                       r.g_ci.p, r.g_val.p, (const double2 *)d_g, (double2 *)d_b);
    TMI_LAUNCH_CHECK();
}

void launch_basis_to_grid(Raster &r, int ncols, const double *d_b, double *d_g, hipStream_t st)
{
    if (!r.has_g) {
        TMI_HIP(hipMemcpyAsync(d_g, d_b, sizeof(double) * 2 * (size_t)r.glen * ncols, hipMemcpyDeviceToDevice, st));
        return;
    }
    r.build_gimap(st);
    const long cnt = r.glen * ncols;
    hipLaunchKernelGGL(k_csr_cols, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, r.glen, ncols, r.gi_rp.p,
                       r.gi_ci.p, r.gi_val.p, (const double2 *)d_b, (double2 *)d_g);
    TMI_LAUNCH_CHECK();
}

void Raster::build(hipStream_t st)
{
    Mesh &m = *mesh;
    const int nn = m.nn;
    GridDesc gd{gdim[0], gdim[1], gdim[2], bbmin[0], bbmin[1], bbmin[2], dW[0], dW[1], dW[2],
                bbmax[0], bbmax[1], bbmax[2]};
    elref.alloc(glen);
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((glen + 255) / 256)), dim3(256), 0, st, elref.p, glen, INT_MAX);
    TMI_LAUNCH_CHECK();
    if (dim == 3)
        hipLaunchKernelGGL(k_elref<3>, dim3((m.elen + 127) / 128), dim3(128), 0, st, m.elen, m.nodes.p, m.elist.p,
                           m.geo.p, gd, elref.p);
    else
        hipLaunchKernelGGL(k_elref<2>, dim3((m.elen + 127) / 128), dim3(128), 0, st, m.elen, m.nodes.p, m.elist.p,
                           m.geo.p, gd, elref.p);
    TMI_LAUNCH_CHECK();
    DBuf<int> flag(glen), pos(glen + 1);
    const dim3 gg((unsigned)((glen + 255) / 256));
    hipLaunchKernelGGL(k_elref_finish, gg, dim3(256), 0, st, elref.p, glen, flag.p);
    TMI_LAUNCH_CHECK();
    TMI_CHECK(glen < INT_MAX, TMI_ERR_UNSUPPORTED, "grid of %ld points is too large", glen);
    launch_exclusive_scan(flag.p, pos.p, (int)glen, st);
    TMI_HIP(hipMemcpyAsync(&gslen, pos.p + glen, sizeof(int), hipMemcpyDeviceToHost, st));
    TMI_HIP(hipStreamSynchronize(st));
    gsol.alloc(gslen > 0 ? gslen : 1);
    hipLaunchKernelGGL(k_compact, gg, dim3(256), 0, st, flag.p, pos.p, glen, gsol.p);
    TMI_LAUNCH_CHECK();
    if (has_g) {
        build_gmap(*this, pos.p, st);
    } else {   // solution voxels = grid points inside the mesh
        slen = gslen;
        sol2basis.alloc(slen > 0 ? slen : 1);
        TMI_HIP(hipMemcpyAsync(sol2basis.p, gsol.p, sizeof(int) * (size_t)slen, hipMemcpyDeviceToDevice, st));
    }
    bcols.alloc((size_t)glen * nn);
    bw.alloc((size_t)glen * nn);
    if (dim == 3)
        hipLaunchKernelGGL(k_bmap<3>, gg, dim3(256), 0, st, glen, elref.p, m.elist.p, m.geo.p, gd, bcols.p, bw.p);
    else
        hipLaunchKernelGGL(k_bmap<2>, gg, dim3(256), 0, st, glen, elref.p, m.elist.p, m.geo.p, gd, bcols.p, bw.p);
    TMI_LAUNCH_CHECK();
    TMI_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------- fields
// out[g][c] = sum_j w[g][j] f[col[g][j]][c] (complex), j ascending as the
// CSR row of B; thread per (g, c) with c fastest
template <int NN>
__global__ void k_raster_map(long glen, int ncols, const int *__restrict__ cols, const double *__restrict__ w,
                             const double2 *__restrict__ f, int ld, double2 *__restrict__ out, long ldo)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= glen * ncols) return;
    const long g = t / ncols;
    const int c = (int)(t % ncols);
    double re = 0.0, im = 0.0;
    for (int j = 0; j < NN; j++) {
        const int nd = cols[g * NN + j];
        if (nd < 0) continue;
        const double2 v = f[(long)nd * ld + c];
        re += w[g * NN + j] * v.x;
        im += w[g * NN + j] * v.y;
    }
    out[g * ldo + c] = make_double2(re, im);
}

void launch_raster_map(const Raster &r, int ncols, const double *d_f, int ld, double *d_out, hipStream_t st)
{
    const long cnt = r.glen * ncols;
    const dim3 g((unsigned)((cnt + 255) / 256));
    if (r.dim == 3)
        hipLaunchKernelGGL(k_raster_map<4>, g, dim3(256), 0, st, r.glen, ncols, r.bcols.p, r.bw.p,
                           (const double2 *)d_f, ld, (double2 *)d_out, (long)ncols);
    else
        hipLaunchKernelGGL(k_raster_map<3>, g, dim3(256), 0, st, r.glen, ncols, r.bcols.p, r.bw.p,
                           (const double2 *)d_f, ld, (double2 *)d_out, (long)ncols);
    TMI_LAUNCH_CHECK();
}

// ImageGradient (jacobian.cc:944-1060) at the solution voxels: for column
// c and solution voxel s, fgs[c][s] = (value, d/dx, d/dy, d/dz) with the
// reference's spacing gsize/n and mask rules (neighbours tested first).
struct GradDesc {
    int nx, ny, nz, dim;
    double ih[3], i2h[3];
};

// columns c >= unmasked_from take their gradients without the mask.  Block
// of GG_V consecutive solution voxels x 64 columns: each wave takes every
// 4th voxel of the block, lanes over the columns, so that the reads of a grid
// point's columns are contiguous (one 1 KB row of fgv per value or
// neighbour); which neighbours count (mask, grid edges) is decided once per
// voxel (wave-uniform).  The records go to an LDS tile [column][voxel][4] and
// leave it column by column: the GG_V voxels of a column are one contiguous
// 1 KB run of fgs.
constexpr int GG_V = 16;

__global__ void __launch_bounds__(256) k_grid_grad(int slen, int ncols, const int *__restrict__ s2b,
                                                   const int *__restrict__ mask, const double2 *__restrict__ fgv,
                                                   GradDesc gd, double2 *__restrict__ fgs, int unmasked_from)
{
    __shared__ double2 tile[64][GG_V][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long ldv = ncols;
    const int len[3] = {gd.nx, gd.ny, gd.nz};
    const long str[3] = {1, gd.nx, (long)gd.nx * gd.ny};
    const int s0 = blockIdx.x * GG_V, nv = min(GG_V, slen - s0);
    for (int c0 = 0; c0 < ncols; c0 += 64) {
        const int c = c0 + lane;
        const int um = c >= unmasked_from ? 1 : 0;
        for (int i = w; i < nv; i += 4) {
            const long g = s2b[s0 + i];
            const int pos[3] = {(int)(g % gd.nx), (int)((g / gd.nx) % gd.ny), (int)(g / ((long)gd.nx * gd.ny))};
            const bool ctm = mask[g] >= 0;
            double2 rec[4];
            const double2 *v = fgv + (c < ncols ? c : 0);   // column c of [g][ncols]
            const double2 vc = v[g * ldv];
            rec[0] = vc;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                double2 r = make_double2(0.0, 0.0);
                if (d < gd.dim) {
                    const bool in_lo = pos[d] > 0, in_hi = pos[d] < len[d] - 1;
                    const bool mlo = in_lo && mask[g - str[d]] >= 0, mhi = in_hi && mask[g + str[d]] >= 0;
                    const bool bl = um ? in_lo : mlo, bh = um ? in_hi : mhi, ct = um ? true : ctm;
                    if (bl && bh) {
                        const double2 a = v[(g + str[d]) * ldv], b = v[(g - str[d]) * ldv];
                        r = make_double2((a.x - b.x) * gd.i2h[d], (a.y - b.y) * gd.i2h[d]);
                    } else if (ct && bl) {
                        const double2 b = v[(g - str[d]) * ldv];
                        r = make_double2((vc.x - b.x) * gd.ih[d], (vc.y - b.y) * gd.ih[d]);
                    } else if (ct && bh) {
                        const double2 a = v[(g + str[d]) * ldv];
                        r = make_double2((a.x - vc.x) * gd.ih[d], (a.y - vc.y) * gd.ih[d]);
                    }
                }
                rec[1 + d] = r;
            }
#pragma unroll
            for (int d = 0; d < 4; d++) tile[lane][i][d] = rec[d];
        }
        __syncthreads();
        // column runs: thread t writes record part e = t + 256 k of the
        // chunk, (column, voxel, part) with the part fastest
        const int nc = min(64, ncols - c0);
        for (int e = threadIdx.x; e < nc * GG_V * 4; e += 256) {
            const int cc = e / (GG_V * 4), r = e % (GG_V * 4), i = r >> 2, d = r & 3;
            if (i < nv) fgs[((long)(c0 + cc) * slen + s0 + i) * 4 + d] = tile[cc][i][d];
        }
        __syncthreads();
    }
}

// J rows through the detectors: block (voxel tile of 64, group of JG_T JG_M
// detectors), thread (voxel, detectors m + t JG_M, t < JG_T) holds the
// adjoint values and gradients of its detectors at its voxel and walks the
// sources q, reading the direct ones once for all of them (shared by the block's
// detectors through the cache, so each adjoint value is read once and each
// direct one once per detector group).  PMDF_mua = -d a, PMDF_kap = -sum_d
// grad_d d grad_d a (jacobian.cc:57-79), divided by the projection for log
// data (one reciprocal of |proj|^2 per link), times the voxel size sz.
constexpr int JG_M = 8;
#ifndef TMI_JG_T
#define TMI_JG_T 4
#endif
constexpr int JG_T = TMI_JG_T;

__global__ void __launch_bounds__(64 * JG_M) k_jac_grid(int slen, int nq, int nm, int dim,
                                                        const double2 *__restrict__ fgs, const int *__restrict__ lk,
                                                        const double2 *__restrict__ proj, int logscale, double sz,
                                                        int nqm, double *__restrict__ J)
{
    const int s = blockIdx.x * 64 + (threadIdx.x & 63);
    const int mb = blockIdx.y * JG_T * JG_M + (threadIdx.x >> 6);
    if (s >= slen) return;
    double2 a[JG_T][4];
    bool mv[JG_T];
#pragma unroll
    for (int t = 0; t < JG_T; t++) {
        const int m = mb + t * JG_M;
        mv[t] = m < nm;
        const double2 *A = fgs + ((long)(nq + min(m, nm - 1)) * slen + s) * 4;
#pragma unroll
        for (int d = 0; d < 4; d++) a[t][d] = A[d];
    }
    const size_t ld = 2 * (size_t)slen;
    for (int q = 0; q < nq; q++) {
        const double2 *D = fgs + ((long)q * slen + s) * 4;
        const double2 d0 = D[0], d1 = D[1], d2 = D[2], d3 = D[3];
#pragma unroll
        for (int t = 0; t < JG_T; t++) {
            if (!mv[t]) continue;
            const int k = lk[(long)q * nm + mb + t * JG_M];
            if (k < 0) continue;
            const double2 a0 = a[t][0], a1 = a[t][1], a2 = a[t][2], a3 = a[t][3];
            double pm_r = -(d0.x * a0.x - d0.y * a0.y), pm_i = -(d0.x * a0.y + d0.y * a0.x);
            double pk_r = 0.0, pk_i = 0.0;
            pk_r -= d1.x * a1.x - d1.y * a1.y;
            pk_i -= d1.x * a1.y + d1.y * a1.x;
            pk_r -= d2.x * a2.x - d2.y * a2.y;
            pk_i -= d2.x * a2.y + d2.y * a2.x;
            if (dim > 2) {
                pk_r -= d3.x * a3.x - d3.y * a3.y;
                pk_i -= d3.x * a3.y + d3.y * a3.x;
            }
            if (logscale) {
                const double2 p = proj[k];
                const double id = 1.0 / (p.x * p.x + p.y * p.y);
                const double ar = (pm_r * p.x + pm_i * p.y) * id, ai = (pm_i * p.x - pm_r * p.y) * id;
                const double br = (pk_r * p.x + pk_i * p.y) * id, bi = (pk_i * p.x - pk_r * p.y) * id;
                pm_r = ar; pm_i = ai; pk_r = br; pk_i = bi;
            }
            // J is written once: streaming stores keep it out of the caches
            __builtin_nontemporal_store(pm_r * sz, J + (size_t)k * ld + s);
            __builtin_nontemporal_store(pm_i * sz, J + (size_t)(k + nqm) * ld + s);
            __builtin_nontemporal_store(pk_r * sz, J + (size_t)k * ld + slen + s);
            __builtin_nontemporal_store(pk_i * sz, J + (size_t)(k + nqm) * ld + slen + s);
        }
    }
}

// CW (GenerateJacobian_cw_grid, jacobian.cc:610-721, then the log row
// scaling of :160-186) from real fields given with zero imaginary parts:
// Jmua[k][s] = (-d a sz) (1/proj_k), Jkap[k][s] = (-sum_d grad_d d grad_d a sz)
// (1/proj_k), the adjoint gradients taken without the mask; either output may
// be absent.  Same thread mapping as k_jac_grid.
__global__ void __launch_bounds__(64 * JG_M) k_jac_grid_cw(int slen, int nq, int nm, int dim,
                                                           const double2 *__restrict__ fgs,
                                                           const int *__restrict__ lk, const double *__restrict__ proj,
                                                           int logscale, double sz, double *__restrict__ Jmua,
                                                           double *__restrict__ Jkap)
{
    const int s = blockIdx.x * 64 + (threadIdx.x & 63);
    const int m = blockIdx.y * JG_M + (threadIdx.x >> 6);
    if (s >= slen || m >= nm) return;
    const double2 *A = fgs + ((long)(nq + m) * slen + s) * 4;
    const double a0 = A[0].x, a1 = A[1].x, a2 = A[2].x, a3 = A[3].x;
    for (int q = 0; q < nq; q++) {
        const int k = lk[(long)q * nm + m];
        if (k < 0) continue;
        const double2 *D = fgs + ((long)q * slen + s) * 4;
        double pm = -(D[0].x * a0) * sz;
        double pk = 0.0;
        pk -= D[1].x * a1;
        pk -= D[2].x * a2;
        if (dim > 2) pk -= D[3].x * a3;
        pk *= sz;
        if (logscale) {
            const double ip = 1.0 / proj[k];
            pm *= ip;
            pk *= ip;
        }
        if (Jmua) __builtin_nontemporal_store(pm, Jmua + (size_t)k * slen + s);
        if (Jkap) __builtin_nontemporal_store(pk, Jkap + (size_t)k * slen + s);
    }
}

// nq direct then nm adjoint complex columns mapped to the grid (r.fgv) and
// their values and gradients at the grid points inside the mesh (r.fgs
// [c][gslen][4]); columns c >= unmasked_from take unmasked gradients.
// Returns the voxel size prod gsize/(bdim-1).
static double grid_fields(Raster &r, int nq, int nm, const double *d_dphi, int ldd, const double *d_aphi, int lda,
                          int unmasked_from, hipStream_t st)
{
    const int ncols = nq + nm;
    TMI_CHECK(r.gslen > 0 && r.slen > 0, TMI_ERR_INVALID_ARG, "the raster has no voxel inside the mesh");
    r.fgv.alloc((size_t)r.glen * ncols * 2);
    r.fgs.alloc((size_t)r.gslen * ncols * 8);
    // fields on the grid: direct columns 0..nq-1, adjoint nq..nq+nm-1
    const long cnt_q = r.glen * nq, cnt_m = r.glen * nm;
    const int NN = r.dim + 1;
    if (NN == 4) {
        hipLaunchKernelGGL(k_raster_map<4>, dim3((unsigned)((cnt_q + 255) / 256)), dim3(256), 0, st, r.glen, nq,
                           r.bcols.p, r.bw.p, (const double2 *)d_dphi, ldd, (double2 *)r.fgv.p, (long)ncols);
        hipLaunchKernelGGL(k_raster_map<4>, dim3((unsigned)((cnt_m + 255) / 256)), dim3(256), 0, st, r.glen, nm,
                           r.bcols.p, r.bw.p, (const double2 *)d_aphi, lda, (double2 *)r.fgv.p + nq, (long)ncols);
    } else {
        hipLaunchKernelGGL(k_raster_map<3>, dim3((unsigned)((cnt_q + 255) / 256)), dim3(256), 0, st, r.glen, nq,
                           r.bcols.p, r.bw.p, (const double2 *)d_dphi, ldd, (double2 *)r.fgv.p, (long)ncols);
        hipLaunchKernelGGL(k_raster_map<3>, dim3((unsigned)((cnt_m + 255) / 256)), dim3(256), 0, st, r.glen, nm,
                           r.bcols.p, r.bw.p, (const double2 *)d_aphi, lda, (double2 *)r.fgv.p + nq, (long)ncols);
    }
    TMI_LAUNCH_CHECK();
    GradDesc gd{r.gdim[0], r.gdim[1], r.gdim[2], r.dim, {0, 0, 0}, {0, 0, 0}};
    for (int d = 0; d < r.dim; d++) {
        const double gsize = r.bbmax[d] - r.bbmin[d];
        const double h = gsize / r.gdim[d];
        gd.ih[d] = 1.0 / h;
        gd.i2h[d] = 0.5 * gd.ih[d];
    }
    hipLaunchKernelGGL(k_grid_grad, dim3((unsigned)((r.gslen + GG_V - 1) / GG_V)), dim3(256), 0, st, r.gslen, ncols,
                       r.gsol.p, r.elref.p, (const double2 *)r.fgv.p, gd, (double2 *)r.fgs.p, unmasked_from);
    TMI_LAUNCH_CHECK();
    return r.voxel_size();
}

// Map_GridToSol (raster_px.cc:102-114,148-151) of Jacobian rows computed at
// the grid points inside the mesh: out[row][s] = (sum_e G[b_s][e]
// in[row][e]) * sz, times 1/proj[row] when proj is given (the log row
// scaling of GenerateJacobian_cw, jacobian.cc:160-186); thread per (row, s)
__global__ void k_jac_to_basis(int rows, int slen, const int *__restrict__ s2b, const int *__restrict__ rp,
                               const int *__restrict__ cs, const double *__restrict__ val,
                               const double *__restrict__ in, long ldin, double *__restrict__ out, long ldout,
                               double sz, const double *__restrict__ proj)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)rows * slen) return;
    const int row = (int)(t / slen), s = (int)(t % slen);
    const int b = s2b[s];
    const double *x = in + (size_t)row * ldin;
    double sum = 0.0;
    for (int e = rp[b]; e < rp[b + 1]; e++) sum += x[cs[e]] * val[e];
    double v = sum * sz;
    if (proj) v *= 1.0 / proj[row];
    out[(size_t)row * ldout + s] = v;
}

static void jac_to_basis(const Raster &r, int rows, const double *in, long ldin, double *out, long ldout, double sz,
                         const double *proj, hipStream_t st)
{
    const long cnt = (long)rows * r.slen;
    if (cnt == 0) return;
    hipLaunchKernelGGL(k_jac_to_basis, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, rows, r.slen,
                       r.sol2basis.p, r.g_rp.p, r.g_cs.p, r.g_val.p, in, ldin, out, ldout, sz, proj);
    TMI_LAUNCH_CHECK();
}

// AddDataGradient's raster part (toastmodule.cc:1813-1836): for each source q
// in order, grad_mua -= Re(d_q a_q) and grad_kappa -= Re(sum_d grad_d d_q
// grad_d a_q) at every solution voxel, a_q the adjoint field of source q;
// thread per voxel, so each voxel's sums run in the reference's order.
__global__ void k_data_grad(int slen, int nq, const double2 *__restrict__ fgs, double *__restrict__ grad)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slen) return;
    double gm = grad[s], gk = grad[slen + s];
    for (int q = 0; q < nq; q++) {
        const double2 *D = fgs + ((long)q * slen + s) * 4, *A = fgs + ((long)(nq + q) * slen + s) * 4;
        gm -= D[0].x * A[0].x - D[0].y * A[0].y;
        double t = 0.0;   // real part of the complex sum over the dimensions
        for (int d = 1; d <= 3; d++) t += D[d].x * A[d].x - D[d].y * A[d].y;
        gk -= t;
    }
    grad[s] = gm;
    grad[slen + s] = gk;
}

// the same through G: per solution voxel and source in order, the G row
// applied to Re(d a) and to Re(sum_d grad_d d grad_d a) at the grid points
// inside the mesh (Map_GridToSol then Re, toastmodule.cc:1821-1832)
__global__ void k_data_grad_g(int slen, int nq, int gslen, const int *__restrict__ s2b, const int *__restrict__ rp,
                              const int *__restrict__ cs, const double *__restrict__ val,
                              const double2 *__restrict__ fgs, double *__restrict__ grad)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slen) return;
    const int b = s2b[s];
    double gm = grad[s], gk = grad[slen + s];
    for (int q = 0; q < nq; q++) {
        double sm = 0.0, sk = 0.0;
        for (int e = rp[b]; e < rp[b + 1]; e++) {
            const double v = val[e];
            const double2 *D = fgs + ((long)q * gslen + cs[e]) * 4, *A = fgs + ((long)(nq + q) * gslen + cs[e]) * 4;
            sm += (D[0].x * A[0].x - D[0].y * A[0].y) * v;
            double t = 0.0;
            for (int d = 1; d <= 3; d++) t += D[d].x * A[d].x - D[d].y * A[d].y;
            sk += t * v;
        }
        gm -= sm;
        gk -= sk;
    }
    grad[s] = gm;
    grad[slen + s] = gk;
}

void launch_data_gradient(Raster &r, int nq, const double *d_dphi, int ldd, const double *d_aphi, int lda,
                          double *d_grad, hipStream_t st)
{
    grid_fields(r, nq, nq, d_dphi, ldd, d_aphi, lda, 2 * nq, st);
    if (r.has_g)
        hipLaunchKernelGGL(k_data_grad_g, dim3((unsigned)((r.slen + 255) / 256)), dim3(256), 0, st, r.slen, nq,
                           r.gslen, r.sol2basis.p, r.g_rp.p, r.g_cs.p, r.g_val.p, (const double2 *)r.fgs.p, d_grad);
    else
        hipLaunchKernelGGL(k_data_grad, dim3((unsigned)((r.slen + 255) / 256)), dim3(256), 0, st, r.slen, nq,
                           (const double2 *)r.fgs.p, d_grad);
    TMI_LAUNCH_CHECK();
}

void launch_jacobian_grid(Raster &r, int nq, int nm, const double *d_dphi, int ldd, const double *d_aphi, int lda,
                          const int *d_lk, int nqm, const double *d_proj, int logscale, double *d_J, hipStream_t st,
                          bool cw, double *d_Jkap)
{
    const double sz = grid_fields(r, nq, nm, d_dphi, ldd, d_aphi, lda, cw ? nq : nq + nm, st);
    const dim3 grid((unsigned)((r.gslen + 63) / 64), (unsigned)((nm + JG_M - 1) / JG_M));   // CW kernel
    const dim3 grid2((unsigned)((r.gslen + 63) / 64), (unsigned)((nm + JG_T * JG_M - 1) / (JG_T * JG_M)));
    const dim3 blk(64 * JG_M);
    if (!r.has_g) {   // solution voxels = grid points inside the mesh
        if (cw)
            hipLaunchKernelGGL(k_jac_grid_cw, grid, blk, 0, st, r.gslen, nq, nm, r.dim, (const double2 *)r.fgs.p,
                               d_lk, d_proj, logscale, sz, d_J, d_Jkap);
        else
            hipLaunchKernelGGL(k_jac_grid, grid2, blk, 0, st, r.gslen, nq, nm, r.dim, (const double2 *)r.fgs.p,
                               d_lk, (const double2 *)d_proj, logscale, sz, nqm, d_J);
        TMI_LAUNCH_CHECK();
        return;
    }
    // PMDFs at the grid points inside the mesh (unit scale, log division of
    // the complex rows included), then G, the voxel size and for CW the
    // 1/proj row scaling
    const size_t gs = (size_t)r.gslen;
    if (cw) {
        DBuf<double> jm(d_J ? (size_t)nqm * gs : 0), jk(d_Jkap ? (size_t)nqm * gs : 0);
        hipLaunchKernelGGL(k_jac_grid_cw, grid, blk, 0, st, r.gslen, nq, nm, r.dim, (const double2 *)r.fgs.p,
                           d_lk, d_proj, 0, 1.0, jm.p, jk.p);
        TMI_LAUNCH_CHECK();
        const double *rs = logscale ? d_proj : nullptr;
        if (d_J) jac_to_basis(r, nqm, jm.p, (long)gs, d_J, r.slen, sz, rs, st);
        if (d_Jkap) jac_to_basis(r, nqm, jk.p, (long)gs, d_Jkap, r.slen, sz, rs, st);
        TMI_HIP(hipStreamSynchronize(st));
        return;
    }
    DBuf<double> jg((size_t)4 * nqm * gs);
    hipLaunchKernelGGL(k_jac_grid, grid2, blk, 0, st, r.gslen, nq, nm, r.dim, (const double2 *)r.fgs.p, d_lk,
                       (const double2 *)d_proj, logscale, 1.0, nqm, jg.p);
    TMI_LAUNCH_CHECK();
    const long ldo = 2 * (long)r.slen;
    jac_to_basis(r, 2 * nqm, jg.p, 2 * (long)gs, d_J, ldo, sz, nullptr, st);                  // mua block
    jac_to_basis(r, 2 * nqm, jg.p + gs, 2 * (long)gs, d_J + r.slen, ldo, sz, nullptr, st);     // kappa block
    TMI_HIP(hipStreamSynchronize(st));
}

}  // namespace tmi
