// Mesh setup, CSR pattern, system-matrix assembly and Gaussian source
// vectors on the GPU.
//
// This file is compiled with -ffp-contract=off: the element integrals and
// the per-entry summation order equal the reference's sequential assembly
// (elements ascending; PFF, then PDD, then BNDPFF, then i*omega*FF), so the
// FP64 system matrix is bit-identical to the reference's where the device
// math functions (sqrt, hypot) round as the host's do.
//
// Data layout: every per-node / per-element array is a flat SoA array in
// HBM; element geometry is an AoS record per element (GEO2 / GEO3 doubles)
// because the gather kernels read whole records of a few elements per row.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fe_integrals.hpp"
#include "mesh.hpp"

namespace tmi {

// ---------------------------------------------------------------- scan
// Exclusive prefix sum of n ints into out[0..n] (out[n] = total), one
// workgroup: each thread scans a contiguous chunk, a wave-level Hillis-Steele
// scan combines the chunk totals.  Setup-time only (n up to a few 1e6).
__global__ void __launch_bounds__(1024) k_scan1(const int *__restrict__ in, int *__restrict__ out, int n)
{
    __shared__ int part[1024];
    int t = threadIdx.x, T = blockDim.x;
    long chunk = ((long)n + T - 1) / T;
    long b = t * chunk, e = b + chunk;
    if (b > n) b = n;
    if (e > n) e = n;
    int s = 0;
    for (long i = b; i < e; i++) s += in[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        int v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;   // exclusive prefix of this chunk
    for (long i = b; i < e; i++) { out[i] = run; run += in[i]; }
    if (t == T - 1) out[n] = part[T - 1];
}

void launch_exclusive_scan(const int *in, int *out, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, st, in, out, n);
    TMI_LAUNCH_CHECK();
}

// ------------------------------------------------- node -> element adjacency
__global__ void k_n2e_count(const int *__restrict__ el, int elen, int nn, int *__restrict__ cnt)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)elen * nn) return;
    atomicAdd(&cnt[el[t]], 1);
}

__global__ void k_n2e_fill(const int *__restrict__ el, int elen, int nn, const int *__restrict__ ptr,
                           int *__restrict__ cursor, int *__restrict__ n2e)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)elen * nn) return;
    int node = el[t];
    int pos = atomicAdd(&cursor[node], 1);
    n2e[ptr[node] + pos] = (int)(t / nn);
}

// the fill order is arbitrary: sort each node's list ascending (lists are short)
__global__ void k_n2e_sort(const int *__restrict__ ptr, int *__restrict__ n2e, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b = ptr[i], e = ptr[i + 1];
    for (int k = b + 1; k < e; k++) {
        int v = n2e[k], j = k - 1;
        while (j >= b && n2e[j] > v) { n2e[j + 1] = n2e[j]; j--; }
        n2e[j + 1] = v;
    }
}

// ------------------------------------------------------------- geometry
__global__ void k_geometry(const double *__restrict__ x, const int *__restrict__ el, int elen, int dim,
                           double *__restrict__ geo)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= elen) return;
    if (dim == 3) {
        const int *nd = el + 4 * (long)e;
        double P[4][3];
        for (int k = 0; k < 4; k++)
            for (int d = 0; d < 3; d++) P[k][d] = x[3 * (long)nd[k] + d];
        double g[GEO3];
        tet_geometry(P[0], P[1], P[2], P[3], g);
        for (int k = 21; k < GEO3; k++) g[k] = 0.0;
        for (int k = 0; k < GEO3; k++) geo[(long)e * GEO3 + k] = g[k];
    } else {
        const int *nd = el + 3 * (long)e;
        double P[3][2];
        for (int k = 0; k < 3; k++)
            for (int d = 0; d < 2; d++) P[k][d] = x[2 * (long)nd[k] + d];
        double g[GEO2];
        tri_geometry(P[0], P[1], P[2], g);
        for (int k = 0; k < GEO2; k++) geo[(long)e * GEO2 + k] = g[k];
    }
}

__device__ inline bool elem_has(const int *nd, int nn, int v)
{
    for (int k = 0; k < nn; k++)
        if (nd[k] == v) return true;
    return false;
}

// MarkBoundary: a side is a boundary side when no other element holds all
// of its nodes (the reference sorts all side tuples and marks the ones that
// equal neither sorted neighbour, mesh.cc:1614-1703; same set).  Boundary
// side nodes become boundary nodes.
__global__ void k_mark_boundary(const int *__restrict__ el, int elen, int dim, const int *__restrict__ n2e_ptr,
                                const int *__restrict__ n2e, int *__restrict__ eflags, int *__restrict__ bnd)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= elen) return;
    int nn = dim + 1, nsd = dim + 1, nsn = dim;
    const int *nd = el + (long)nn * e;
    int flags = 0;
    for (int s = 0; s < nsd; s++) {
        int f[3];
        for (int k = 0; k < nsn; k++) f[k] = nd[dim == 3 ? tet_side_node(s, k) : tri_side_node(s, k)];
        bool shared = false;
        for (int q = n2e_ptr[f[0]]; q < n2e_ptr[f[0] + 1] && !shared; q++) {
            int e2 = n2e[q];
            if (e2 == e) continue;
            const int *nd2 = el + (long)nn * e2;
            bool all = true;
            for (int k = 1; k < nsn; k++) all = all && elem_has(nd2, nn, f[k]);
            shared = all;
        }
        if (!shared) {
            flags |= 1 << s;
            for (int k = 0; k < nsn; k++) atomicOr(&bnd[f[k]], 1);
        }
    }
    if (flags) flags |= BNDEL_BIT;
    eflags[e] = flags;
}

__global__ void k_tri_bndff(const double *__restrict__ x, const int *__restrict__ el, int elen,
                            const int *__restrict__ eflags, double *__restrict__ geo)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= elen) return;
    int fl = eflags[e];
    if (!(fl & BNDEL_BIT)) return;
    const int *nd = el + 3 * (long)e;
    double P[3][2];
    for (int k = 0; k < 3; k++)
        for (int d = 0; d < 2; d++) P[k][d] = x[2 * (long)nd[k] + d];
    double g[GEO2];
    for (int k = 0; k < GEO2; k++) g[k] = geo[(long)e * GEO2 + k];
    tri_bndff(P[0], P[1], P[2], fl & 7, g);
    for (int k = 10; k < 16; k++) geo[(long)e * GEO2 + k] = g[k];
}

void Mesh::setup(hipStream_t st)
{
    nn = dim + 1;
    TMI_CHECK(dim == 2 || dim == 3, TMI_ERR_INVALID_ARG, "mesh dimension must be 2 or 3 (got %d)", dim);
    TMI_CHECK(n > 0 && elen > 0, TMI_ERR_INVALID_ARG, "empty mesh (%d nodes, %d elements)", n, elen);
    for (size_t k = 0; k < h_elist.size(); k++)
        TMI_CHECK(h_elist[k] >= 0 && h_elist[k] < n, TMI_ERR_INVALID_ARG,
                  "element %zu references node %d outside [0,%d)", k / nn, h_elist[k], n);
    nodes.upload(h_nodes.data(), h_nodes.size(), st);
    elist.upload(h_elist.data(), h_elist.size(), st);

    // node -> element adjacency
    DBuf<int> cnt(n), cursor(n);
    TMI_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), st));
    TMI_HIP(hipMemsetAsync(cursor.p, 0, cursor.bytes(), st));
    long tot = (long)elen * nn;
    hipLaunchKernelGGL(k_n2e_count, dim3(grid_for(tot, 256)), dim3(256), 0, st, elist.p, elen, nn, cnt.p);
    TMI_LAUNCH_CHECK();
    n2e_ptr.alloc(n + 1);
    launch_exclusive_scan(cnt.p, n2e_ptr.p, n, st);
    n2e.alloc(tot);
    hipLaunchKernelGGL(k_n2e_fill, dim3(grid_for(tot, 256)), dim3(256), 0, st, elist.p, elen, nn, n2e_ptr.p,
                       cursor.p, n2e.p);
    TMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_n2e_sort, dim3(grid_for(n, 256)), dim3(256), 0, st, n2e_ptr.p, n2e.p, n);
    TMI_LAUNCH_CHECK();

    // geometry, boundary
    geo.alloc((size_t)elen * geo_stride());
    hipLaunchKernelGGL(k_geometry, dim3(grid_for(elen, 256)), dim3(256), 0, st, nodes.p, elist.p, elen, dim, geo.p);
    TMI_LAUNCH_CHECK();
    eflags.alloc(elen);
    bnd.alloc(n);
    TMI_HIP(hipMemsetAsync(bnd.p, 0, bnd.bytes(), st));
    hipLaunchKernelGGL(k_mark_boundary, dim3(grid_for(elen, 256)), dim3(256), 0, st, elist.p, elen, dim, n2e_ptr.p,
                       n2e.p, eflags.p, bnd.p);
    TMI_LAUNCH_CHECK();
    if (dim == 2) {
        hipLaunchKernelGGL(k_tri_bndff, dim3(grid_for(elen, 256)), dim3(256), 0, st, nodes.p, elist.p, elen,
                           eflags.p, geo.p);
        TMI_LAUNCH_CHECK();
    }
    h_bnd = bnd.to_host(st);
    nbnd = 0;
    for (int v : h_bnd) nbnd += v;
}

// ---------------------------------------------------------- CSR pattern
// Row i of the pattern = sorted unique nodes of the elements holding node i
// (mesh.cc:422-549 sorts all element node pairs and drops duplicates; the
// result is the same set in the same order).  Each thread walks its row's
// element list repeatedly, taking the next larger neighbour each time: no
// per-thread arrays, deterministic output.
__device__ inline int next_neighbour(const int *el, int nn, const int *n2e, int b, int e, int last)
{
    int cand = 0x7fffffff;
    for (int q = b; q < e; q++) {
        const int *nd = el + (long)nn * n2e[q];
        for (int k = 0; k < nn; k++) {
            int v = nd[k];
            if (v > last && v < cand) cand = v;
        }
    }
    return cand;
}

__global__ void k_csr_count(const int *__restrict__ el, int nn, const int *__restrict__ n2e_ptr,
                            const int *__restrict__ n2e, int n, int *__restrict__ rowlen)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b = n2e_ptr[i], e = n2e_ptr[i + 1];
    int cnt = 0, last = -1;
    for (;;) {
        int v = next_neighbour(el, nn, n2e, b, e, last);
        if (v == 0x7fffffff) break;
        cnt++;
        last = v;
    }
    rowlen[i] = cnt;
}

__global__ void k_csr_fill(const int *__restrict__ el, int nn, const int *__restrict__ n2e_ptr,
                           const int *__restrict__ n2e, int n, const int *__restrict__ rowptr,
                           int *__restrict__ colidx, int *__restrict__ rowidx, int *__restrict__ diagpos)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b = n2e_ptr[i], e = n2e_ptr[i + 1];
    int k = rowptr[i], last = -1, kend = rowptr[i + 1];
    diagpos[i] = -1;
    while (k < kend) {
        int v = next_neighbour(el, nn, n2e, b, e, last);
        if (v == 0x7fffffff) break;
        if (v == i) diagpos[i] = k;
        colidx[k] = v;
        rowidx[k] = i;
        k++;
        last = v;
    }
}

void Mesh::build_csr(hipStream_t st)
{
    if (have_csr) return;
    DBuf<int> rowlen(n);
    hipLaunchKernelGGL(k_csr_count, dim3(grid_for(n, 128)), dim3(128), 0, st, elist.p, nn, n2e_ptr.p, n2e.p, n,
                       rowlen.p);
    TMI_LAUNCH_CHECK();
    rowptr.alloc(n + 1);
    launch_exclusive_scan(rowlen.p, rowptr.p, n, st);
    TMI_HIP(hipMemcpyAsync(&nnz, rowptr.p + n, sizeof(int), hipMemcpyDeviceToHost, st));
    TMI_HIP(hipStreamSynchronize(st));
    colidx.alloc(nnz);
    rowidx.alloc(nnz);
    diagpos.alloc(n);
    hipLaunchKernelGGL(k_csr_fill, dim3(grid_for(n, 128)), dim3(128), 0, st, elist.p, nn, n2e_ptr.p, n2e.p, n,
                       rowptr.p, colidx.p, rowidx.p, diagpos.p);
    TMI_LAUNCH_CHECK();
    have_csr = true;
}

// ------------------------------------------------------------- assembly
// One thread per nonzero (i, j): walk the elements of row i in ascending
// order and add each element's (i, j) contribution, pass by pass, exactly as
// the reference adds them into the CSR slot (AddToElMatrix -> Add).
// Nodal coefficients (elbasis = 0): IntPFF / IntPDD / BndIntPFF of the
// nodal values; per-element coefficients (elbasis = 1, ASSEMBLE_*_EL,
// mesh.cc:1985-1993, :2112-2120): the element's value times IntFF / IntDD /
// BndIntFF.
__device__ inline int local_index(const int *nd, int nn, int v)
{
    for (int k = 0; k < nn; k++)
        if (nd[k] == v) return k;
    return -1;
}

template <int DIM>
__global__ void k_assemble(const int *__restrict__ el, const double *__restrict__ geo, const int *__restrict__ eflags,
                           const int *__restrict__ n2e_ptr, const int *__restrict__ n2e,
                           const int *__restrict__ rowidx, const int *__restrict__ colidx, int nnz,
                           const double *__restrict__ cmua, const double *__restrict__ ckap,
                           const double *__restrict__ c2a, double omega, int complex_valued, int elbasis,
                           double *__restrict__ val64, void *__restrict__ val32)
{
    constexpr int NN = DIM + 1;
    constexpr int GS = DIM == 3 ? GEO3 : GEO2;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    int i = rowidx[k], j = colidx[k];
    int b = n2e_ptr[i], e = n2e_ptr[i + 1];
    double re = 0.0, im = 0.0;
    for (int pass = 0; pass < 4; pass++) {
        if (pass == 3 && !complex_valued) break;
        const double *P = pass == 0 ? cmua : (pass == 1 ? ckap : c2a);
        for (int q = b; q < e; q++) {
            int elx = n2e[q];
            const int *nd = el + (long)NN * elx;
            int lj = local_index(nd, NN, j);
            if (lj < 0) continue;
            int li = local_index(nd, NN, i);
            const double *g = geo + (long)GS * elx;
            if (elbasis && pass < 3) {
                const double pe = P[elx];
                if (DIM == 3) {
                    if (pass == 0) re += tet_intff(g, li, lj) * pe;
                    else if (pass == 1) re += tet_intdd(g, li, lj) * pe;
                    else {
                        int fl = eflags[elx];
                        if (fl & BNDEL_BIT) re += tet_bndff(g, fl, li, lj) * pe;
                    }
                } else {
                    if (pass == 0) re += tri_intff(g, li, lj) * pe;
                    else if (pass == 1) re += tri_intdd(g, li, lj) * pe;
                    else re += tri_bndff(g, li, lj) * pe;
                }
                continue;
            }
            if (DIM == 3) {
                if (pass == 3) { im += tet_intff(g, li, lj) * omega; continue; }
                double p0 = P[nd[0]], p1 = P[nd[1]], p2 = P[nd[2]], p3 = P[nd[3]];
                if (pass == 0) re += tet_intpff(g, li, lj, p0, p1, p2, p3);
                else if (pass == 1) re += tet_intpdd(g, li, lj, p0, p1, p2, p3);
                else {
                    int fl = eflags[elx];
                    if (fl & BNDEL_BIT) {
                        double p4[4] = {p0, p1, p2, p3};
                        re += tet_bndintpff(g, fl, li, lj, p4);
                    }
                }
            } else {
                if (pass == 3) { im += tri_intff(g, li, lj) * omega; continue; }
                double p0 = P[nd[0]], p1 = P[nd[1]], p2 = P[nd[2]];
                if (pass == 0) re += tri_intpff(g, li, lj, p0, p1, p2);
                else if (pass == 1) re += tri_intpdd(g, li, lj, p0, p1, p2);
                else re += tri_bndintpff(g, li, lj, p0, p1, p2);
            }
        }
    }
    if (complex_valued) {
        val64[2 * (long)k] = re;
        val64[2 * (long)k + 1] = im;
        if (val32) { ((float *)val32)[2 * (long)k] = (float)re; ((float *)val32)[2 * (long)k + 1] = (float)im; }
    } else {
        val64[k] = re;
        if (val32) ((float *)val32)[k] = (float)re;
    }
}

void launch_assemble(const Mesh &m, const double *d_cmua, const double *d_ckap, const double *d_c2a, double omega,
                     int complex_valued, double *d_val64, void *d_val32, hipStream_t st, int elbasis)
{
    dim3 g(grid_for(m.nnz, 256)), b(256);
    if (m.dim == 3)
        hipLaunchKernelGGL(k_assemble<3>, g, b, 0, st, m.elist.p, m.geo.p, m.eflags.p, m.n2e_ptr.p, m.n2e.p,
                           m.rowidx.p, m.colidx.p, m.nnz, d_cmua, d_ckap, d_c2a, omega, complex_valued, elbasis,
                           d_val64, d_val32);
    else
        hipLaunchKernelGGL(k_assemble<2>, g, b, 0, st, m.elist.p, m.geo.p, m.eflags.p, m.n2e_ptr.p, m.n2e.p,
                           m.rowidx.p, m.colidx.p, m.nnz, d_cmua, d_ckap, d_c2a, omega, complex_valued, elbasis,
                           d_val64, d_val32);
    TMI_LAUNCH_CHECK();
}

// ------------------------------------------------------ Gaussian sources
// QVec_Gaussian (source.cc:4-41) as a gather: the value at node j sums, over
// the elements holding j in ascending order and their nodes i in ascending
// order, q_i * BndIntFF(i, j) (Neumann) or q_i * IntFF(i, j) (isotropic) --
// the order in which the reference's element loop adds into qvec[j].
template <int DIM>
__global__ void k_qvec_gaussian(const double *__restrict__ x, const int *__restrict__ el,
                                const double *__restrict__ geo, const int *__restrict__ eflags,
                                const int *__restrict__ n2e_ptr, const int *__restrict__ n2e,
                                const int *__restrict__ bnd, int n, int npts, const double *__restrict__ centres,
                                double width, int neumann, const double *__restrict__ scale,
                                double *__restrict__ out)
{
    constexpr int NN = DIM + 1;
    constexpr int GS = DIM == 3 ? GEO3 : GEO2;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * npts) return;
    int j = (int)(t % n), s = (int)(t / n);
    double v = 0.0;
    if (!neumann || bnd[j]) {
        const double PI = 3.1415926535897932384626433832795;
        double w2 = width * width;
        double fac1 = 1.0 / (sqrt(2.0 * PI) * width);
        if (DIM == 3) fac1 /= (sqrt(2.0 * PI) * width);
        if (!neumann) fac1 /= (sqrt(2.0 * PI) * width);
        double fac2 = -0.5 / w2;
        const double *c = centres + (long)DIM * s;
        for (int q = n2e_ptr[j]; q < n2e_ptr[j + 1]; q++) {
            int elx = n2e[q];
            const int *nd = el + (long)NN * elx;
            const double *g = geo + (long)GS * elx;
            int fl = eflags[elx];
            int lj = local_index(nd, NN, j);
            for (int i = 0; i < NN; i++) {
                int is = nd[i];
                if (neumann && !bnd[is]) continue;
                double d2 = 0.0;
                for (int d = 0; d < DIM; d++) d2 += (c[d] - x[(long)DIM * is + d]) * (c[d] - x[(long)DIM * is + d]);
                double dd = sqrt(d2);
                double qq = exp(dd * dd * fac2) * fac1;
                double w;
                if (neumann) w = DIM == 3 ? tet_bndff(g, fl, i, lj) : tri_bndff(g, i, lj);
                else w = DIM == 3 ? tet_intff(g, i, lj) : tri_intff(g, i, lj);
                v += qq * w;
            }
        }
        if (scale) v = v * scale[j];
    }
    out[t] = v;
}

void launch_qvec_gaussian(const Mesh &m, int npts, const double *d_centres, double width, int neumann,
                          const double *d_scale, double *d_out, hipStream_t st)
{
    long work = (long)m.n * npts;
    dim3 g(grid_for(work, 256)), b(256);
    if (m.dim == 3)
        hipLaunchKernelGGL(k_qvec_gaussian<3>, g, b, 0, st, m.nodes.p, m.elist.p, m.geo.p, m.eflags.p, m.n2e_ptr.p,
                           m.n2e.p, m.bnd.p, m.n, npts, d_centres, width, neumann, d_scale, d_out);
    else
        hipLaunchKernelGGL(k_qvec_gaussian<2>, g, b, 0, st, m.nodes.p, m.elist.p, m.geo.p, m.eflags.p, m.n2e_ptr.p,
                           m.n2e.p, m.bnd.p, m.n, npts, d_centres, width, neumann, d_scale, d_out);
    TMI_LAUNCH_CHECK();
}

}  // namespace tmi

namespace tmi {

// ----------------------------------------------------------- projections
// ProjectSingle/ProjectAll (fwdsolver.cc:1415-1447): proj[k] = sum_j
// phi_q[j] * mvec[m, j] for the link k = (q, m), over the nonzeros j of mvec
// row m (the reference's dense dot, vector_imp.hpp:1045-1052, whose other
// terms are exact zeros).  phi element (q, j) is at phi[q*sq + j*sj]
// (complex: interleaved).
//
// A detector's row is split into segments of S nonzeros; one wave per
// (segment, detector).  Its lanes are QW sources (consecutive q: with
// node-major fields, sq = 1, one 16*QW-byte run of a field row per nonzero)
// times EP = 64 / QW interleaved entry streams; each lane sums its entries in
// ascending order, the EP streams are combined by a butterfly, and the
// segment sums of each link are added in segment order by k_project_fin.
// The summation order is fixed by (S, QW): results are reproducible run to
// run.
template <bool CPLX>
__global__ void __launch_bounds__(64) k_project_seg(int nq, int nm, int S, int qw, const int *__restrict__ lk,
                                                    const int *__restrict__ mrp, const int *__restrict__ mci,
                                                    const double *__restrict__ mval, const double *__restrict__ phi,
                                                    long sq, long sj, double *__restrict__ partial)
{
    constexpr int CW = CPLX ? 2 : 1;
    const int seg = blockIdx.x, m = blockIdx.y, nseg = gridDim.x;
    const int r0 = mrp[m], r1 = mrp[m + 1];
    const int e0 = r0 + seg * S;
    if (e0 >= r1) return;   // the row ends before this segment
    const int e1 = min(r1, e0 + S);
    const int lane = threadIdx.x, ep = lane / qw, EP = 64 / qw;
    for (int qb = 0; qb < nq; qb += qw) {
        const int q = qb + lane % qw;
        const bool qon = q < nq;
        const long qo = (long)(qon ? q : 0) * sq;
        double sr = 0.0, si = 0.0;
        constexpr int U = 8;
        for (int e = e0 + ep; e < e1; e += U * EP) {
            long jj[U];
            double c[U], d[U], a[U], b[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int eu = min(e + u * EP, e1 - 1);
                jj[u] = mci[eu];
                c[u] = mval[CW * (long)eu];
                d[u] = CPLX ? mval[CW * (long)eu + 1] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const long at = qo + jj[u] * sj;
                a[u] = phi[CW * at];
                b[u] = CPLX ? phi[CW * at + 1] : 0.0;
            }
            // the batch's loads stay ahead of the ordered sums
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool on = e + u * EP < e1;   // entries past the segment end are not added
                if (CPLX) {
                    const double tr = a[u] * c[u] - b[u] * d[u], ti = a[u] * d[u] + b[u] * c[u];
                    sr = on ? sr + tr : sr;
                    si = on ? si + ti : si;
                } else {
                    const double tr = a[u] * c[u];
                    sr = on ? sr + tr : sr;
                }
            }
        }
        // combine the EP entry streams of each source (same order in every lane)
        for (int off = qw; off < 64; off <<= 1) {
            sr += __shfl_xor(sr, off);
            if (CPLX) si += __shfl_xor(si, off);
        }
        if (ep == 0 && qon && lk[(long)m * nq + q] >= 0) {
            double *o = partial + CW * (((long)m * nseg + seg) * nq + q);
            o[0] = sr;
            if (CPLX) o[1] = si;
        }
    }
}

// proj[k] of every linked (m, q): the row's segment sums in segment order
template <bool CPLX>
__global__ void k_project_fin(int nq, int nm, int S, int nseg, const int *__restrict__ lk,
                              const int *__restrict__ mrp, const double *__restrict__ partial,
                              double *__restrict__ proj)
{
    constexpr int CW = CPLX ? 2 : 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nq * nm) return;
    const int m = (int)(t / nq), q = (int)(t % nq);
    const int k = lk[t];
    if (k < 0) return;
    const int len = mrp[m + 1] - mrp[m], ns = (len + S - 1) / S;
    double sr = 0.0, si = 0.0;
    // loads in batches of B (independent), sums in segment order
    constexpr int B = 16;
    const double *o = partial + CW * ((long)m * nseg * nq + q);
    const long step = (long)CW * nq;
    for (int g0 = 0; g0 < ns; g0 += B) {
        double a[B], b[B];
#pragma unroll
        for (int u = 0; u < B; u++) {
            const int g = min(g0 + u, ns - 1);
            a[u] = o[g * step];
            b[u] = CPLX ? o[g * step + 1] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < B; u++) {
            if (g0 + u < ns) {
                sr += a[u];
                if (CPLX) si += b[u];
            }
        }
    }
    proj[CW * (long)k] = sr;
    if (CPLX) proj[CW * (long)k + 1] = si;
}

ProjectPlan project_plan(int nq, int nm, int maxrow, int cplx)
{
    ProjectPlan p;
    p.qw = 1;
    while (p.qw < nq && p.qw < 64) p.qw *= 2;
    // segments: enough waves to fill the GPU (about 2048), at least 64
    // entries per lane stream
    const int ep = 64 / p.qw;
    p.S = 4096;
    while (p.S > 64 * ep && (long)nm * ((maxrow + p.S - 1) / p.S) < 2048) p.S /= 2;
    p.nseg = std::max(1, (maxrow + p.S - 1) / p.S);
    p.partial_doubles = (size_t)nm * p.nseg * nq * (cplx ? 2 : 1);
    return p;
}

void launch_project(const ProjectPlan &pl, int nq, int nm, const int *d_lk, const int *mrp, const int *mci,
                    const double *mval, const double *phi, long sq, long sj, int cplx, double *partial, double *proj,
                    hipStream_t st)
{
    TMI_CHECK(pl.qw >= 1 && pl.qw <= 64 && (64 % pl.qw) == 0 && pl.S > 0 && pl.nseg > 0, TMI_ERR_INVALID_ARG,
              "bad projection plan");
    const dim3 g(pl.nseg, nm);
    if (cplx)
        hipLaunchKernelGGL(k_project_seg<true>, g, dim3(64), 0, st, nq, nm, pl.S, pl.qw, d_lk, mrp, mci, mval, phi,
                           sq, sj, partial);
    else
        hipLaunchKernelGGL(k_project_seg<false>, g, dim3(64), 0, st, nq, nm, pl.S, pl.qw, d_lk, mrp, mci, mval, phi,
                           sq, sj, partial);
    TMI_LAUNCH_CHECK();
    const long pairs = (long)nq * nm;
    if (cplx)
        hipLaunchKernelGGL(k_project_fin<true>, dim3(grid_for(pairs, 256)), dim3(256), 0, st, nq, nm, pl.S, pl.nseg,
                           d_lk, mrp, partial, proj);
    else
        hipLaunchKernelGGL(k_project_fin<false>, dim3(grid_for(pairs, 256)), dim3(256), 0, st, nq, nm, pl.S,
                           pl.nseg, d_lk, mrp, partial, proj);
    TMI_LAUNCH_CHECK();
}

}  // namespace tmi
