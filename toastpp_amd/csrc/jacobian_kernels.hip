// Nodal-basis Jacobian (GenerateJacobian_mesh, jacobian.cc:441-497) on the GPU.
//
// For node i with neighbour set N(i) (its CSR row) the reference's PMDFs
//   pmdf_mua_i(q,m) = sum_el sum_jk IntFFF(i,j,k) f_j g_k   (IntFG, jacobian.cc:806-835)
//   pmdf_kap_i(q,m) = sum_el sum_jk IntFDD(i,j,k) f_j g_k   (IntGradFGradG, :839-872)
// with f = dphi_q, g = aphi_m are bilinear forms f^T A_i g and f^T B_i g,
// where A_i, B_i (|N(i)| x |N(i)|) collect the element integrals of the
// elements holding i.  Over all source/detector pairs they are two small
// GEMMs per node: P_i = F_i (A_i G_i^T), F_i = dphi on N(i) (nq x d),
// G_i = aphi on N(i) (nm x d).  One workgroup per node: it builds A_i, B_i
// in LDS from the element integrals, then for each chunk of detectors
// gathers G into LDS and forms H = G A, K = G B, and for each chunk of
// sources gathers F and forms P for the linked pairs of the two chunks,
// written node-major (one contiguous row of the intermediate per node).  A tiled transpose kernel
// then scales by 1/proj (log data), negates and writes J in the reference
// layout [dlnA/dmua | dlnA/dkappa ; dphase/dmua | dphase/dkappa].
//
// This file is compiled without floating-point contraction so that A_i and
// B_i use the exact element integrals of the assembly; the GEMM loops opt in
// to FMA contraction with a pragma.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "fe_integrals.hpp"
#include "jacobian.hpp"

namespace tmi {

// threads per node (four waves: every SIMD of a CU gets work even at the
// two or three nodes per CU that the LDS allows)
constexpr int JAC_NT = 256;
#ifndef TMI_JAC_LDS_KIB
#define TMI_JAC_LDS_KIB 40
#endif
constexpr int JAC_LDS_KIB = TMI_JAC_LDS_KIB;

// J and the node-major intermediate are written once and read at most once:
// streaming (nontemporal) accesses keep them out of the caches
typedef double jac_dv2 __attribute__((ext_vector_type(2)));
__device__ inline void st_stream(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ inline void st_stream(double2 *p, double2 v)
{
    __builtin_nontemporal_store(__builtin_bit_cast(jac_dv2, v), (jac_dv2 *)p);
}
__device__ inline double2 ld_stream(const double2 *p)
{
    return __builtin_bit_cast(double2, __builtin_nontemporal_load((const jac_dv2 *)p));
}

struct JacDims {
    int n, nq, nm, nqm, dmax, chunk, ldd, lda;
};

__device__ inline int find_pos(const int *cols, int d, int v)
{
    for (int t = 0; t < d; t++)
        if (cols[t] == v) return t;
    return -1;
}

// LDS layout per node (doubles): A[dmax^2] B[dmax^2] W[C dmax] (complex: the
// detector chunk G, then the source chunk F) H[C dmax] Kk[C dmax] cols[dmax]
// (ints), C = source/detector chunk
// the W/H/Kk region also holds the element-integral scratch (2 doubles and
// an int per thread) before the GEMMs start
__host__ __device__ inline size_t jac_region_doubles(int dmax, int chunk)
{
    const size_t g = (size_t)6 * chunk * dmax, e = (size_t)3 * JAC_NT;
    return g > e ? g : e;
}

static size_t jac_lds_doubles(int dmax, int chunk)
{
    return (size_t)2 * dmax * dmax + jac_region_doubles(dmax, chunk) + (dmax + 1) / 2 + 1;
}

// Output of the direct variant: J (complex: 2nqm x 2n, [Re; Im] x [mua |
// kappa]) or the CW pair Jmua / Jkap (nqm x n each, either may be absent),
// scaled by 1/proj (log data) and negated, as k_jac_transpose /
// k_jac_transpose_cw write them from the intermediate.
struct JacOut {
    double *J, *Jkap;
    const double *proj;
    int logscale, cw;
};

__device__ inline void jac_emit(const JacOut &o, int n, int nqm, int i, int k, double mr, double mi, double pr,
                                double pi)
{
    if (o.cw) {
        double a = -mr, b = -pr;
        if (o.logscale) {
            const double ip = 1.0 / o.proj[k];
            a *= ip;
            b *= ip;
        }
        if (o.J) o.J[(size_t)k * n + i] = a;
        if (o.Jkap) o.Jkap[(size_t)k * n + i] = b;
        return;
    }
    double2 a = make_double2(mr, mi), b = make_double2(pr, pi);
    if (o.logscale) {
        const double2 p = ((const double2 *)o.proj)[k];
        const double den = p.x * p.x + p.y * p.y;
        a = make_double2((a.x * p.x + a.y * p.y) / den, (a.y * p.x - a.x * p.y) / den);
        b = make_double2((b.x * p.x + b.y * p.y) / den, (b.y * p.x - b.x * p.y) / den);
    }
    const size_t ld = 2 * (size_t)n;
    o.J[(size_t)k * ld + i] = -a.x;
    o.J[(size_t)(k + nqm) * ld + i] = -a.y;
    o.J[(size_t)k * ld + n + i] = -b.x;
    o.J[(size_t)(k + nqm) * ld + n + i] = -b.y;
}

// node of workgroup b such that each XCD (workgroups b = x mod 8) walks a
// contiguous run of nodes: the J entries of neighbouring nodes share cache
// lines, which then fill in that XCD's L2 before they are written back
__device__ inline int xcd_node(int b, int n)
{
    const int x = b & 7, s = b >> 3, q = n >> 3, r = n & 7;
    return x * q + min(x, r) + s;
}

template <int DIM, bool DIRECT>
__global__ void __launch_bounds__(JAC_NT) k_jac_nodes(JacDims jd, const int *__restrict__ rowptr,
                                                  const int *__restrict__ colidx, const int *__restrict__ n2e_ptr,
                                                  const int *__restrict__ n2e, const int *__restrict__ el,
                                                  const double *__restrict__ geo,
                                                  const double2 *__restrict__ dphi,   // [n][ldd]
                                                  const double2 *__restrict__ aphi,   // [n][lda]
                                                  const int *__restrict__ lk,         // [nq][nm] link or -1
                                                  double2 *__restrict__ Pt,           // [n][nqm][2]
                                                  JacOut jo)
{
    constexpr int NN = DIM + 1;
    constexpr int GS = DIM == 3 ? GEO3 : GEO2;
    extern __shared__ double lds[];
    const int i = DIRECT ? xcd_node(blockIdx.x, jd.n) : blockIdx.x;
    const int tid = threadIdx.x;
    const int dmax = jd.dmax, nq = jd.nq, nm = jd.nm, C = jd.chunk;
    double *A = lds, *B = A + dmax * dmax;
    double2 *W = (double2 *)(B + dmax * dmax);
    double2 *H = W + C * dmax, *Kk = H + C * dmax;
    int *cols = (int *)((double *)W + jac_region_doubles(dmax, C));
    const int rb = rowptr[i], d = rowptr[i + 1] - rb;
    for (int t = tid; t < d; t += JAC_NT) cols[t] = colidx[rb + t];
    for (int t = tid; t < dmax * dmax; t += JAC_NT) { A[t] = 0.0; B[t] = 0.0; }
    __syncthreads();
    // element integrals: the elements of node i in rounds of R, thread
    // (element, a, b) evaluating IntFFF(li, a, b) and IntFDD(a, b) into a
    // scratch in the W/H/Kk region (not in use yet), then threads (a, b)
    // adding them into A, B one element after another (deterministic sums,
    // the same order as a serial loop over the elements)
    {
        double *ef = (double *)W, *ed = ef + (size_t)JAC_NT;
        int *ep = (int *)(ed + (size_t)JAC_NT);
        constexpr int R = JAC_NT / (NN * NN);
        const int q0 = n2e_ptr[i], q1 = n2e_ptr[i + 1];
        for (int qb = q0; qb < q1; qb += R) {
            const int ne = min(R, q1 - qb);
            if (tid < ne * NN * NN) {
                const int le = tid / (NN * NN), ab = tid % (NN * NN);
                const int e = n2e[qb + le];
                const int *nd = el + (long)NN * e;
                const double *g = geo + (long)GS * e;
                const int a = ab / NN, b = ab % NN;
                int li = 0;
                for (int k = 0; k < NN; k++)
                    if (nd[k] == i) li = k;
                const int pa = find_pos(cols, d, nd[a]), pb = find_pos(cols, d, nd[b]);
                ef[tid] = DIM == 3 ? tet_intfff(g, li, a, b) : tri_intfff(g, li, a, b);
                ed[tid] = DIM == 3 ? tet_intfdd(g, a, b) : tri_intfdd(g, a, b);
                // the row of i holds every node of its elements: pa, pb >= 0
                ep[tid] = (pa >= 0 && pb >= 0) ? pa * dmax + pb : -1;
            }
            __syncthreads();
            if (tid < NN * NN) {   // one wave: the (a, b) of an element hit distinct entries
                for (int le = 0; le < ne; le++) {
                    const int at = le * NN * NN + tid, pos = ep[at];
                    if (pos >= 0) {
                        A[pos] += ef[at];
                        B[pos] += ed[at];
                    }
                }
            }
            __syncthreads();
        }
    }
    double2 *out = Pt + (size_t)i * jd.nqm * 2;
    for (int m0 = 0; m0 < nm; m0 += C) {
        const int mc = min(C, nm - m0);
        // detector chunk G on the neighbour set
        for (int t = tid; t < mc * d; t += JAC_NT) {
            const int mm = t / d, j = t % d;
            W[mm * dmax + j] = aphi[(size_t)cols[j] * jd.lda + m0 + mm];
        }
        __syncthreads();
        {
#pragma clang fp contract(fast)
            // H = G A^T, Kk = G B^T  (A, B symmetric)
            for (int t = tid; t < mc * d; t += JAC_NT) {
                const int mm = t / d, j = t % d;
                double hr = 0.0, hi = 0.0, kr = 0.0, ki = 0.0;
                for (int k = 0; k < d; k++) {
                    const double2 gv = W[mm * dmax + k];
                    const double a = A[j * dmax + k], b = B[j * dmax + k];
                    hr += a * gv.x; hi += a * gv.y;
                    kr += b * gv.x; ki += b * gv.y;
                }
                H[mm * dmax + j] = make_double2(hr, hi);
                Kk[mm * dmax + j] = make_double2(kr, ki);
            }
        }
        __syncthreads();
        for (int q0 = 0; q0 < nq; q0 += C) {
            const int qc = min(C, nq - q0);
            for (int t = tid; t < qc * d; t += JAC_NT) {
                const int qq = t / d, j = t % d;
                W[qq * dmax + j] = dphi[(size_t)cols[j] * jd.ldd + q0 + qq];
            }
            __syncthreads();
            {
#pragma clang fp contract(fast)
                // P(q,m) = F_q . H_m: one thread per pair of the chunk pair,
                // linked pairs only (lk[q nm + m] = link or -1)
                for (int b = tid; b < qc * mc; b += JAC_NT) {
                    const int r = b / mc, cc = b % mc;
                    const int k = lk[(size_t)(q0 + r) * nm + m0 + cc];
                    if (k < 0) continue;
                    const double2 *f = W + r * dmax, *h = H + cc * dmax, *kk = Kk + cc * dmax;
                    double mr = 0.0, mi = 0.0, pr = 0.0, pi = 0.0;
                    for (int j = 0; j < d; j++) {
                        const double2 fv = f[j], hv = h[j], kv = kk[j];
                        mr += fv.x * hv.x - fv.y * hv.y;
                        mi += fv.x * hv.y + fv.y * hv.x;
                        pr += fv.x * kv.x - fv.y * kv.y;
                        pi += fv.x * kv.y + fv.y * kv.x;
                    }
                    if constexpr (DIRECT) {
                        jac_emit(jo, jd.n, jd.nqm, i, k, mr, mi, pr, pi);
                    } else {
                        st_stream(out + 2 * k, make_double2(mr, mi));
                        st_stream(out + 2 * k + 1, make_double2(pr, pi));
                    }
                }
            }
            __syncthreads();
        }
    }
}

// J[k][i] = -Re(P_mua/proj_k), J[k+nqm][i] = -Im(.), J[k][n+i] = -Re(P_kap/proj_k),
// J[k+nqm][n+i] = -Im(.)  (DATA_LOG; without log the division is skipped).
// Tiles of 64 nodes x 16 measurements through LDS: the reads take 512-byte
// runs of one node's measurements, the writes 512-byte runs of one J row.
constexpr int TJK = 16;
__global__ void __launch_bounds__(256) k_jac_transpose(int n, int nqm, const double2 *__restrict__ Pt,
                                                       const double2 *__restrict__ proj, int logscale,
                                                       double *__restrict__ J)
{
    __shared__ double2 tm[64][TJK + 1], tk[64][TJK + 1];
    const int i0 = blockIdx.x * 64, k0 = blockIdx.y * TJK;
    const int t = threadIdx.x;
    // read: node r, measurement c (2 values each)
    for (int e = t; e < 64 * TJK; e += 256) {
        const int r = e / TJK, c = e % TJK;
        const int i = i0 + r, k = k0 + c;
        if (i < n && k < nqm) {
            tm[r][c] = ld_stream(Pt + ((size_t)i * nqm + k) * 2);
            tk[r][c] = ld_stream(Pt + ((size_t)i * nqm + k) * 2 + 1);
        }
    }
    __syncthreads();
    const size_t ld = 2 * (size_t)n;
    for (int e = t; e < 64 * TJK; e += 256) {
        const int c = e / 64, r = e % 64;
        const int k = k0 + c, i = i0 + r;
        if (k < nqm && i < n) {
            double2 a = tm[r][c], b = tk[r][c];
            if (logscale) {
                const double2 p = proj[k];
                const double den = p.x * p.x + p.y * p.y;
                a = make_double2((a.x * p.x + a.y * p.y) / den, (a.y * p.x - a.x * p.y) / den);
                b = make_double2((b.x * p.x + b.y * p.y) / den, (b.y * p.x - b.x * p.y) / den);
            }
            st_stream(J + (size_t)k * ld + i, -a.x);
            st_stream(J + (size_t)(k + nqm) * ld + i, -a.y);
            st_stream(J + (size_t)k * ld + n + i, -b.x);
            st_stream(J + (size_t)(k + nqm) * ld + n + i, -b.y);
        }
    }
}

// CW (GenerateJacobian_cw_mesh, jacobian.cc:727-790, then the log row scaling
// of :160-186): from the intermediate of real fields, Jmua[k][i] =
// -P_mua * (1/proj_k), Jkap[k][i] = -P_kap * (1/proj_k) (without log the
// factor is dropped); either output may be absent.  Same tiling as above.
__global__ void __launch_bounds__(256) k_jac_transpose_cw(int n, int nqm, const double2 *__restrict__ Pt,
                                                          const double *__restrict__ proj, int logscale,
                                                          double *__restrict__ Jmua, double *__restrict__ Jkap)
{
    __shared__ double tm[64][TJK + 1], tk[64][TJK + 1];
    const int i0 = blockIdx.x * 64, k0 = blockIdx.y * TJK;
    const int t = threadIdx.x;
    for (int e = t; e < 64 * TJK; e += 256) {
        const int r = e / TJK, c = e % TJK;
        const int i = i0 + r, k = k0 + c;
        if (i < n && k < nqm) {
            tm[r][c] = ld_stream(Pt + ((size_t)i * nqm + k) * 2).x;
            tk[r][c] = ld_stream(Pt + ((size_t)i * nqm + k) * 2 + 1).x;
        }
    }
    __syncthreads();
    for (int e = t; e < 64 * TJK; e += 256) {
        const int c = e / 64, r = e % 64;
        const int k = k0 + c, i = i0 + r;
        if (k < nqm && i < n) {
            double a = -tm[r][c], b = -tk[r][c];
            if (logscale) {
                const double ip = 1.0 / proj[k];
                a *= ip;
                b *= ip;
            }
            if (Jmua) st_stream(Jmua + (size_t)k * n + i, a);
            if (Jkap) st_stream(Jkap + (size_t)k * n + i, b);
        }
    }
}

// Element-basis Jacobian (GenerateJacobian_mesh with elbasis = true,
// jacobian.cc:441-497; IntFG_el :879-906, IntGradFGradG_el :911-939): one
// thread per (element, source).  The element's IntFF / IntDD entries and the
// direct field at its nodes stay in registers; for every detector linked to
// the source the thread gathers the adjoint field at the nodes, forms
//   P_mua = sum_i sum_j (f_i g_j) IntFF(i,j),  P_kap = sum_i sum_j (f_i g_j) IntDD(i,j)
// in the reference's order (i outer, j inner, complex product first), scales
// by 1/proj (log data) and writes the link's four J entries -- rows k and
// k + nqm, columns e and elen + e -- coalesced over the workgroup's
// consecutive elements.
// mvec_n > 0 reproduces the mvec overload of GenerateJacobian_mesh, whose
// kappa PMDF is a view at offset n instead of elen (jacobian.cc:397): mua
// columns n .. elen-1 then hold the kappa PMDF of elements 0 .. elen-n-1.
// cw: real fields given with zero imaginary parts; Jmua / Jkap (nqm x elen)
// take -P * (1/proj) (GenerateJacobian_cw_mesh :727-790 + the row scaling
// of :160-186).
template <int DIM, bool CW>
__global__ void __launch_bounds__(256) k_jac_el(int elen, int nqm, const int *__restrict__ el,
                                                const double *__restrict__ geo, const double2 *__restrict__ dphi,
                                                int ldd, const double2 *__restrict__ aphi, int lda,
                                                const int *__restrict__ lptr, const int *__restrict__ lm,
                                                const double2 *__restrict__ proj, const double *__restrict__ projr,
                                                int logscale, int mvec_n, double *__restrict__ J,
                                                double *__restrict__ Jkap)
{
    constexpr int NN = DIM + 1;
    constexpr int GS = DIM == 3 ? GEO3 : GEO2;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (e >= elen) return;
    int nd[NN];
    for (int i = 0; i < NN; i++) nd[i] = el[(long)NN * e + i];
    const double *g = geo + (long)GS * e;
    double ff[NN][NN], dd[NN][NN];
    for (int i = 0; i < NN; i++)
        for (int j = 0; j < NN; j++) {
            ff[i][j] = DIM == 3 ? tet_intff(g, i, j) : tri_intff(g, i, j);
            dd[i][j] = DIM == 3 ? tet_intdd(g, i, j) : tri_intdd(g, i, j);
        }
    double2 f[NN];
    for (int i = 0; i < NN; i++) f[i] = dphi[(long)nd[i] * ldd + q];
    const size_t ld = CW ? (size_t)elen : 2 * (size_t)elen;
    for (int k = lptr[q]; k < lptr[q + 1]; k++) {
        const int mi = lm[k];
        double2 a[NN];
        for (int j = 0; j < NN; j++) a[j] = aphi[(long)nd[j] * lda + mi];
        double mr = 0.0, mim = 0.0, kr = 0.0, kim = 0.0;
        for (int i = 0; i < NN; i++)
            for (int j = 0; j < NN; j++) {
                const double pr = f[i].x * a[j].x - f[i].y * a[j].y;
                const double pi = f[i].x * a[j].y + f[i].y * a[j].x;
                mr += pr * ff[i][j];
                mim += pi * ff[i][j];
                kr += pr * dd[i][j];
                kim += pi * dd[i][j];
            }
        if (CW) {
            double am = -mr, ak = -kr;
            if (logscale) {
                const double ip = 1.0 / projr[k];
                am *= ip;
                ak *= ip;
            }
            if (J) st_stream(J + (size_t)k * ld + e, am);
            if (Jkap) st_stream(Jkap + (size_t)k * ld + e, ak);
            continue;
        }
        if (logscale) {
            const double2 p = proj[k];
            const double den = p.x * p.x + p.y * p.y;
            const double ar = (mr * p.x + mim * p.y) / den, ai = (mim * p.x - mr * p.y) / den;
            const double br = (kr * p.x + kim * p.y) / den, bi = (kim * p.x - kr * p.y) / den;
            mr = ar; mim = ai; kr = br; kim = bi;
        }
        double *R0 = J + (size_t)k * ld, *R1 = J + (size_t)(k + nqm) * ld;
        if (mvec_n <= 0 || e < mvec_n) {
            st_stream(R0 + e, -mr);
            st_stream(R1 + e, -mim);
        }
        st_stream(R0 + elen + e, -kr);
        st_stream(R1 + elen + e, -kim);
        if (mvec_n > 0 && e + mvec_n < elen) {
            st_stream(R0 + e + mvec_n, -kr);
            st_stream(R1 + e + mvec_n, -kim);
        }
    }
}

void launch_jacobian_el(const Mesh &m, int nq, const double *d_dphi, int ldd, const double *d_aphi, int lda,
                        const int *d_lptr, const int *d_lm, int nqm, const double *d_proj, int logscale, int mvec_n,
                        double *d_J, hipStream_t st, bool cw, double *d_Jkap)
{
    dim3 g((m.elen + 255) / 256, nq);
    const double2 *pc = cw ? nullptr : (const double2 *)d_proj;
    const double *pr = cw ? d_proj : nullptr;
    if (m.dim == 3) {
        if (cw)
            hipLaunchKernelGGL((k_jac_el<3, true>), g, dim3(256), 0, st, m.elen, nqm, m.elist.p, m.geo.p,
                               (const double2 *)d_dphi, ldd, (const double2 *)d_aphi, lda, d_lptr, d_lm, pc, pr,
                               logscale, 0, d_J, d_Jkap);
        else
            hipLaunchKernelGGL((k_jac_el<3, false>), g, dim3(256), 0, st, m.elen, nqm, m.elist.p, m.geo.p,
                               (const double2 *)d_dphi, ldd, (const double2 *)d_aphi, lda, d_lptr, d_lm, pc, pr,
                               logscale, mvec_n, d_J, nullptr);
    } else {
        if (cw)
            hipLaunchKernelGGL((k_jac_el<2, true>), g, dim3(256), 0, st, m.elen, nqm, m.elist.p, m.geo.p,
                               (const double2 *)d_dphi, ldd, (const double2 *)d_aphi, lda, d_lptr, d_lm, pc, pr,
                               logscale, 0, d_J, d_Jkap);
        else
            hipLaunchKernelGGL((k_jac_el<2, false>), g, dim3(256), 0, st, m.elen, nqm, m.elist.p, m.geo.p,
                               (const double2 *)d_dphi, ldd, (const double2 *)d_aphi, lda, d_lptr, d_lm, pc, pr,
                               logscale, mvec_n, d_J, nullptr);
    }
    TMI_LAUNCH_CHECK();
}

// real fields (node-major, row strides ldd / lda) into the complex layout of
// the nodal kernel, imaginary parts zero
__global__ void k_real_to_cplx(long rows, int cols, const double *__restrict__ src, int ld, double2 *__restrict__ dst)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * cols) return;
    const long r = t / cols;
    const int c = (int)(t % cols);
    dst[t] = make_double2(src[r * ld + c], 0.0);
}

void launch_real_to_cplx(long rows, int cols, const double *src, int ld, double *dst, hipStream_t st)
{
    const long cnt = rows * cols;
    hipLaunchKernelGGL(k_real_to_cplx, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, rows, cols, src, ld,
                       (double2 *)dst);
    TMI_LAUNCH_CHECK();
}

// How the nodal Jacobian kernel's output reaches J (TMI_JAC_MODE, read
// once): by default the node kernel writes J itself (nodes dealt to the XCDs
// in contiguous runs, so that the 8-byte entries of neighbouring nodes fill
// cache lines in one L2); "two" keeps the earlier layout, a node-major
// intermediate of n nqm 32 B (streaming stores) and one tiled transpose.
bool jacobian_two_pass()
{
    static const bool two = getenv("TMI_JAC_MODE") && !strcmp(getenv("TMI_JAC_MODE"), "two");
    return two;
}

void launch_jacobian_nodal(const Mesh &m, int nq, int nm, const double *d_dphi, int ldd, const double *d_aphi, int lda,
                           int nqm, const double *d_proj, int logscale, double *d_Pt, double *d_J, int dmax,
                           hipStream_t st, bool cw, double *d_Jkap)
{
    // LDS row stride: the neighbour count rounded up to odd, so that the
    // rows of A, B, W, H read by consecutive lanes start in different banks
    dmax |= 1;
    // source/detector chunk: all of them if a node's LDS stays within
    // JAC_LDS_KIB (several nodes per CU), else the largest chunk that fits
    // (at least 8 when the LDS of one CU allows)
    const int want = std::max(nq, nm);
    int chunk = want;
    while (chunk > 1 && jac_lds_doubles(dmax, chunk) * sizeof(double) > JAC_LDS_KIB * 1024) chunk--;
    if (chunk < std::min(want, 8))
        while (chunk < want && chunk < 8 && jac_lds_doubles(dmax, chunk + 1) * sizeof(double) <= 160 * 1024) chunk++;
    const size_t lds = jac_lds_doubles(dmax, chunk) * sizeof(double);
    TMI_CHECK(lds <= 160 * 1024, TMI_ERR_UNSUPPORTED,
              "nodal Jacobian: %d neighbours per node exceed the LDS of one CU", dmax);
    const JacDims jd{m.n, nq, nm, nqm, dmax, chunk, ldd, lda};
    const JacOut jo{d_J, d_Jkap, d_proj, logscale, cw ? 1 : 0};
    const bool two = jacobian_two_pass();
#define JN(D, DIR)                                                                                                  \
    hipLaunchKernelGGL((k_jac_nodes<D, DIR>), dim3(m.n), dim3(JAC_NT), lds, st, jd, m.rowptr.p, m.colidx.p,          \
                       m.n2e_ptr.p, m.n2e.p, m.elist.p, m.geo.p, (const double2 *)d_dphi, (const double2 *)d_aphi,     \
                       m.d_lk.p, (double2 *)d_Pt, jo);
    if (m.dim == 3) { if (two) JN(3, false) else JN(3, true) }
    else { if (two) JN(2, false) else JN(2, true) }
#undef JN
    TMI_LAUNCH_CHECK();
    if (!two) return;
    dim3 g((m.n + 63) / 64, (nqm + TJK - 1) / TJK);
    if (cw)
        hipLaunchKernelGGL(k_jac_transpose_cw, g, dim3(256), 0, st, m.n, nqm, (const double2 *)d_Pt, d_proj, logscale,
                           d_J, d_Jkap);
    else
        hipLaunchKernelGGL(k_jac_transpose, g, dim3(256), 0, st, m.n, nqm, (const double2 *)d_Pt,
                           (const double2 *)d_proj, logscale, d_J);
    TMI_LAUNCH_CHECK();
}

}  // namespace tmi
