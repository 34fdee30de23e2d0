// Block Krylov kernels for gfx950.
//
// Layout: block vectors are node-major [n][K] (K = columns of the tile, a
// power of two <= 64).  The SpMM maps one wavefront lane to one column and
// 64/K rows to a wavefront, so each nonzero of a CSR row costs one coalesced
// load of K neighbour values; matrix values and column indices are wave
// uniform for K = 64 (readfirstlane -> scalar loads).  For a single column
// (K = 1) a second SpMV maps LPR lanes to the nonzeros of one row.
//
// Every row-parallel kernel owns a contiguous row range per workgroup and
// remaps workgroup ids so that each XCD (workgroups are dealt round-robin
// over the 8 XCDs) walks one contiguous slice of the mesh: the neighbour
// rows gathered by the SpMM then sit in that XCD's L2.
//
// Reductions (inner products per column) are accumulated in FP64 per lane,
// combined across lanes, waves and workgroups in a fixed order: results are
// run-to-run deterministic.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "krylov.hpp"

namespace tmi {

// ------------------------------------------------------------ value types
__device__ inline float2 vadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline double2 vadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline float vadd(float a, float b) { return a + b; }
__device__ inline double vadd(double a, double b) { return a + b; }
__device__ inline float2 vsub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline double2 vsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline float vsub(float a, float b) { return a - b; }
__device__ inline double vsub(double a, double b) { return a - b; }
__device__ inline float2 vmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 vmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline float vmul(float a, float b) { return a * b; }
// acc + a * b with four fused multiply-adds (complex) / one (real)
__device__ inline float2 vfma(float2 acc, float2 a, float2 b)
{
    acc.x = __builtin_fmaf(a.x, b.x, acc.x);
    acc.x = __builtin_fmaf(-a.y, b.y, acc.x);
    acc.y = __builtin_fmaf(a.x, b.y, acc.y);
    acc.y = __builtin_fmaf(a.y, b.x, acc.y);
    return acc;
}
__device__ inline double2 vfma(double2 acc, double2 a, double2 b)
{
    acc.x = __builtin_fma(a.x, b.x, acc.x);
    acc.x = __builtin_fma(-a.y, b.y, acc.x);
    acc.y = __builtin_fma(a.x, b.y, acc.y);
    acc.y = __builtin_fma(a.y, b.x, acc.y);
    return acc;
}
__device__ inline float vfma(float acc, float a, float b) { return __builtin_fmaf(a, b, acc); }
__device__ inline double vfma(double acc, double a, double b) { return __builtin_fma(a, b, acc); }
__device__ inline double vmul(double a, double b) { return a * b; }

// acc + c x as four fused multiply-adds (complex) or one (real): the
// accumulation of the SpMM kernels
__device__ inline float2 vmac(float2 acc, float2 c, float2 x)
{
    acc.x = fmaf(c.x, x.x, acc.x);
    acc.x = fmaf(-c.y, x.y, acc.x);
    acc.y = fmaf(c.x, x.y, acc.y);
    acc.y = fmaf(c.y, x.x, acc.y);
    return acc;
}
__device__ inline double2 vmac(double2 acc, double2 c, double2 x)
{
    acc.x = fma(c.x, x.x, acc.x);
    acc.x = fma(-c.y, x.y, acc.x);
    acc.y = fma(c.x, x.y, acc.y);
    acc.y = fma(c.y, x.x, acc.y);
    return acc;
}
__device__ inline float vmac(float acc, float c, float x) { return fmaf(c, x, acc); }
__device__ inline double vmac(double acc, double c, double x) { return fma(c, x, acc); }

// scalar type of a vector element type
template <class V> struct ScalarOf { using type = V; };
template <> struct ScalarOf<float2> { using type = float; };
template <> struct ScalarOf<double2> { using type = double; };
// Multi-frequency operator (MF): the stored coefficient holds the real part
// and the imaginary part at omega = 1 (the mass term); a column at omega w
// uses re + i w im.  Real types and MF = false pass the coefficient through.
template <bool MF, class V>
__device__ inline V mfc(V c, typename ScalarOf<V>::type w)
{
    if constexpr (MF && (std::is_same<V, float2>::value || std::is_same<V, double2>::value)) {
        c.y *= w;
        return c;
    } else {
        (void)w;
        return c;
    }
}
// the column's omega (1 without MF)
template <bool MF, class V>
__device__ inline typename ScalarOf<V>::type mf_omega(const double *omg, int c)
{
    using S = typename ScalarOf<V>::type;
    if constexpr (MF) return (S)omg[c];
    else return (S)1;
}
__device__ inline float2 vscale(float2 a, double s) { float f = (float)s; return make_float2(a.x * f, a.y * f); }
__device__ inline double2 vscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ inline float vscale(float a, double s) { return a * (float)s; }
__device__ inline double vscale(double a, double s) { return a * s; }
__device__ inline double dotR(float2 a, float2 b) { return (double)a.x * (double)b.x + (double)a.y * (double)b.y; }
__device__ inline double dotR(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
__device__ inline double dotR(float a, float b) { return (double)a * (double)b; }
__device__ inline double dotR(double a, double b) { return a * b; }
// SpMM epilogue dots of one entry: mode 0: d1 += <w, y>_R; mode 1: also
// d2 += <y, y>_R; mode DOT_BIL (complex): d1, d2 += Re, Im of w y
// (unconjugated, COCG's p^T A p)
template <class V>
__device__ inline void dot_epi(V w, V y, int mode, double &d1, double &d2)
{
    if constexpr (std::is_same<V, float2>::value || std::is_same<V, double2>::value) {
        if (mode == DOT_BIL) {
            d1 += (double)w.x * (double)y.x - (double)w.y * (double)y.y;
            d2 += (double)w.x * (double)y.y + (double)w.y * (double)y.x;
            return;
        }
    }
    d1 += dotR(w, y);
    if (mode) d2 += dotR(y, y);
}

// sign pattern of the initial residual (-1, 0, +1 per real component), the
// shadow residual of the mixed-precision BiCGSTAB: any fixed shadow vector
// with <rd, r0> != 0 defines the method, and <sign(r0), r0> = ||r0||_1 > 0.
// On the diffusion systems it needs fewer iterations than rd = r0 (100 vs
// 109 per source at 250k nodes, 61 vs 66 at 103k), and its entries are
// exact in any precision.
__device__ inline float sgn3(float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f); }
__device__ inline double sgn3(double v) { return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0); }
__device__ inline float2 shadow_sign(float2 v) { return make_float2(sgn3(v.x), sgn3(v.y)); }
__device__ inline double2 shadow_sign(double2 v) { return make_double2(sgn3(v.x), sgn3(v.y)); }
__device__ inline float shadow_sign(float v) { return sgn3(v); }
__device__ inline double shadow_sign(double v) { return sgn3(v); }

// The shadow residual of a 64-column complex FP32 tile packed to 4 bits per
// entry (2 per real component: 0 -> 0, 1 -> +1, 3 -> -1), 32 bytes per row
// instead of 512: the kernels that stream it every iteration (the dot of
// the first SpMM, the x/r update) read the packed copy, the others the FP32
// one; both hold the same values, so every dot is bit-identical.
__device__ inline unsigned sh_code(float x) { return x > 0.f ? 1u : (x < 0.f ? 3u : 0u); }
__device__ inline float sh_val(unsigned b) { return (b & 1u) ? ((b & 2u) ? -1.f : 1.f) : 0.f; }
__device__ inline float2 sh_unpack(const unsigned char *__restrict__ q, size_t row, int c)
{
    const unsigned nib = (q[row * 32 + (c >> 1)] >> ((c & 1) * 4)) & 15u;
    return make_float2(sh_val(nib & 3u), sh_val(nib >> 2));
}

// q[row][j] = the packed entries (row, 2j) and (row, 2j+1) of rd [n][64]
__global__ void k_pack_shadow(int n, const float2 *__restrict__ rd, unsigned char *__restrict__ q)
{
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= (size_t)n * 32) return;
    const float2 a = rd[2 * b], c = rd[2 * b + 1];
    q[b] = (unsigned char)(sh_code(a.x) | (sh_code(a.y) << 2) | (sh_code(c.x) << 4) | (sh_code(c.y) << 6));
}

template <class V> __device__ inline V vzero();
template <> __device__ inline float2 vzero<float2>() { return make_float2(0.f, 0.f); }
template <> __device__ inline double2 vzero<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ inline float vzero<float>() { return 0.f; }
template <> __device__ inline double vzero<double>() { return 0.0; }
__device__ inline float2 to32(double2 a) { return make_float2((float)a.x, (float)a.y); }
__device__ inline float to32(double a) { return (float)a; }
__device__ inline double2 to64(float2 a) { return make_double2(a.x, a.y); }
__device__ inline double to64(float a) { return a; }
__device__ inline double2 to64(double2 a) { return a; }
__device__ inline double to64(double a) { return a; }
__device__ inline float2 cvt32(double2 a) { return make_float2((float)a.x, (float)a.y); }
__device__ inline float cvt32(double a) { return (float)a; }
template <class To> __device__ inline To cvt(double2 a);
template <> __device__ inline float2 cvt<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }
template <> __device__ inline double2 cvt<double2>(double2 a) { return a; }
template <class To> __device__ inline To cvt(double a);
template <> __device__ inline float cvt<float>(double a) { return (float)a; }
template <> __device__ inline double cvt<double>(double a) { return a; }
__device__ inline float2 shfl_add(float2 a, int off) { a.x += __shfl_xor(a.x, off); a.y += __shfl_xor(a.y, off); return a; }
__device__ inline double2 shfl_add(double2 a, int off) { a.x += __shfl_xor(a.x, off); a.y += __shfl_xor(a.y, off); return a; }
__device__ inline float shfl_add(float a, int off) { return a + __shfl_xor(a, off); }
__device__ inline double shfl_add(double a, int off) { return a + __shfl_xor(a, off); }

constexpr int NXCD = 8;
// neighbour-row gathers in flight per wave and batch in the 64-column and
// narrow-tile SpMM kernels (build-time tuning parameters)
#ifndef TMI_SPMM64_BATCH
#define TMI_SPMM64_BATCH 16
#endif
#ifndef TMI_NK_BATCH
#define TMI_NK_BATCH 16
#endif
// waves per workgroup of the 64-column SpMM (512 threads: 3-4 % faster than
// 256 on MI355X at the same 8 waves per SIMD); the other kernels use 4
constexpr int SPMM64_WAVES = 8;
// workgroup id -> logical id so that XCD x (which receives ids x, x+8, ...)
// owns the contiguous logical range [x*G/8, (x+1)*G/8).  G is a multiple of 8.
__device__ inline int xcd_block(int b, int G) { return (b % NXCD) * (G / NXCD) + b / NXCD; }

// Row schedule of the SpMM kernels: the rows are cut into 8 contiguous
// regions, one per XCD (workgroup b runs on XCD b % 8), and inside a region
// the XCD's waves take rows round-robin.  All waves of an XCD therefore work
// on one moving window of consecutive rows, so the neighbour rows a CSR row
// gathers (i +- 1, i +- layer, i +- slab in a structured numbering) were
// loaded into that XCD's L2 by the window just before.
struct RowSched {
    int start, end, first, stride;
};
__device__ inline RowSched row_sched(int n, int rows_per_wave)
{
    const int G = gridDim.x;
    const int xcd = blockIdx.x % NXCD, bx = blockIdx.x / NXCD, nbx = G / NXCD;
    const int nw = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int per = (n + NXCD - 1) / NXCD;
    RowSched r;
    r.start = min(n, xcd * per);
    r.end = min(n, r.start + per);
    r.first = r.start + (bx * nw + wave) * rows_per_wave;
    r.stride = nbx * nw * rows_per_wave;
    return r;
}

// Reduce two per-lane sums over the lanes that share a column (lane % K) and
// over the waves of the workgroup; write partial[(slot*K + c)*G + blk] (each
// column's G partials contiguous for the finalize).
template <int K>
__device__ inline void block_reduce2(double d1, double d2, double *partial, int blk, int G, bool want2)
{
    for (int off = K; off < 64; off <<= 1) {
        d1 += __shfl_xor(d1, off);
        if (want2) d2 += __shfl_xor(d2, off);
    }
    __shared__ double s1[SPMM64_WAVES][K], s2[SPMM64_WAVES][K];   // up to 8 waves
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane < K) { s1[wave][lane] = d1; s2[wave][lane] = d2; }
    __syncthreads();
    if ((int)threadIdx.x < K) {
        double t1 = 0.0, t2 = 0.0;
        for (int w = 0; w < nw; w++) { t1 += s1[w][threadIdx.x]; t2 += s2[w][threadIdx.x]; }
        partial[((size_t)0 * K + threadIdx.x) * G + blk] = t1;
        if (want2) partial[((size_t)1 * K + threadIdx.x) * G + blk] = t2;
    }
}

// Three sums per column (COCG's complex r^T z and ||r||^2), in the order of
// block_reduce2, into partial slots 0, 1 and 2.
template <int K>
__device__ inline void block_reduce3(double d1, double d2, double d3, double *partial, int blk, int G)
{
    for (int off = K; off < 64; off <<= 1) {
        d1 += __shfl_xor(d1, off);
        d2 += __shfl_xor(d2, off);
        d3 += __shfl_xor(d3, off);
    }
    __shared__ double s1[SPMM64_WAVES][K], s2[SPMM64_WAVES][K], s3[SPMM64_WAVES][K];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane < K) { s1[wave][lane] = d1; s2[wave][lane] = d2; s3[wave][lane] = d3; }
    __syncthreads();
    if ((int)threadIdx.x < K) {
        double t1 = 0.0, t2 = 0.0, t3 = 0.0;
        for (int w = 0; w < nw; w++) { t1 += s1[w][threadIdx.x]; t2 += s2[w][threadIdx.x]; t3 += s3[w][threadIdx.x]; }
        partial[((size_t)0 * K + threadIdx.x) * G + blk] = t1;
        partial[((size_t)1 * K + threadIdx.x) * G + blk] = t2;
        partial[((size_t)2 * K + threadIdx.x) * G + blk] = t3;
    }
}

// ----------------------------------------------------------- block SpMM
// Gathers are issued in chunks of CH nonzeros (indices and values first,
// then all CH neighbour loads, then the multiply-adds) so that a wave keeps
// CH independent memory requests in flight instead of one per nonzero.
#ifndef TMI_DOT_CH
#define TMI_DOT_CH 16
#endif
constexpr int CH = TMI_DOT_CH;

template <class V, int K, bool PRE, bool MF = false>
__device__ inline V row_gather(int kb, int ke, int c, const int *__restrict__ colidx, const V *__restrict__ val,
                               const V *__restrict__ idiag, const V *__restrict__ X,
                               typename ScalarOf<V>::type w = 1)
{
    V acc = vzero<V>();
    for (int k0 = kb; k0 < ke; k0 += CH) {
        int jj[CH];
        V aa[CH], xx[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            // tail of the row: repeat the last nonzero with a zero weight so
            // every load stays inside the row
            const int kk = min(k0 + u, ke - 1);
            jj[u] = colidx[kk];
            aa[u] = (k0 + u < ke) ? val[kk] : vzero<V>();
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            xx[u] = X[(size_t)jj[u] * K + c];
            if (PRE) xx[u] = vmul(idiag[jj[u]], xx[u]);
        }
#pragma unroll
        for (int u = 0; u < CH; u++) acc = vmac(acc, mfc<MF>(aa[u], w), xx[u]);
    }
    return acc;
}

template <class V, int K, bool PRE, bool MF = false>
__global__ void __launch_bounds__(256) k_spmm_dot(int n, const int *__restrict__ rowptr,
                                                  const int *__restrict__ colidx, const V *__restrict__ val,
                                                  const V *__restrict__ idiag, const V *__restrict__ X,
                                                  V *__restrict__ Y, const V *__restrict__ W, int want_yy,
                                                  const int *__restrict__ active, double *__restrict__ partial,
                                                  const double *__restrict__ omg = nullptr)
{
    constexpr int R = 64 / K;
    const int G = gridDim.x;
    const int blk = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int c = lane % K, rsub = lane / K;
    const bool act = active[c] != 0;
    const RowSched rs = row_sched(n, R);
    double d1 = 0.0, d2 = 0.0;
    for (int base = rs.first; base < rs.end; base += rs.stride) {
        const int i = base + rsub;
        if (i < rs.end && act) {
            V acc = row_gather<V, K, PRE, MF>(rowptr[i], rowptr[i + 1], c, colidx, val, idiag, X,
                                              mf_omega<MF, V>(omg, c));
            Y[(size_t)i * K + c] = acc;
            dot_epi(W[(size_t)i * K + c], acc, want_yy, d1, d2);
        }
    }
    block_reduce2<K>(d1, d2, partial, blk, G, want_yy != 0);
}

// ------------------------------------------- block SpMM, 64 columns
// K = 64: one lane per column, one wave per contiguous run of rows.  The
// row pointers, column indices and values of the run are fetched with
// coalesced vector loads (64 entries per load) and broadcast with readlane,
// so the only memory accesses on the critical path are the neighbour-row
// gathers, issued B at a time.  Each XCD owns one contiguous region of
// rows; inside it, R > 0 deals runs of R rows round-robin over the XCD's
// waves (the waves of an XCD then sweep the region together, so their
// neighbour rows overlap in that XCD's L2), R <= 0 gives every wave one
// contiguous share of the region, processed in runs of up to 63 rows.
// a[i] with a wave-uniform row pointer a kept in scalar registers (the
// load then takes the scalar-base + lane-offset form, no per-lane 64-bit
// address arithmetic)
// (loads go through integer types, bit-cast to the element type)
typedef unsigned long long ldrow_u2 __attribute__((ext_vector_type(2)));
template <class V>
__device__ inline V ld_row(const V *a, unsigned i)
{
    static_assert(sizeof(V) == 8 || sizeof(V) == 16, "ld_row: 8- or 16-byte elements");
    const unsigned long long p = (unsigned long long)a;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)p);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(p >> 32));
    const unsigned long long u = ((unsigned long long)hi << 32) | lo;
    if constexpr (sizeof(V) == 8)
        return __builtin_bit_cast(V, ((const __attribute__((address_space(1))) unsigned long long *)u)[i]);
    else
        return __builtin_bit_cast(V, ((const __attribute__((address_space(1))) ldrow_u2 *)u)[i]);
}

__device__ inline float2 rdlane(float2 v, int l)
{
    return make_float2(__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v.x), l)),
                       __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v.y), l)));
}
__device__ inline double rdlane_d(double v, int l)
{
    long long b = __builtin_bit_cast(long long, v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double2 rdlane(double2 v, int l) { return make_double2(rdlane_d(v.x, l), rdlane_d(v.y, l)); }
__device__ inline float rdlane(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ inline double rdlane(double v, int l) { return rdlane_d(v, l); }

template <class V, int MODE, bool MF = false>   // MODE 0: Y = A X, dots <W,Y> [, <Y,Y>]; MODE 1: Y = W - A X, dot <Y,Y>
__global__ void __launch_bounds__(64 * SPMM64_WAVES) k_spmm64(int n, const int *__restrict__ rowptr, const int *__restrict__ colidx,
                                                const V *__restrict__ val, const V *__restrict__ X,
                                                V *__restrict__ Y, const V *__restrict__ W, int want_yy,
                                                const int *__restrict__ active, double *__restrict__ partial, int R,
                                                const double *__restrict__ omg = nullptr)
{
    constexpr int B = TMI_SPMM64_BATCH;
    const int lane = threadIdx.x & 63;
    const int c = lane;
    const bool act = active[c] != 0;
    const auto wmf = mf_omega<MF, V>(omg, c);
    const int G = gridDim.x, nw = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int xcd = blockIdx.x % NXCD, bx = blockIdx.x / NXCD, nbx = G / NXCD;
    const int per_x = (n + NXCD - 1) / NXCD;
    const int xs = min(n, xcd * per_x), xe = min(n, xs + per_x);
    const int wpx = nbx * nw, wid = bx * nw + wave;
    int rstart, rend_all, rstep, rlen;
    if (R > 0) {
        rlen = min(R, 63);
        rstart = xs + wid * rlen;
        rend_all = xe;
        rstep = wpx * rlen;
    } else {
        const int per_w = (xe - xs + wpx - 1) / wpx;
        rstart = min(xe, xs + wid * per_w);
        rend_all = min(xe, rstart + per_w);
        rstep = rlen = 63;
    }
    double d1 = 0.0, d2 = 0.0;
    for (int rb = rstart; rb < rend_all; rb += rstep) {
        const int nr = min(rlen, rend_all - rb);
        const int rpv = rowptr[rb + min(lane, nr)];
        const int kb = __builtin_amdgcn_readlane(rpv, 0), ke = __builtin_amdgcn_readlane(rpv, nr);
        int row = 0;
        int rend = __builtin_amdgcn_readlane(rpv, 1);
        V acc = vzero<V>();
        auto flush = [&]() {
            const size_t at = (size_t)(rb + row) * 64 + c;
            if (act) {
                if (MODE == 0) {
                    Y[at] = acc;
                    dot_epi(W[at], acc, want_yy, d1, d2);
                } else {
                    V r = vsub(W[at], acc);
                    Y[at] = r;
                    d1 += dotR(r, r);
                }
            }
            acc = vzero<V>();
            row++;
            rend = __builtin_amdgcn_readlane(rpv, min(row + 1, nr));
        };
        for (int seg = kb; seg < ke; seg += 64) {
            const int kk = min(seg + lane, ke - 1);
            const int cv = colidx[kk];
            const V vv = val[kk];
            const int m = min(64, ke - seg);
            for (int t0 = 0; t0 < m; t0 += B) {
                V xb[B];
#pragma unroll
                for (int u = 0; u < B; u++) {
                    const int t = min(t0 + u, m - 1);
                    const int j = __builtin_amdgcn_readlane(cv, t);
                    xb[u] = act ? X[(size_t)j * 64 + c] : vzero<V>();
                }
#pragma unroll
                for (int u = 0; u < B; u++) {
                    if (t0 + u < m) {
                        const int k = seg + t0 + u;
                        while (k >= rend && row < nr) flush();
                        acc = vmac(acc, mfc<MF>(rdlane(vv, t0 + u), wmf), xb[u]);
                    }
                }
            }
        }
        while (row < nr) flush();
    }
    block_reduce2<64>(d1, d2, partial, blockIdx.x, G, MODE == 0 && want_yy != 0);
}

// ------------------------------------ block SpMM, K < 64 columns
// 64 / K row streams per wave: stream s (lanes sK .. sK+K-1, lane = column)
// works through a run of RS consecutive rows as one flat list of nonzeros.
// Row pointers of the wave's S*RS rows come with one coalesced load; each
// stream loads the column indices and values of its next K nonzeros with one
// coalesced load and hands them to its lanes with lane shuffles, so the
// memory accesses per nonzero step are the one gather of the neighbour row
// (K values per stream, S rows per instruction).  Shuffles run with every
// lane active; only the row flushes branch per stream.  Each row accumulates
// acc + val * x in ascending order, like the 64-column kernel.
template <class V>
__device__ inline V shfl_v(V v, int src);
template <>
__device__ inline float2 shfl_v<float2>(float2 v, int src)
{
    return make_float2(__shfl(v.x, src), __shfl(v.y, src));
}
template <>
__device__ inline double2 shfl_v<double2>(double2 v, int src)
{
    return make_double2(__shfl(v.x, src), __shfl(v.y, src));
}
template <>
__device__ inline float shfl_v<float>(float v, int src) { return __shfl(v, src); }
template <>
__device__ inline double shfl_v<double>(double v, int src) { return __shfl(v, src); }

constexpr int NK_RSMAX = 4;   // rows per stream per run (at most)
// K of the narrow kernel for a dispatch over all tile widths (the 1- and
// 64-column cases never reach it).  Below NK_KMIN columns the per-lane
// gather kernel is faster on MI355X (K = 4: 23 us against 46 us per launch
// at 250k rows: with 16 streams per wave the per-stream row flushes diverge
// at almost every nonzero).
constexpr int NK_KMIN = 16;
__host__ __device__ constexpr int nk_k(int K) { return K >= NK_KMIN && K <= 32 ? K : NK_KMIN; }
__host__ __device__ constexpr int nk_rows(int K) { return (64 / K) * NK_RSMAX <= 63 ? NK_RSMAX : 63 / (64 / K); }

template <class V, int K, int MODE, bool MF = false>   // MODE 0: Y = A X, dots <W,Y> [, <Y,Y>]; MODE 1: Y = W - A X, dot <Y,Y>
__global__ void __launch_bounds__(256) k_spmm_nk(int n, const int *__restrict__ rowptr, const int *__restrict__ colidx,
                                                 const V *__restrict__ val, const V *__restrict__ X,
                                                 V *__restrict__ Y, const V *__restrict__ W, int want_yy,
                                                 const int *__restrict__ active, double *__restrict__ partial,
                                                 const double *__restrict__ omg = nullptr)
{
    constexpr int S = 64 / K, RS = nk_rows(K), RW = S * RS, B = K < TMI_NK_BATCH ? K : TMI_NK_BATCH;
    static_assert(K >= 2 && K <= 32 && RW <= 63, "narrow SpMM: 2 <= K <= 32");
    const int lane = threadIdx.x & 63, c = lane % K, s = lane / K, sb = s * K;
    const bool act = active[c] != 0;
    const auto wmf = mf_omega<MF, V>(omg, c);
    const int G = gridDim.x, nw = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int xcd = blockIdx.x % NXCD, bx = blockIdx.x / NXCD, nbx = G / NXCD;
    const int per_x = (n + NXCD - 1) / NXCD;
    const int xs = min(n, xcd * per_x), xe = min(n, xs + per_x);
    const int wpx = nbx * nw, wid = bx * nw + wave;
    double d1 = 0.0, d2 = 0.0;
    for (int rb = xs + wid * RW; rb < xe; rb += wpx * RW) {
        const int nrun = min(RW, xe - rb);
        const int rpv = rowptr[rb + min(lane, nrun)];   // row pointers of the run
        // stream s: rows rb + s RS .. + nr (nr may be 0 at the region's end)
        const int f = min(nrun, s * RS), nr = min(nrun, f + RS) - f;
        int rpa[RS + 1];
#pragma unroll
        for (int i = 0; i <= RS; i++) rpa[i] = __shfl(rpv, min(f + i, f + nr));
        const int kb = rpa[0], len = rpa[RS] - kb;   // rpa[i] for i > nr repeats the end
        int row = 0;
        auto rend_of = [&](int r) {
            int e = rpa[RS];
#pragma unroll
            for (int i = RS; i >= 1; i--)
                if (r + 1 == i) e = rpa[i];
            return e;
        };
        int rend = rend_of(0);
        V acc = vzero<V>();
        auto flush = [&]() {
            const size_t at = (size_t)(rb + f + row) * K + c;
            if (act) {
                if (MODE == 0) {
                    Y[at] = acc;
                    dot_epi(W[at], acc, want_yy, d1, d2);
                } else {
                    V r = vsub(W[at], acc);
                    Y[at] = r;
                    d1 += dotR(r, r);
                }
            }
            acc = vzero<V>();
            row++;
            rend = rend_of(row);
        };
        for (int o = 0; __any(o < len); o += K) {
            const int m = min(K, len - o);   // nonzeros of this segment (<= 0: stream done)
            const int kk = max(0, min(kb + o + c, kb + len - 1));
            const int cv = colidx[kk];
            const V vv = val[kk];
            for (int t0 = 0; t0 < K && __any(t0 < m); t0 += B) {
                V xb[B], ab[B];
#pragma unroll
                for (int u = 0; u < B; u++) {
                    const int src = sb + max(0, min(t0 + u, m - 1));
                    const int j = __shfl(cv, src);
                    ab[u] = shfl_v(vv, src);
                    xb[u] = X[(size_t)j * K + c];
                }
#pragma unroll
                for (int u = 0; u < B; u++) {
                    if (t0 + u < m) {
                        const int k = kb + o + t0 + u;
                        while (k >= rend && row < nr) flush();
                        acc = vmac(acc, mfc<MF>(ab[u], wmf), xb[u]);
                    }
                }
            }
        }
        while (row < nr) flush();
    }
    block_reduce2<K>(d1, d2, partial, blockIdx.x, G, MODE == 0 && want_yy != 0);
}

// ------------------------------- row-block SpMM, 64 columns
// One wave per block of RB_ROWS consecutive rows (blocks of an XCD's region
// dealt round-robin to its waves).  The block's rows share most neighbours
// (Morton order): the wave walks the ascending union of the block's columns
// and gathers each neighbour row of X once for all block rows holding it
// (the column's row mask says which), instead of once per nonzero -- 54 % of
// the gathers at 8 rows per block on the 250k-node tetrahedral mesh.  Column
// indices and masks are loaded 64 at a time, coefficients (stored per block
// in the same column-then-row order) through a 64-entry window that is
// refilled from a prefetched next window, all broadcast with readlane.  Each
// row accumulates acc + val * x over its columns in ascending order, the
// same operation sequence as the per-row kernels.  KC = 32: tiles of 32
// columns, lanes 32-63 idle (the walk costs the same per union entry, so
// this beats the per-row 32-column kernel once the blocks share gathers).
template <class V, int MODE, bool MF = false, int KC = 64>   // MODE 0: Y = A X, dots <W,Y> [, <Y,Y>]; MODE 1: Y = W - A X, dot <Y,Y>
__global__ void __launch_bounds__(64 * SPMM64_WAVES) k_spmm_rb(int n, const int *__restrict__ rowptr,
                                                           const int *__restrict__ bofs, const int *__restrict__ ucol,
                                                           const rb_mask_t *__restrict__ umask,
                                                           const V *__restrict__ rbval, const V *__restrict__ X,
                                                           V *__restrict__ Y, const V *__restrict__ W, int want_yy,
                                                           const int *__restrict__ active, double *__restrict__ partial,
                                                           const unsigned char *__restrict__ Wq,
                                                           const double *__restrict__ omg = nullptr)
{
    constexpr int B = TMI_RB_BATCH;
    static_assert(B * RB_ROWS <= 64, "a batch's coefficients must fit one 64-lane window");
    static_assert(KC == 64 || KC == 32, "row-block tiles of 64 or 32 columns");
    const int lane = threadIdx.x & 63;
    const bool live = KC == 64 || lane < KC;
    const int c = live ? lane : 0;   // idle lanes load column 0 (in bounds) and store nothing
    const bool act = live && active[c] != 0;
    const auto wmf = mf_omega<MF, V>(omg, c);
    const int G = gridDim.x, nw = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int xcd = blockIdx.x % NXCD, bx = blockIdx.x / NXCD, nbx = G / NXCD;
    const int nb = (n + RB_ROWS - 1) / RB_ROWS;
    const int per_x = (nb + NXCD - 1) / NXCD;
    const int xs = min(nb, xcd * per_x), xe = min(nb, xs + per_x);
    const int wpx = nbx * nw, wid = bx * nw + wave;
    double d1 = 0.0, d2 = 0.0;
    for (int b = xs + wid; b < xe; b += wpx) {
        const int u0 = bofs[b], u1 = bofs[b + 1];
        const int r0 = b * RB_ROWS, nr = min(RB_ROWS, n - r0);
        const int kb = rowptr[r0], ke = rowptr[r0 + nr];
        int kc = kb;   // the block's next coefficient (they follow the union entries)
        V acc[RB_ROWS];
#pragma unroll
        for (int r = 0; r < RB_ROWS; r++) acc[r] = vzero<V>();
        for (int seg = u0; seg < u1; seg += 64) {
            const int m = min(64, u1 - seg);
            const int uu = min(seg + lane, u1 - 1);
            const int jv = ucol[uu];
            const int mv = umask[uu];
            for (int t0 = 0; t0 < m; t0 += B) {
                // a batch of B <= 8 entries holds at most 8 B <= 64
                // coefficients: one window load, issued with the gathers
                const V cw = ld_row(rbval + kc, (unsigned)min(lane, ke - 1 - kc));
                int kk = 0;
                V xb[B];
#pragma unroll
                for (int u = 0; u < B; u++) {
                    // unconditional: j is a column of the block (in bounds);
                    // an inactive column's sums are never stored
                    const int j = __builtin_amdgcn_readlane(jv, min(t0 + u, m - 1));
                    xb[u] = ld_row(X + (size_t)j * KC, (unsigned)c);
                }
#pragma unroll
                for (int u = 0; u < B; u++) {
                    if (t0 + u < m) {
                        const int mk = __builtin_amdgcn_readlane(mv, t0 + u);
#pragma unroll
                        for (int r = 0; r < RB_ROWS; r++) {
                            if (mk & (1 << r)) {
                                acc[r] = vmac(acc[r], mfc<MF>(rdlane(cw, kk), wmf), xb[u]);
                                kk++;
                            }
                        }
                    }
                }
                kc += kk;
            }
        }
#pragma unroll
        for (int r = 0; r < RB_ROWS; r++) {
            if (r < nr && act) {
                const size_t at = (size_t)(r0 + r) * KC + c;
                if constexpr (MODE == 0) {
                    Y[at] = acc[r];
                    V w;
                    if constexpr (std::is_same<V, float2>::value)
                        w = (KC == 64 && Wq) ? sh_unpack(Wq, (size_t)(r0 + r), c) : W[at];   // Wq: packed shadow residual
                    else
                        w = W[at];
                    dot_epi(w, acc[r], want_yy, d1, d2);
                } else {
                    const V y = vsub(W[at], acc[r]);
                    Y[at] = y;
                    d1 += dotR(y, y);
                }
            }
        }
    }
    block_reduce2<KC>(d1, d2, partial, blockIdx.x, G, MODE == 0 && want_yy != 0);
}

// ------------------------- pipelined row-block SpMM, 64 columns (variant 4)
// The row-block kernel's arithmetic (each row sums val * x over its columns
// in ascending order, so Y is bit-identical to it) with its memory traffic
// restructured for latency hiding inside each wave:
//  * a batch is up to RBP_B distinct columns of one block; its loads are the
//    RBP_B neighbour-row gathers plus ONE 64-lane load of the batch's
//    coefficients (at most RBP_B x RB_ROWS = 64, contiguous: the block's
//    coefficients are stored column then row), located by the packed word
//    of its first column (row mask | offset of the column's first
//    coefficient in the block << 16);
//  * column indices and words come through the scalar cache (uniform
//    addresses), so they never wait behind the vector gathers;
//  * batches are double-buffered: the gathers of batch t+1 -- including the
//    first batch of the wave's next block -- are in flight while batch t is
//    summed, and every batch issues the same number of loads (indices past
//    the block are clamped), so each wait covers only the batch it needs;
//  * the packed shadow residual of a block's rows is loaded with the block's
//    first batch, ahead of its use in the epilogue.
#ifndef TMI_RBP_B
#define TMI_RBP_B 8
#endif
constexpr int RBP_B = TMI_RBP_B;
static_assert(RBP_B * RB_ROWS <= 64, "a batch's coefficients must fit one 64-lane window");
static_assert(RBP_B + 1 < 64, "s_waitcnt counts at most 63 vector loads");
// a wave-uniform pointer kept in scalar registers: both halves pass through
// readfirstlane, so the optimizer cannot fold a per-lane offset into it and
// loads p[lane] use the scalar-base + 32-bit lane-offset addressing
// (global address space, so that the loads stay global loads).  Loads go
// through integer types, bit-cast to the element type.
template <class I>
__device__ inline const __attribute__((address_space(1))) I *uniform_gptr(const void *p)
{
    const unsigned long long a = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    return (const __attribute__((address_space(1))) I *)(((unsigned long long)hi << 32) | lo);
}
__device__ inline float2 ld_row_f2(const float2 *row, unsigned i)
{
    return __builtin_bit_cast(float2, uniform_gptr<unsigned long long>(row)[i]);
}
__device__ inline unsigned ld_row_u8(const unsigned char *row, unsigned i)
{
    return uniform_gptr<unsigned char>(row)[i];
}
// s_waitcnt immediate (gfx9 family) waiting until at most n vector-memory
// operations are outstanding; export and LGKM counts are left unconstrained
constexpr unsigned waitcnt_vm(unsigned n) { return (n & 0xfu) | (0x7u << 4) | (0xfu << 8) | ((n >> 4) << 14); }

struct RbpBatch {
    int b;        // block (-1: none)
    int p, cnt;   // first distinct column (index into the union arrays) and count
    int u1;       // end of the block's union
    int kb, ke;   // the block's coefficient range
};

// the wave's next batch after (blk, p): the rest of the block, else the
// first batch of the wave's next block (b + step), else none
__device__ inline RbpBatch rbp_batch(int b, int p, int xe, int step, int n, const int *__restrict__ bofs,
                                     const int *__restrict__ rowptr, RbpBatch cur)
{
    RbpBatch nx;
    if (b >= 0 && p < cur.u1) {
        nx = cur;
        nx.p = p;
    } else {
        const int nb = b < 0 ? b : b + step;
        nx.b = (b >= 0 && nb < xe) ? nb : -1;
        if (nx.b < 0) {
            nx.p = nx.u1 = nx.kb = nx.ke = 0;
            nx.cnt = 0;
            return nx;
        }
        nx.b = __builtin_amdgcn_readfirstlane(nx.b);
        nx.p = bofs[nx.b];
        nx.u1 = bofs[nx.b + 1];
        const int r0 = nx.b * RB_ROWS, nr = min(RB_ROWS, n - r0);
        nx.kb = rowptr[r0];
        nx.ke = rowptr[r0 + nr];
    }
    nx.cnt = min(RBP_B, nx.u1 - nx.p);
    return nx;
}

template <int MODE, bool PACKED>   // MODE 0: Y = A X, dots <W,Y> [, <Y,Y>]; PACKED: W given as Wq
__global__ void __launch_bounds__(64 * SPMM64_WAVES) k_spmm_rbp(int n, const int *__restrict__ rowptr,
                                                            const int *__restrict__ bofs, const int *__restrict__ ucol,
                                                            const unsigned *__restrict__ uword,
                                                            const float2 *__restrict__ rbval,
                                                            const float2 *__restrict__ X, float2 *__restrict__ Y,
                                                            const float2 *__restrict__ W, int want_yy,
                                                            const int *__restrict__ active,
                                                            double *__restrict__ partial,
                                                            const unsigned char *__restrict__ Wq)
{
    constexpr int B = RBP_B;
    const int lane = threadIdx.x & 63, c = lane;
    const bool act = active[c] != 0;
    const int G = gridDim.x, nw = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int xcd = blockIdx.x % NXCD, bx = blockIdx.x / NXCD, nbx = G / NXCD;
    const int nb = (n + RB_ROWS - 1) / RB_ROWS;
    const int per_x = (nb + NXCD - 1) / NXCD;
    const int xs = min(nb, xcd * per_x), xe = min(nb, xs + per_x);
    const int wpx = nbx * nw, wid = bx * nw + wave;
    double d1 = 0.0, d2 = 0.0;

    float2 xa[B], xn[B];   // gathered rows of the batch being summed / in flight
    float2 ca, cn;         // their coefficient windows (lane l: coefficient l of the batch)
    unsigned wa[B], wn[B]; // packed words
    float2 wr[RB_ROWS];    // W of the rows of a finishing block
    unsigned qr[RB_ROWS];  // ... or its packed bytes (decoded where used)
    float2 acc[RB_ROWS];
#pragma unroll
    for (int r = 0; r < RB_ROWS; r++) acc[r] = vzero<float2>();
#pragma unroll
    for (int u = 0; u < B; u++) xa[u] = xn[u] = vzero<float2>();

    // issue the loads of batch t into (x, cw, w).  Every batch issues the
    // same B + 1 vector loads: the words and column indices of B consecutive
    // union entries (one scalar load each; entries past the batch belong to
    // the next block or to the arrays' 64-entry padding and are valid rows
    // whose values are not used), the B gathered rows for all lanes, and the
    // coefficient window.  Nothing is conditional and no register is reset,
    // so each batch's wait covers exactly the loads it needs.
    auto issue = [&](const RbpBatch &t, float2 *x, float2 &cw, unsigned *w) {
        const int p = t.p;
        int j[B];
#pragma unroll
        for (int u = 0; u < B; u++) {
            w[u] = uword[p + u];
            j[u] = ucol[p + u];
        }
        // uniform row address + lane offset: scalar-base loads, no 64-bit
        // per-lane address registers
#pragma unroll
        for (int u = 0; u < B; u++) x[u] = ld_row_f2(X + (size_t)j[u] * 64, c);
        const int base = __builtin_amdgcn_readfirstlane(t.kb + (int)(w[0] >> 16));
        cw = ld_row_f2(rbval + base, min(lane, max(t.ke - 1 - base, 0)));
    };
    // W (or its packed shadow form) of the rows of batch t's block, if t is
    // the block's last batch: loaded before the next batch's gathers, used
    // after them
    auto load_w = [&](const RbpBatch &t) {
        if (t.p + t.cnt >= t.u1) {
            const int r0 = t.b * RB_ROWS;
#pragma unroll
            for (int r = 0; r < RB_ROWS; r++) {
                const int row = min(r0 + r, n - 1);
                if (PACKED) qr[r] = ld_row_u8(Wq + (size_t)row * 32, c >> 1);
                else wr[r] = ld_row_f2(W + (size_t)row * 64, c);
            }
        }
    };
    // sum batch t (buffers x, cw, w); at the block's last batch write its rows
    auto sum = [&](const RbpBatch &t, const float2 *x, const float2 &cw, const unsigned *w) {
        // this batch's loads (and a finishing block's W) are all older than
        // the next batch's B + 1 vector loads
        __builtin_amdgcn_s_waitcnt(waitcnt_vm(B + 1));
        const int off0 = (int)(w[0] >> 16);
#pragma unroll
        for (int u = 0; u < B; u++) {
            if (u < t.cnt) {
                const unsigned mk = w[u] & 0xffffu;
                int k = (int)(w[u] >> 16) - off0;
#pragma unroll
                for (int r = 0; r < RB_ROWS; r++) {
                    if (mk & (1u << r)) {
                        acc[r] = vmac(acc[r], rdlane(cw, k), x[u]);
                        k++;
                    }
                }
            }
        }
        if (t.p + t.cnt >= t.u1) {   // the block is complete
            const int r0 = t.b * RB_ROWS, nr = min(RB_ROWS, n - r0);
#pragma unroll
            for (int r = 0; r < RB_ROWS; r++) {
                if (r < nr && act) {
                    Y[(size_t)(r0 + r) * 64 + c] = acc[r];
                    float2 wv = wr[r];
                    if (PACKED) {
                        const unsigned s = (qr[r] >> ((c & 1) * 4)) & 0xfu;
                        wv = make_float2(sh_val(s & 3u), sh_val(s >> 2));
                    }
                    dot_epi(wv, acc[r], want_yy, d1, d2);
                }
                acc[r] = vzero<float2>();
            }
        }
    };

    RbpBatch ta;
    ta.b = -1;
    ta.p = ta.cnt = ta.u1 = ta.kb = ta.ke = 0;
    {   // the wave's first block
        const int b0 = xs + wid;
        if (b0 < xe) {
            ta.b = __builtin_amdgcn_readfirstlane(b0);
            ta.p = bofs[ta.b];
            ta.u1 = bofs[ta.b + 1];
            const int r0 = ta.b * RB_ROWS, nr = min(RB_ROWS, n - r0);
            ta.kb = rowptr[r0];
            ta.ke = rowptr[r0 + nr];
            ta.cnt = min(B, ta.u1 - ta.p);
        }
    }
    issue(ta, xa, ca, wa);
    while (ta.b >= 0) {
        const RbpBatch tn = rbp_batch(ta.b, ta.p + ta.cnt, xe, wpx, n, bofs, rowptr, ta);
        load_w(ta);
        issue(tn, xn, cn, wn);
        sum(ta, xa, ca, wa);
        if (tn.b < 0) break;
        ta = rbp_batch(tn.b, tn.p + tn.cnt, xe, wpx, n, bofs, rowptr, tn);
        load_w(tn);
        issue(ta, xa, ca, wa);
        sum(tn, xn, cn, wn);
    }
    block_reduce2<64>(d1, d2, partial, blockIdx.x, G, MODE == 0 && want_yy != 0);
}

// --------------------- row-block SpMM on the matrix cores (variant 5)
// The block of RB_ROWS = 8 rows times the neighbour rows of its column
// union is a small dense product: Y_b (8 x KC, complex) = C_b X_b, with C_b
// the 8 x U coefficients (zero where a row lacks the column) and X_b the U
// gathered rows of X.  In real form, with [Y_re; Y_im] as one 16-row
// matrix,
//     [Y_re; Y_im] += [C_re; C_im] X_re + [-C_im; C_re] X_im,
// which is two v_mfma_f32_16x16x4_f32 per 4 union entries and 16-column
// tile (KC = 16 NT).  Operand layout (lane l):
//  * A: row m = l % 16 of the 16 (m < 8: real part of block row m, else the
//    imaginary part of row m - 8), entry k = l / 16 of the group; the
//    second product's A is the first's rotated by 8 lanes inside each
//    16-lane row (DPP), the real-part rows negated;
//  * B: entry l / 16, tile column l % 16; tile t's column n is the tile's
//    column NT n + t, so lane l needs columns NT (l % 16) .. + NT - 1 of the
//    neighbour row, NT x 8 contiguous bytes (16-byte loads; 16 lanes read
//    the whole 512-byte row of a 64-column tile);
//  * D: register v holds row 4 (l / 16) + v of the 16, tile column l % 16.
// Vector-memory instructions are what bound the row-block kernels, so the
// loads are wide: per group of 4 entries NT / 2 16-byte gathers (one
// 8-byte one when NT = 1), per batch of 8 entries one 64-lane window of the
// block's coefficients handed to the lanes with ds_bpermute, and in the
// epilogue the real- and imaginary-part lanes (l, l ^ 32) swap halves so
// that each stores and loads whole complex numbers, 16 bytes a lane.
// The coefficients come from the block's (column, row) ordered list at
// offset (word >> 16) + popcount of the column's mask bits below the row;
// the words are the pipelined kernel's.  The sums run in the matrix cores'
// order, not the per-row kernels': results agree to rounding, not bit for
// bit.
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
template <int NT, bool PACKED>   // NT 16-column tiles; W given as Wq when PACKED
__global__ void __launch_bounds__(64 * SPMM64_WAVES) k_spmm_rbm(int n, const int *__restrict__ rowptr,
                                                            const int *__restrict__ bofs, const int *__restrict__ ucol,
                                                            const unsigned *__restrict__ uword,
                                                            const float2 *__restrict__ rbv, const float2 *__restrict__ X,
                                                            float2 *__restrict__ Y, const float2 *__restrict__ W,
                                                            int want_yy, const int *__restrict__ active,
                                                            double *__restrict__ partial,
                                                            const unsigned char *__restrict__ Wq)
{
    static_assert(RB_ROWS == 8, "the 16-row operand holds 8 block rows, real and imaginary parts");
    static_assert(NT == 1 || NT == 2 || NT == 4, "tiles of 16, 32 or 64 columns");
    constexpr int KC = 16 * NT;
    constexpr int MB = 2;                  // groups per batch: 8 entries, at most 64 coefficients
    constexpr int NC = NT > 1 ? NT / 2 : 1;   // complex columns per lane in the epilogue
    const int lane = threadIdx.x & 63;
    const int m = lane & 15, q = lane >> 4;
    const int rr = m & 7, part = m >> 3;            // this lane's A row: block row, real/imaginary part
    const unsigned below = (1u << rr) - 1u;
    const float sgn = part ? 1.f : -1.f;            // second A operand: sgn * (first A of lane l ^ 8)
    const int dpart = q >> 1, drow0 = 4 * (q & 1);  // D rows 4q..4q+3: part, first block row
    // epilogue columns of this lane: NC consecutive from cb (NT = 1: the
    // real-part lane holds the column, its partner none)
    const int cb = NT > 1 ? NT * m + dpart * NC : m;
    const bool owner = NT > 1 || dpart == 0;
    bool act[NC];
#pragma unroll
    for (int e = 0; e < NC; e++) act[e] = owner && active[cb + e] != 0;
    const int G = gridDim.x, nw = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int xcd = blockIdx.x % NXCD, bx = blockIdx.x / NXCD, nbx = G / NXCD;
    const int nb = (n + RB_ROWS - 1) / RB_ROWS;
    const int per_x = (nb + NXCD - 1) / NXCD;
    const int xs = min(nb, xcd * per_x), xe = min(nb, xs + per_x);
    const int wpx = nbx * nw, wid = bx * nw + wave;
    double d1[NC], d2[NC];
#pragma unroll
    for (int e = 0; e < NC; e++) d1[e] = d2[e] = 0.0;
    // column indices and words of entries s0 .. s0 + 63 (past s1: the last
    // entry's, every load in bounds and unconditional; the words of entries
    // past the union are dropped where they are used, not here, so that
    // nothing waits on these loads until then)
    auto meta = [&](int s0, int s1, int &jv, unsigned &wv) {
        const int u = s0 + min(lane, s1 - s0 - 1);
        jv = ucol[u];
        wv = uword[u];
    };
    struct Batch {
        float2 cw;
        float2 xv[MB][NT];
        int widx[MB];
        unsigned hasb;
    };
    // the loads of the batch at entry g of a segment of cnt entries (column
    // indices and words jv, wv; the block's coefficients [kb, ke))
    auto issue = [&](int g, int cnt, int jv, unsigned wv, int kb, int ke, Batch &B) {
        // the batch's coefficients: at most 8 x 8 from the first entry's
        const int w0 = kb + (int)(__builtin_amdgcn_readlane(wv, g) >> 16);
        B.cw = rbv[min(w0 + lane, ke - 1)];
        B.hasb = 0;
#pragma unroll
        for (int gi = 0; gi < MB; gi++) {
            // entries past the union: its last column, no rows (in bounds)
            const int idx = g + 4 * gi + q;   // <= 63: g <= 64 - 4 MB
            // (both permutes with every lane active: a lane left out of a
            // ds_bpermute provides no data to the lanes reading from it)
            const int j = __builtin_amdgcn_ds_bpermute(4 * idx, jv);
            const unsigned wraw = (unsigned)__builtin_amdgcn_ds_bpermute(4 * idx, (int)wv);
            const unsigned w = wraw & (0u - (unsigned)(idx < cnt));   // no rows past the union
            const unsigned mk = w & 0xffu;
            const unsigned has = (mk >> rr) & 1u;
            B.hasb |= has << gi;
            // window lane of this lane's coefficient (0 when absent)
            B.widx[gi] = has ? kb + (int)(w >> 16) + __popc(mk & below) - w0 : 0;
            const float2 *xr = X + (size_t)j * KC + NT * m;
            if constexpr (NT == 4) {
                const float4 v0 = *reinterpret_cast<const float4 *>(xr);
                const float4 v1 = *reinterpret_cast<const float4 *>(xr + 2);
                B.xv[gi][0] = make_float2(v0.x, v0.y);
                B.xv[gi][1] = make_float2(v0.z, v0.w);
                B.xv[gi][2] = make_float2(v1.x, v1.y);
                B.xv[gi][3] = make_float2(v1.z, v1.w);
            } else if constexpr (NT == 2) {
                const float4 v0 = *reinterpret_cast<const float4 *>(xr);
                B.xv[gi][0] = make_float2(v0.x, v0.y);
                B.xv[gi][1] = make_float2(v0.z, v0.w);
            } else {
                B.xv[gi][0] = xr[0];
            }
        }
    };
    mfma_f4 acc[NT];
    // (a group wholly past the union has zero coefficients: its products add
    // zeros; not branching over it keeps its loads ahead of the barrier)
    auto compute = [&](const Batch &B) {
#pragma unroll
        for (int gi = 0; gi < MB; gi++) {
            const float cre = __builtin_bit_cast(
                float, __builtin_amdgcn_ds_bpermute(4 * B.widx[gi], __builtin_bit_cast(int, B.cw.x)));
            const float cim = __builtin_bit_cast(
                float, __builtin_amdgcn_ds_bpermute(4 * B.widx[gi], __builtin_bit_cast(int, B.cw.y)));
            const float a = ((B.hasb >> gi) & 1u) ? (part ? cim : cre) : 0.f;
            // row_ror:8 inside each 16-lane row: lane l takes lane l ^ 8
            const float ab = sgn * __builtin_bit_cast(
                                       float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, a), 0x128, 0xf, 0xf,
                                                                       false));
#pragma unroll
            for (int t = 0; t < NT; t++)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, B.xv[gi][t].x, acc[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NT; t++)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ab, B.xv[gi][t].y, acc[t], 0, 0, 0);
        }
    };
    // The wave walks its blocks as one stream of batches, double-buffered:
    // the loads of the next batch are issued before the products of the
    // current one, the two buffers in turn over pairs of batches (no copies:
    // a copy would wait on the loads; no exit between a batch's loads and its
    // products: the compiler would sink the loads past the other buffer's
    // products).  A segment with an odd batch count ends with a padding
    // batch (the last one's loads again, no coefficients); the spare load
    // slot of a segment's last pair takes the next segment's first batch --
    // for a block's last segment the first batch of the wave's next block,
    // whose column indices and words are loaded one block ahead (past the
    // wave's last block: a stand-in, in bounds, never used).
    int b = xs + wid;
    if (b < xe) {
        int jn;
        unsigned wn;
        meta(bofs[b], bofs[b + 1], jn, wn);
        Batch A, Bn;
        issue(0, min(64, bofs[b + 1] - bofs[b]), jn, wn, rowptr[b * RB_ROWS], rowptr[min(b * RB_ROWS + RB_ROWS, n)], A);
        for (; b < xe; b += wpx) {
            const int u0 = bofs[b], u1 = bofs[b + 1];
            const int r0 = b * RB_ROWS, nr = min(RB_ROWS, n - r0);
            const int kb = rowptr[r0], ke = rowptr[r0 + nr];
            // W of the block's rows, for the epilogue (rows past n: row n - 1, unused)
            float2 wpre[4][NC];
            if constexpr (!PACKED) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const float2 *wr = W + (size_t)min(r0 + drow0 + v, n - 1) * KC + cb;
                    if constexpr (NC == 2) {
                        const float4 w4 = *reinterpret_cast<const float4 *>(wr);
                        wpre[v][0] = make_float2(w4.x, w4.y);
                        wpre[v][1] = make_float2(w4.z, w4.w);
                    } else {
                        wpre[v][0] = wr[0];
                    }
                }
            }
            int jv = jn;
            unsigned wv = wn;
            const int bn = min(b + wpx, xe - 1);
            const int kbn = rowptr[bn * RB_ROWS], ken = rowptr[min(bn * RB_ROWS + RB_ROWS, n)];
            const int cntn = min(64, bofs[bn + 1] - bofs[bn]);
            meta(bofs[bn], bofs[bn + 1], jn, wn);
#pragma unroll
            for (int t = 0; t < NT; t++) acc[t] = mfma_f4{0.f, 0.f, 0.f, 0.f};
            // segments of up to 64 entries (one in almost every block); A holds
            // the segment's first batch
            for (int seg = u0; seg < u1; seg += 64) {
                const int cnt = min(64, u1 - seg);
                if (seg > u0) {
                    meta(seg, u1, jv, wv);
                    issue(0, cnt, jv, wv, kb, ke, A);
                }
                const int glast = (cnt - 1) / (4 * MB) * (4 * MB);
                const int npair = (glast / (4 * MB) + 2) / 2;
                for (int pi = 0; pi + 1 < npair; pi++) {
                    const int g1 = (2 * pi + 1) * 4 * MB;
                    issue(min(g1, glast), cnt, jv, wv, kb, ke, Bn);
                    __builtin_amdgcn_sched_barrier(0);
                    compute(A);
                    issue(g1 + 4 * MB, cnt, jv, wv, kb, ke, A);
                    __builtin_amdgcn_sched_barrier(0);
                    compute(Bn);
                }
                {
                    // the segment's last pair
                    const int g1 = (2 * npair - 1) * 4 * MB;
                    issue(min(g1, glast), cnt, jv, wv, kb, ke, Bn);
                    if (g1 > glast) Bn.hasb = 0;
                    __builtin_amdgcn_sched_barrier(0);
                    compute(A);
                    if (seg + 64 >= u1) issue(0, cntn, jn, wn, kbn, ken, A);   // the next block's first batch
                    else issue(glast, cnt, jv, wv, kb, ke, A);                 // stand-in: reloaded above
                    __builtin_amdgcn_sched_barrier(0);
                    compute(Bn);
                }
            }
            // lane l holds part dpart of rows drow0 .. drow0 + 3, tile column m;
            // with its partner l ^ 32 (the other part of the same rows) it trades
            // halves: columns cb .. cb + NC - 1 whole
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = drow0 + v;
                float2 y[NC];
#pragma unroll
                for (int e2 = 0; e2 < NC; e2++) {
                    const float mine = NT > 1 ? acc[dpart ? NC + e2 : e2][v] : acc[0][v];
                    const float send = NT > 1 ? acc[dpart ? e2 : NC + e2][v] : acc[0][v];
                    const float recv = __shfl_xor(send, 32);
                    y[e2] = dpart ? make_float2(recv, mine) : make_float2(mine, recv);
                }
                if (row < nr) {
                    const size_t at = (size_t)(r0 + row) * KC + cb;
                    float2 w[NC];
                    if constexpr (NC == 2) {
                        if (act[0] && act[1]) {
                            *reinterpret_cast<float4 *>(Y + at) = make_float4(y[0].x, y[0].y, y[1].x, y[1].y);
                        } else {
                            if (act[0]) Y[at] = y[0];
                            if (act[1]) Y[at + 1] = y[1];
                        }
                    } else {
                        if (act[0]) Y[at] = y[0];
                    }
#pragma unroll
                    for (int e2 = 0; e2 < NC; e2++)
                        w[e2] = PACKED ? sh_unpack(Wq, (size_t)(r0 + row), cb + e2) : wpre[v][e2];
#pragma unroll
                    for (int e2 = 0; e2 < NC; e2++)
                        if (act[e2]) dot_epi(w[e2], y[e2], want_yy, d1[e2], d2[e2]);
                }
            }
        }
    }
    // per column: the lanes of rows 0-3 and 4-7 (l, l ^ 16), then the waves
    const bool want2 = want_yy != 0;
#pragma unroll
    for (int e = 0; e < NC; e++) {
        d1[e] += __shfl_xor(d1[e], 16);
        if (want2) d2[e] += __shfl_xor(d2[e], 16);
    }
    __shared__ double s1[SPMM64_WAVES][KC], s2[SPMM64_WAVES][KC];
    if ((q & 1) == 0 && owner) {
#pragma unroll
        for (int e = 0; e < NC; e++) {
            s1[wave][cb + e] = d1[e];
            s2[wave][cb + e] = d2[e];
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < KC) {
        double t1 = 0.0, t2 = 0.0;
        for (int w = 0; w < nw; w++) { t1 += s1[w][threadIdx.x]; t2 += s2[w][threadIdx.x]; }
        partial[((size_t)0 * KC + threadIdx.x) * G + blockIdx.x] = t1;
        if (want2) partial[((size_t)1 * KC + threadIdx.x) * G + blockIdx.x] = t2;
    }
}

// ------------------------------- LDS-staged block SpMM, 64 columns
// One workgroup per block of consecutive rows (blocks of one XCD's region
// dealt round-robin to that XCD's workgroups).  The workgroup first copies
// the block's neighbour rows of X -- each distinct row once -- and the
// block's rows of W from global memory into LDS, all loads in flight
// together, then every wave computes a contiguous share of the block's rows
// from LDS: the vector-memory path carries each staged row once per block
// instead of once per nonzero, and no row waits on a global load.
template <int NW>
__device__ inline void block_reduce2_nw(double d1, double d2, double *partial, int blk, int G, bool want2)
{
    __shared__ double s1[NW][64], s2[NW][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    s1[wave][lane] = d1;
    s2[wave][lane] = d2;
    __syncthreads();
    if ((int)threadIdx.x < 64) {
        double t1 = 0.0, t2 = 0.0;
        for (int w = 0; w < NW; w++) { t1 += s1[w][threadIdx.x]; t2 += s2[w][threadIdx.x]; }
        partial[((size_t)0 * 64 + threadIdx.x) * G + blk] = t1;
        if (want2) partial[((size_t)1 * 64 + threadIdx.x) * G + blk] = t2;
    }
}

template <class V, int MODE, int UROWS, int NW, int NZMAX>
__global__ void __launch_bounds__(64 * NW) k_spmm_lds(int nb, const int *__restrict__ brow, const int *__restrict__ bu,
                                                       const int *__restrict__ ucols, const int *__restrict__ rowptr,
                                                       const unsigned short *__restrict__ lcol,
                                                       const V *__restrict__ val, const V *__restrict__ X,
                                                       V *__restrict__ Y, const V *__restrict__ W, int want_yy,
                                                       const int *__restrict__ active, double *__restrict__ partial)
{
    constexpr int NT = 64 * NW;
    constexpr int SB = UROWS / NW;   // staged rows per wave, all in flight
    constexpr int LB = 8;            // nonzeros per batch of LDS reads
    constexpr int RPMAX = UROWS;     // row pointers of a block (rows < UROWS)
    static_assert(SB * NW == UROWS && SB <= 64, "LDS rows must split evenly over the waves");
    __shared__ V xs[UROWS * 64];                 // X neighbour rows, then W rows of the block
    __shared__ V vs[NZMAX];                      // matrix values of the block
    __shared__ unsigned short ls[NZMAX];         // their positions in xs
    __shared__ int rs[RPMAX + 1];                // block-local row pointers
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool act = active[lane] != 0;
    const int G = gridDim.x;
    const int xcd = blockIdx.x % NXCD, gi = blockIdx.x / NXCD, gstep = G / NXCD;
    const int per = (nb + NXCD - 1) / NXCD;
    const int b0 = min(nb, xcd * per), b1 = min(nb, b0 + per);
    double d1 = 0.0, d2 = 0.0;
    for (int b = b0 + gi; b < b1; b += gstep) {
        const int u0 = bu[b], nu = bu[b + 1] - u0;
        const int r0 = brow[b], nra = brow[b + 1] - r0;
        const int z0 = rowptr[r0], nz = rowptr[r0 + nra] - z0;
        const int ns = nu + nra;   // staged rows: X neighbours, then W of the block's rows
        __syncthreads();           // the previous block's readers are done with the buffers
        {
            // row pointers, values and positions of the block (one pass of
            // the workgroup each: nra < RPMAX <= NT, nz <= NZMAX)
            int rp = 0;
            if (tid <= nra) rp = rowptr[r0 + tid] - z0;
            V vv[(NZMAX + NT - 1) / NT];
            unsigned short lv[(NZMAX + NT - 1) / NT];
#pragma unroll
            for (int i = 0; i < (NZMAX + NT - 1) / NT; i++) {
                const int k = min(tid + i * NT, nz - 1);
                vv[i] = val[z0 + k];
                lv[i] = lcol[z0 + k];
            }
            // wave w stages entries w, w + NW, ... w + NW (SB - 1) < UROWS;
            // entries past ns repeat the last one into LDS rows never read,
            // so that every load is unconditional and all are in flight together
            const int e = min(wave + NW * lane, ns - 1);
            const int src = e < nu ? ucols[u0 + e] : -1 - (r0 + e - nu);
            V tmp[SB];
#pragma unroll
            for (int i = 0; i < SB; i++) {
                const int sj = __builtin_amdgcn_readlane(src, i);
                const V *p = sj >= 0 ? X + (size_t)sj * 64 : W + (size_t)(-1 - sj) * 64;
                tmp[i] = p[lane];
            }
            if (tid <= nra) rs[tid] = rp;
#pragma unroll
            for (int i = 0; i < (NZMAX + NT - 1) / NT; i++) {
                const int k = tid + i * NT;
                if (k < nz) { vs[k] = vv[i]; ls[k] = lv[i]; }
            }
#pragma unroll
            for (int i = 0; i < SB; i++) xs[(wave + NW * i) * 64 + lane] = tmp[i];
        }
        __syncthreads();
        const int pw = (nra + NW - 1) / NW;
        const int la = min(nra, wave * pw), lz = min(nra, (wave + 1) * pw);   // block-local rows
        // one row at a time; a nonzero's value and position are broadcast
        // LDS reads (every lane the same address), its X row a 512-byte read
        for (int lr = la; lr < lz; lr++) {
            const int kb = __builtin_amdgcn_readfirstlane(rs[lr]);
            const int ke = __builtin_amdgcn_readfirstlane(rs[lr + 1]);
            V acc = vzero<V>();
            for (int k0 = kb; k0 < ke; k0 += LB) {
                V xb[LB], vb[LB];
#pragma unroll
                for (int u = 0; u < LB; u++) {
                    const int kk = min(k0 + u, ke - 1);
                    vb[u] = vs[kk];
                    xb[u] = xs[(int)ls[kk] * 64 + lane];
                }
                if (k0 + LB <= ke) {
#pragma unroll
                    for (int u = 0; u < LB; u++) acc = vfma(acc, vb[u], xb[u]);
                } else {
#pragma unroll
                    for (int u = 0; u < LB; u++)
                        if (k0 + u < ke) acc = vfma(acc, vb[u], xb[u]);
                }
            }
            const V wv = xs[(nu + lr) * 64 + lane];
            if (act) {
                const size_t at = (size_t)(r0 + lr) * 64 + lane;
                if (MODE == 0) {
                    Y[at] = acc;
                    dot_epi(wv, acc, want_yy, d1, d2);
                } else {
                    const V r = vsub(wv, acc);
                    Y[at] = r;
                    d1 += dotR(r, r);
                }
            }
        }
    }
    block_reduce2_nw<NW>(d1, d2, partial, blockIdx.x, G, MODE == 0 && want_yy != 0);
}

// ------------------- LDS-staged block SpMM with LDS DMA, 64 columns
// One workgroup per CU walks the row blocks of its XCD's region with two LDS
// buffers and two kinds of waves.  DMA_PW producer waves copy block b + step
// into one buffer with the global->LDS DMA (global_load_lds_dwordx4: 16
// bytes per lane straight into LDS, no registers) -- the block's distinct
// neighbour rows of X and its own rows of W (512 B each, two rows per
// instruction), its matrix values, their positions in the staged list and
// its row pointers -- while DMA_CW compute waves work on block b from the
// other buffer, touching LDS only (plus the Y stores): per row run a wave
// reads 64 values and positions with one coalesced LDS read each,
// broadcasts them with readlane and gathers the X rows from LDS.  Both
// groups meet at one barrier per block.  Each nonzero accumulates as
// acc + val * x in ascending order, the same operation sequence as the
// direct kernel.
constexpr int DMA_CW = 8, DMA_PW = 2, DMA_NT = 64 * (DMA_CW + DMA_PW);
struct DmaBuf {
    float2 x[LDS_ROWS_C32 * 64];          // X neighbour rows, then W rows of the block
    float2 v[LDS_NZ_C32 + 2];             // values, staged from an even index
    unsigned short l[LDS_NZ_C32 + 8];     // staged-row positions, staged from a multiple of 8
    int rp[LDS_ROWS_C32 + 4];             // row pointers, staged from a multiple of 4
};
static_assert(sizeof(DmaBuf) % 16 == 0, "LDS buffer parts must stay 16-byte aligned");

typedef __attribute__((address_space(3))) void *lds_ptr;
// Lane i copies 16 bytes from g into LDS at base + 16 i (base wave-uniform).
// Only producer waves issue it; each waits for all of its copies (vmcnt(0))
// before the barrier that hands the buffer to the compute waves.
__device__ inline void dma16(const void *g, void *base)
{
    __builtin_amdgcn_global_load_lds(g, (lds_ptr)base, 16, 0, 0);
}

struct DmaBlock {
    int u0, nu, r0, nra, z0, nz;
};
// block bounds; the nonzero range (z0, nz) only when `nonzeros` (the
// producers read it from the row pointers in global memory, the compute
// waves from the staged copy, so that their loop issues no vector load:
// a vector load's wait would also wait for their pending Y stores)
__device__ inline DmaBlock dma_block(int b, const int *__restrict__ brow, const int *__restrict__ bu,
                                     const int *__restrict__ rowptr, bool nonzeros)
{
    DmaBlock d;
    d.u0 = bu[b];
    d.nu = bu[b + 1] - d.u0;
    d.r0 = brow[b];
    d.nra = brow[b + 1] - d.r0;
    d.z0 = nonzeros ? rowptr[d.r0] : 0;
    d.nz = nonzeros ? rowptr[d.r0 + d.nra] - d.z0 : 0;
    return d;
}

// producer wave pw issues its share of the DMA of block d into B and waits
// for it
__device__ inline void dma_stage(const DmaBlock &d, DmaBuf &B, const int *__restrict__ ucols,
                                 const int *__restrict__ rowptr, const unsigned short *__restrict__ lcol,
                                 const float2 *__restrict__ val, const float2 *__restrict__ X,
                                 const float2 *__restrict__ W, int pw, int lane)
{
    const int ns = d.nu + d.nra, npair = (ns + 1) / 2;
    const int half = lane >> 5, q16 = (lane & 31) * 2;   // float2 offset of the lane's 16 bytes
    // pairs p = pw + PW i; per chunk of 32 of them, lane q looks up the
    // source of row 2 p + q % 2 for i = i0 + q / 2: >= 0 an X row, < 0 the
    // W row -1 - src
    for (int i0 = 0; pw + DMA_PW * i0 < npair; i0 += 32) {
        int src = 0;
        {
            const int r = 2 * (pw + DMA_PW * (i0 + (lane >> 1))) + (lane & 1);
            if (r < ns) src = r < d.nu ? ucols[d.u0 + r] : -1 - (d.r0 + r - d.nu);
        }
        for (int i = 0; i < 32; i++) {
            const int p = pw + DMA_PW * (i0 + i);
            if (p >= npair) break;
            const int sa = __builtin_amdgcn_readlane(src, 2 * i), sb = __builtin_amdgcn_readlane(src, 2 * i + 1);
            const int sj = half ? sb : sa;
            const float2 *g = (sj >= 0 ? X + (size_t)sj * 64 : W + (size_t)(-1 - sj) * 64) + q16;
            if (2 * p + half < ns) dma16(g, &B.x[2 * p * 64]);
        }
    }
    // metadata, 8 instructions at most, dealt round-robin over the producers:
    // m = 0..4 values (2 per 16-byte chunk from index z0 & ~1), m = 5, 6
    // positions (8 per chunk from z0 & ~7), m = 7 row pointers (4 per chunk
    // from r0 & ~3)
    const int nvc = ((d.z0 & 1) + d.nz + 1) / 2, nlc = ((d.z0 & 7) + d.nz + 7) / 8;
    const int nrc = ((d.r0 & 3) + d.nra + 1 + 3) / 4;
    for (int m = pw; m < 8; m += DMA_PW) {
        if (m < 5) {
            const int q = 64 * m + lane;
            if (64 * m < nvc && q < nvc) dma16(val + (d.z0 & ~1) + 2 * q, &B.v[128 * m]);
        } else if (m < 7) {
            const int q = 64 * (m - 5) + lane;
            if (64 * (m - 5) < nlc && q < nlc) dma16(lcol + (d.z0 & ~7) + 8 * q, &B.l[512 * (m - 5)]);
        } else {
            if (lane < nrc) dma16(rowptr + (d.r0 & ~3) + 4 * lane, &B.rp[0]);
        }
    }
    __builtin_amdgcn_s_waitcnt(0x0f70);   // vmcnt(0): this wave's copies have landed
}

// compute wave cw: its contiguous share of the block's rows from B
template <bool WANT_YY>
__device__ inline void dma_compute(DmaBlock d, const DmaBuf &B, float2 *__restrict__ Y, bool act, int cw,
                                   int lane, double &d1, double &d2)
{
    constexpr int BB = 16;
    d.z0 = B.rp[d.r0 & 3];
    const int pw = (d.nra + DMA_CW - 1) / DMA_CW;
    const int la = min(d.nra, cw * pw), nr = min(d.nra, la + pw) - la;
    if (nr <= 0) return;
    const int roff = (d.r0 & 3) + la, voff = d.z0 & 1, loff = d.z0 & 7;
    const int rpv = B.rp[roff + min(lane, nr)] - d.z0;   // block-local row pointers of the run
    const int kb = __builtin_amdgcn_readlane(rpv, 0), ke = __builtin_amdgcn_readlane(rpv, nr);
    int row = 0;
    int rend = __builtin_amdgcn_readlane(rpv, 1);
    float2 acc = make_float2(0.f, 0.f);
    auto flush = [&]() {
        const float2 wv = B.x[(d.nu + la + row) * 64 + lane];
        if (act) {
            Y[(size_t)(d.r0 + la + row) * 64 + lane] = acc;
            d1 += dotR(wv, acc);
            if (WANT_YY) d2 += dotR(acc, acc);
        }
        acc = make_float2(0.f, 0.f);
        row++;
        rend = __builtin_amdgcn_readlane(rpv, min(row + 1, nr));
    };
    for (int seg = kb; seg < ke; seg += 64) {
        const int kk = min(seg + lane, ke - 1);
        const int lc = B.l[loff + kk];
        const float2 vv = B.v[voff + kk];
        const int m = min(64, ke - seg);
        for (int t0 = 0; t0 < m; t0 += BB) {
            float2 xb[BB];
#pragma unroll
            for (int u = 0; u < BB; u++) {
                const int j = __builtin_amdgcn_readlane(lc, min(t0 + u, m - 1));
                xb[u] = B.x[j * 64 + lane];
            }
#pragma unroll
            for (int u = 0; u < BB; u++) {
                if (t0 + u < m) {
                    const int k = seg + t0 + u;
                    while (k >= rend && row < nr) flush();
                    acc = vadd(acc, vmul(rdlane(vv, t0 + u), xb[u]));
                }
            }
        }
    }
    while (row < nr) flush();
}

template <bool WANT_YY>
__global__ void __launch_bounds__(DMA_NT) k_spmm_dma(int nb, const int *__restrict__ brow,
                                                     const int *__restrict__ bu, const int *__restrict__ ucols,
                                                     const int *__restrict__ rowptr,
                                                     const unsigned short *__restrict__ lcol,
                                                     const float2 *__restrict__ val, const float2 *__restrict__ X,
                                                     float2 *__restrict__ Y, const float2 *__restrict__ W,
                                                     const int *__restrict__ active, double *__restrict__ partial,
                                                     int nblk)
{
    __shared__ DmaBuf buf[2];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool producer = wave >= DMA_CW;
    const int pw = wave - DMA_CW;
    const bool act = active[lane] != 0;
    const int G = gridDim.x;
    const int xcd = blockIdx.x % NXCD, gi = blockIdx.x / NXCD, gstep = G / NXCD;
    const int per = (nb + NXCD - 1) / NXCD;
    const int bx0 = min(nb, xcd * per), b1 = min(nb, bx0 + per);
    double d1 = 0.0, d2 = 0.0;
    int b = bx0 + gi, s = 0;
    if (producer && b < b1) dma_stage(dma_block(b, brow, bu, rowptr, true), buf[0], ucols, rowptr, lcol, val, X, W, pw, lane);
    for (; b < b1; b += gstep, s ^= 1) {
        __syncthreads();   // block b is in buf[s]; buf[s ^ 1] is free
        if (producer) {
            if (b + gstep < b1)
                dma_stage(dma_block(b + gstep, brow, bu, rowptr, true), buf[s ^ 1], ucols, rowptr, lcol, val, X, W, pw,
                          lane);
        } else {
            dma_compute<WANT_YY>(dma_block(b, brow, bu, rowptr, false), buf[s], Y, act, wave, lane, d1, d2);
        }
    }
    // fixed-order reduction over the compute waves (producers hold zeros)
    __syncthreads();
    double *red = (double *)&buf[0];   // the buffers are free now
    if (!producer) {
        red[wave * 64 + lane] = d1;
        red[(DMA_CW + wave) * 64 + lane] = d2;
    }
    __syncthreads();
    if (tid < 64) {
        double t1 = 0.0, t2 = 0.0;
        for (int w = 0; w < DMA_CW; w++) { t1 += red[w * 64 + tid]; t2 += red[(DMA_CW + w) * 64 + tid]; }
        partial[((size_t)0 * 64 + tid) * nblk + blockIdx.x] = t1;
        if (WANT_YY) partial[((size_t)1 * 64 + tid) * nblk + blockIdx.x] = t2;
        // the finalize sums nblk partials per column: zero the ones no workgroup owns
        for (int e = blockIdx.x + G; e < nblk; e += G) {
            partial[((size_t)0 * 64 + tid) * nblk + e] = 0.0;
            if (WANT_YY) partial[((size_t)1 * 64 + tid) * nblk + e] = 0.0;
        }
    }
}

// single-column SpMV: LPR lanes share one row
template <class V, int LPR>
__global__ void __launch_bounds__(256) k_spmv_dot(int n, const int *__restrict__ rowptr,
                                                  const int *__restrict__ colidx, const V *__restrict__ val,
                                                  const V *__restrict__ idiag, const V *__restrict__ X,
                                                  V *__restrict__ Y, const V *__restrict__ W, int want_yy,
                                                  const int *__restrict__ active, double *__restrict__ partial)
{
    constexpr int RPW = 64 / LPR;   // rows per wave
    const int G = gridDim.x;
    const int blk = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int sub = lane % LPR, rw = lane / LPR;
    const bool act = active[0] != 0;
    const RowSched rs = row_sched(n, RPW);
    double d1 = 0.0, d2 = 0.0;
    for (int base = rs.first; base < rs.end; base += rs.stride) {
        const int i = base + rw;
        V acc = vzero<V>();
        if (i < rs.end && act) {
            const int kb = rowptr[i], ke = rowptr[i + 1];
            for (int k = kb + sub; k < ke; k += LPR) {
                const int j = colidx[k];
                V xv = X[j];
                if (idiag) xv = vmul(idiag[j], xv);
                acc = vmac(acc, val[k], xv);
            }
        }
        // sum the LPR partial rows (all lanes take part in the shuffles)
        for (int off = LPR / 2; off > 0; off >>= 1) acc = shfl_add(acc, off);
        if (i < rs.end && act && sub == 0) {
            Y[i] = acc;
            dot_epi(W[i], acc, want_yy, d1, d2);
        }
    }
    block_reduce2<1>(d1, d2, partial, blk, G, want_yy != 0);
}

// --------------------------------------------------------- FP64 residual
template <class V, int K, bool MF = false>
__global__ void __launch_bounds__(256) k_resid(int n, const int *__restrict__ rowptr, const int *__restrict__ colidx,
                                               const V *__restrict__ val, const V *__restrict__ X,
                                               const V *__restrict__ B, V *__restrict__ Rr,
                                               const int *__restrict__ active, double *__restrict__ partial,
                                               const double *__restrict__ omg = nullptr)
{
    constexpr int R = 64 / K;
    const int G = gridDim.x;
    const int blk = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int c = lane % K, rsub = lane / K;
    const bool act = active[c] != 0;
    const RowSched rs = row_sched(n, R);
    double d1 = 0.0;
    for (int base = rs.first; base < rs.end; base += rs.stride) {
        const int i = base + rsub;
        if (i < rs.end && act) {
            V acc = row_gather<V, K, false, MF>(rowptr[i], rowptr[i + 1], c, colidx, val, nullptr, X,
                                                mf_omega<MF, V>(omg, c));
            V r = vsub(B[(size_t)i * K + c], acc);
            Rr[(size_t)i * K + c] = r;
            d1 += dotR(r, r);
        }
    }
    block_reduce2<K>(d1, 0.0, partial, blk, G, false);
}

// -------------------------------------------------- element-wise updates
// Block vectors are walked as flat arrays of n*K elements; element t belongs
// to column t % K.  Kernels without a reduction give every thread one
// element (grid of n*K/256 workgroups): each thread issues its data loads
// first (index clamped into the array), the first wave meanwhile copies the
// per-column flags and scalars into LDS, and the element is finished after
// one barrier.  Measured on MI355X for 3 reads + 1 write over 128 MB arrays:
// 6.1 TB/s this way, 4.8 TB/s when every wave loads its own column
// parameters (a dependent vector load per wave).  Kernels with a per-column
// reduction keep a fixed grid of G workgroups (one partial per workgroup)
// and sweep the arrays grid-stride, EW_U elements per thread per sweep: the
// loads of all EW_U elements are issued before the first use, and the
// sweep's stride is a multiple of 256, hence of K, so a thread's column is
// fixed.
#define EW1_PROLOGUE                                                     \
    const size_t tot = (size_t)n * K;                                    \
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;             \
    const size_t tl = min(t, tot - 1);                                   \
    const int c = (int)(threadIdx.x % K);

// per-column parameters of a one-element kernel, staged in LDS
template <int K, int NI, int ND>
struct ColPar {
    int i[NI][K];
    double d[ND][K];
};

#define EW_PROLOGUE                                                      \
    const int G = gridDim.x;                                             \
    const int blk = blockIdx.x;                                          \
    const size_t tot = (size_t)n * K;                                    \
    const int c = (int)(threadIdx.x % K);

constexpr int EW_U = 4;
// ld(t) returns the loaded operands of element t as one value, use(t, v)
// consumes them; per-thread sums accumulate in ascending t.
template <class LD, class USE>
__device__ __forceinline__ void ew_for(size_t tot, LD ld, USE use)
{
    using L = decltype(ld(0));
    const size_t stride = (size_t)gridDim.x * 256 * EW_U;
    for (size_t base = (size_t)blockIdx.x * 256 * EW_U + threadIdx.x; base < tot; base += stride) {
        L v[EW_U];
#pragma unroll
        for (int u = 0; u < EW_U; u++)
            if (base + u * 256 < tot) v[u] = ld(base + u * 256);
#pragma unroll
        for (int u = 0; u < EW_U; u++)
            if (base + u * 256 < tot) use(base + u * 256, v[u]);
    }
}

// Nontemporal loads and stores of the streamed block vectors.  Measured on
// MI355X for 3 reads + 1 write over 128 MB arrays (one element per thread,
// working set beyond the last-level cache): 6.8 TB/s with both loads and
// stores nontemporal, 6.1 TB/s with neither or only one of them.
#ifndef TMI_EW_NT
#define TMI_EW_NT 1
#endif
typedef double tmi_dv2 __attribute__((ext_vector_type(2)));
template <class V>
__device__ inline V ldnt(const V *p)
{
    if constexpr (!TMI_EW_NT) return *p;
    else if constexpr (sizeof(V) == 8 && alignof(V) == 8 && !std::is_same<V, double>::value)
        return __builtin_bit_cast(V, __builtin_nontemporal_load((const unsigned long long *)p));
    else if constexpr (sizeof(V) == 16) {
        const tmi_dv2 u = __builtin_nontemporal_load((const tmi_dv2 *)p);
        return __builtin_bit_cast(V, u);
    } else
        return __builtin_nontemporal_load(p);
}
template <class V>
__device__ inline void stnt(V *p, V v)
{
    if constexpr (!TMI_EW_NT) *p = v;
    else if constexpr (sizeof(V) == 8 && alignof(V) == 8 && !std::is_same<V, double>::value)
        __builtin_nontemporal_store(__builtin_bit_cast(unsigned long long, v), (unsigned long long *)p);
    else if constexpr (sizeof(V) == 16)
        __builtin_nontemporal_store(__builtin_bit_cast(tmi_dv2, v), (tmi_dv2 *)p);
    else
        __builtin_nontemporal_store(v, p);
}

template <class V, int K>
__global__ void __launch_bounds__(256) k_bicg_p(int n, const V *__restrict__ r, V *__restrict__ p,
                                                const V *__restrict__ v, const double *__restrict__ sc,
                                                const int *__restrict__ ic)
{
    EW1_PROLOGUE
    __shared__ ColPar<K, 2, 2> cp;
    const V rv = ldnt(r + tl), pv = ldnt(p + tl), vv = ldnt(v + tl);
    if (threadIdx.x < K) {
        cp.i[0][threadIdx.x] = ic[I_ACTIVE * K + threadIdx.x];
        cp.i[1][threadIdx.x] = ic[I_FIRST * K + threadIdx.x];
        cp.d[0][threadIdx.x] = sc[S_BETA * K + threadIdx.x];
        cp.d[1][threadIdx.x] = sc[S_OMEGA * K + threadIdx.x];
    }
    __syncthreads();
    if (t >= tot || !cp.i[0][c]) return;
    stnt(p + t, cp.i[1][c] ? rv : vadd(rv, vscale(vsub(pv, vscale(vv, cp.d[1][c])), cp.d[0][c])));
}

template <class V, int K>
__global__ void __launch_bounds__(256) k_bicg_s(int n, const V *__restrict__ r, const V *__restrict__ v,
                                                V *__restrict__ s, const double *__restrict__ sc,
                                                const int *__restrict__ ic)
{
    EW1_PROLOGUE
    __shared__ ColPar<K, 1, 1> cp;
    const V rv = ldnt(r + tl), vv = ldnt(v + tl);
    if (threadIdx.x < K) {
        cp.i[0][threadIdx.x] = ic[I_ACTIVE * K + threadIdx.x];
        cp.d[0][threadIdx.x] = sc[S_ALPHA * K + threadIdx.x];
    }
    __syncthreads();
    if (t >= tot || !cp.i[0][c]) return;
    stnt(s + t, vsub(rv, vscale(vv, cp.d[0][c])));
}

// x += D^-1 (alpha p + omega s); r = s - omega t; partials ||r||^2, <rd, r>
template <class V, int K>
__global__ void __launch_bounds__(256) k_bicg_xr(int n, const V *__restrict__ idiag, V *__restrict__ x,
                                                 const V *__restrict__ p, const V *__restrict__ s,
                                                 const V *__restrict__ tv, V *__restrict__ r,
                                                 const V *__restrict__ rd, const double *__restrict__ sc,
                                                 const int *__restrict__ ic, double *__restrict__ partial,
                                                 const unsigned char *__restrict__ rdq)
{
    EW_PROLOGUE
    double d1 = 0.0, d2 = 0.0;
    if (ic[I_ACTIVE * K + c]) {
        const double alpha = sc[S_ALPHA * K + c], omega = sc[S_OMEGA * K + c];
        struct L { V s, p, x, t, rd, di; };
        ew_for(
            tot,
            [&](size_t t) {
                V rdv;
                if constexpr (std::is_same<V, float2>::value && K == 64)
                    rdv = rdq ? sh_unpack(rdq, t / 64, (int)(t % 64)) : ldnt(rd + t);   // packed shadow residual
                else
                    rdv = ldnt(rd + t);
                return L{ldnt(s + t), ldnt(p + t), ldnt(x + t), ldnt(tv + t), rdv, idiag ? idiag[t / K] : V{}};
            },
            [&](size_t t, const L &a) {
                V upd = vadd(vscale(a.p, alpha), vscale(a.s, omega));
                if (idiag) upd = vmul(a.di, upd);
                stnt(x + t, vadd(a.x, upd));
                V rn = vsub(a.s, vscale(a.t, omega));
                stnt(r + t, rn);
                d1 += dotR(rn, rn);
                d2 += dotR(a.rd, rn);
            });
    }
    block_reduce2<K>(d1, d2, partial, blk, G, true);
}

// PCG: x += alpha d; r -= alpha q; partials <r, D^-1 r> and ||r||^2
template <class V, int K>
__global__ void __launch_bounds__(256) k_cg_xr(int n, const V *__restrict__ idiag, V *__restrict__ x,
                                               const V *__restrict__ d, const V *__restrict__ q, V *__restrict__ r,
                                               const double *__restrict__ sc, const int *__restrict__ ic,
                                               double *__restrict__ partial)
{
    EW_PROLOGUE
    double d1 = 0.0, d2 = 0.0;
    if (ic[I_ACTIVE * K + c]) {
        const double alpha = sc[S_ALPHA * K + c];
        struct L { V x, d, r, q, di; };
        ew_for(
            tot, [&](size_t t) { return L{x[t], d[t], r[t], q[t], idiag ? idiag[t / K] : V{}}; },
            [&](size_t t, const L &a) {
                x[t] = vadd(a.x, vscale(a.d, alpha));
                V rn = vsub(a.r, vscale(a.q, alpha));
                r[t] = rn;
                V z = idiag ? vmul(a.di, rn) : rn;
                d1 += dotR(rn, z);
                d2 += dotR(rn, rn);
            });
    }
    block_reduce2<K>(d1, d2, partial, blk, G, true);
}

// PCG: d = beta d + D^-1 r (beta = 0 on the first call); optional partials
// <r, D^-1 r>, ||r||^2 (used to start the iteration)
template <class V, int K>
__global__ void __launch_bounds__(256) k_cg_d(int n, const V *__restrict__ idiag, V *__restrict__ d,
                                              const V *__restrict__ r, const double *__restrict__ sc,
                                              const int *__restrict__ ic, double *__restrict__ partial)
{
    EW_PROLOGUE
    double d1 = 0.0, d2 = 0.0;
    if (ic[I_ACTIVE * K + c]) {
        const bool first = ic[I_FIRST * K + c] != 0;
        const double beta = sc[S_BETA * K + c];
        struct L { V r, d, di; };
        ew_for(
            tot,
            [&](size_t t) { return L{r[t], first ? V{} : d[t], idiag ? idiag[t / K] : V{}}; },
            [&](size_t t, const L &a) {
                V z = idiag ? vmul(a.di, a.r) : a.r;
                d[t] = first ? z : vadd(vscale(a.d, beta), z);
                if (partial) { d1 += dotR(a.r, z); d2 += dotR(a.r, a.r); }
            });
    }
    if (partial) block_reduce2<K>(d1, d2, partial, blk, G, true);
}

// ---- COCG: PCG for complex symmetric systems with the unconjugated
// bilinear form (van der Vorst & Melissen, IEEE Trans. Magn. 26 (1990)
// 706-708): alpha = r^T z / p^T A p, beta = r_new^T z_new / r^T z, z = D^-1 r,
// all complex; one SpMM per iteration.
// complex value times a complex scalar (ar + i ai)
__device__ inline float2 cscale(float2 v, double ar, double ai)
{
    const float a = (float)ar, b = (float)ai;
    return make_float2(v.x * a - v.y * b, v.x * b + v.y * a);
}
__device__ inline double2 cscale(double2 v, double ar, double ai)
{
    return make_double2(v.x * ar - v.y * ai, v.x * ai + v.y * ar);
}
// unconjugated product terms: re and im of a b
template <class V>
__device__ inline void bil_add(V a, V b, double &re, double &im)
{
    if constexpr (std::is_same<V, float2>::value || std::is_same<V, double2>::value) {
        re += (double)a.x * (double)b.x - (double)a.y * (double)b.y;
        im += (double)a.x * (double)b.y + (double)a.y * (double)b.x;
    } else {
        re += (double)a * (double)b;
    }
}

// z = D^-1 r for column c.  Without omg, di is the row's inverse diagonal;
// with omg (multi-frequency solve), di is the row's diagonal entry with its
// imaginary part at omega = 1 and the column's diagonal is re + i omg[c] im.
template <class V>
__device__ inline V prec_apply(V di, V r, const double *__restrict__ omg, int c)
{
    if constexpr (std::is_same<V, float2>::value || std::is_same<V, double2>::value) {
        if (omg) {
            using S = typename ScalarOf<V>::type;
            const S a = di.x, b = di.y * (S)omg[c], den = a * a + b * b;
            V z;
            z.x = (r.x * a + r.y * b) / den;
            z.y = (r.y * a - r.x * b) / den;
            return z;
        }
    }
    return vmul(di, r);
}

// x += alpha d; r -= alpha q; partials r^T D^-1 r (re, im) and ||r||^2
// XD: the x update x += alpha d is left to the direction update (k_cocg_d<.., true>),
// which reads d anyway; this kernel then streams only r and q
template <class V, int K, bool XD = false>
__global__ void __launch_bounds__(256) k_cocg_xr(int n, const V *__restrict__ idiag, V *__restrict__ x,
                                                 const V *__restrict__ d, const V *__restrict__ q,
                                                 V *__restrict__ r, const double *__restrict__ sc,
                                                 const int *__restrict__ ic, double *__restrict__ partial,
                                                 const double *__restrict__ omg)
{
    EW_PROLOGUE
    double d1 = 0.0, d2 = 0.0, d3 = 0.0;
    if (ic[I_ACTIVE * K + c]) {
        const double ar = sc[S_ALPHA * K + c], ai = sc[S_ALPHA_I * K + c];
        struct L { V x, d, r, q, di; };
        ew_for(
            tot,
            [&](size_t t) {
                return L{XD ? V{} : ldnt(x + t), XD ? V{} : ldnt(d + t), ldnt(r + t), ldnt(q + t),
                         idiag ? idiag[t / K] : V{}};
            },
            [&](size_t t, const L &a) {
                if (!XD) stnt(x + t, vadd(a.x, cscale(a.d, ar, ai)));
                const V rn = vsub(a.r, cscale(a.q, ar, ai));
                stnt(r + t, rn);
                const V z = idiag ? prec_apply(a.di, rn, omg, c) : rn;
                bil_add(rn, z, d1, d2);
                d3 += dotR(rn, rn);
            });
    }
    block_reduce3<K>(d1, d2, d3, partial, blk, G);
}

// d = beta d + D^-1 r (d = D^-1 r on the first call); with partials (start
// of the iteration): r^T D^-1 r (re, im) and ||r||^2
template <class V, int K, bool XD = false>
__global__ void __launch_bounds__(256) k_cocg_d(int n, const V *__restrict__ idiag, V *__restrict__ d,
                                                const V *__restrict__ r, const double *__restrict__ sc,
                                                const int *__restrict__ ic, double *__restrict__ partial,
                                                const double *__restrict__ omg, V *__restrict__ x)
{
    EW_PROLOGUE
    double d1 = 0.0, d2 = 0.0, d3 = 0.0;
    const bool act = ic[I_ACTIVE * K + c] != 0;
    // XD: the x update of the iteration, for the columns whose alpha step ran
    // (I_XPEND; a column the beta step just stopped still takes it)
    const bool xp = XD && ic[I_XPEND * K + c] != 0;
    if (act || xp) {
        const bool first = ic[I_FIRST * K + c] != 0;
        const double br = sc[S_BETA * K + c], bi = sc[S_BETA_I * K + c];
        const double ar = XD ? sc[S_ALPHA * K + c] : 0.0, ai = XD ? sc[S_ALPHA_I * K + c] : 0.0;
        struct L { V r, d, di, x; };
        ew_for(
            tot,
            [&](size_t t) {
                return L{act ? ldnt(r + t) : V{}, (first && !xp) ? V{} : ldnt(d + t),
                         (act && idiag) ? idiag[t / K] : V{}, xp ? ldnt(x + t) : V{}};
            },
            [&](size_t t, const L &a) {
                if (xp) stnt(x + t, vadd(a.x, cscale(a.d, ar, ai)));
                if (act) {
                    const V z = idiag ? prec_apply(a.di, a.r, omg, c) : a.r;
                    stnt(d + t, first ? z : vadd(cscale(a.d, br, bi), z));
                    if (partial) {
                        bil_add(a.r, z, d1, d2);
                        d3 += dotR(a.r, a.r);
                    }
                }
            });
    }
    if (partial) block_reduce3<K>(d1, d2, d3, partial, blk, G);
}

// r_in = r64 * scale ; rd_in = r_in (sign_shadow: its sign pattern) ;
// x_in = 0 ; partials <r_in, r_in> and <rd_in, r_in>
template <class V64, class V, int K>
__global__ void __launch_bounds__(256) k_to_inner(int n, const V64 *__restrict__ r64, V *__restrict__ rin,
                                                  V *__restrict__ rdin, V *__restrict__ xin,
                                                  const double *__restrict__ scale, const int *__restrict__ oflags,
                                                  double *__restrict__ partial, int sign_shadow)
{
    EW_PROLOGUE
    double d1 = 0.0, d2 = 0.0;
    if (oflags[c]) {
        const double s = scale[c];
        ew_for(
            tot, [&](size_t t) { return r64[t]; },
            [&](size_t t, const V64 &a) {
                V v = cvt<V>(vscale(a, s));
                rin[t] = v;
                const V rd = sign_shadow ? shadow_sign(v) : v;
                if (rdin) rdin[t] = rd;
                if (xin) xin[t] = vzero<V>();
                d1 += dotR(v, v);
                d2 += dotR(rd, v);   // rho_0 = <rd, r_0>
            });
    }
    block_reduce2<K>(d1, d2, partial, blk, G, true);
}

// reliable update: r_in = r64 * scale, x_in = 0; partials <r_in, r_in> and
// <rd, r_in> (BiCGSTAB shadow residual) or <r_in, D^-1 r_in> (PCG)
template <class V64, class V, int K>
__global__ void __launch_bounds__(256) k_reliable(int n, const V64 *__restrict__ r64, V *__restrict__ rin,
                                                  V *__restrict__ xin, const V *__restrict__ rd,
                                                  const V *__restrict__ idiag, const double *__restrict__ scale,
                                                  const int *__restrict__ oflags, double *__restrict__ partial,
                                                  int bil, const double *__restrict__ omg)
{
    EW_PROLOGUE
    double d1 = 0.0, d2 = 0.0, d3 = 0.0;
    if (oflags[c] && bil) {
        // COCG: r^T D^-1 r (re, im) and ||r||^2
        const double s = scale[c];
        struct L { V64 r; V di; };
        ew_for(
            tot, [&](size_t t) { return L{r64[t], idiag ? idiag[t / K] : V{}}; },
            [&](size_t t, const L &a) {
                V v = cvt<V>(vscale(a.r, s));
                rin[t] = v;
                xin[t] = vzero<V>();
                bil_add(v, idiag ? prec_apply(a.di, v, omg, c) : v, d1, d2);
                d3 += dotR(v, v);
            });
    } else if (oflags[c]) {
        const double s = scale[c];
        struct L { V64 r; V rd, di; };
        ew_for(
            tot,
            [&](size_t t) { return L{r64[t], rd ? rd[t] : V{}, (!rd && idiag) ? idiag[t / K] : V{}}; },
            [&](size_t t, const L &a) {
                V v = cvt<V>(vscale(a.r, s));
                rin[t] = v;
                xin[t] = vzero<V>();
                d1 += dotR(v, v);
                if (rd) d2 += dotR(a.rd, v);
                else d2 += dotR(v, idiag ? vmul(a.di, v) : v);
            });
    }
    if (bil) block_reduce3<K>(d1, d2, d3, partial, blk, G);
    else block_reduce2<K>(d1, d2, partial, blk, G, true);
}

template <class V, int K>
__global__ void __launch_bounds__(256) k_norm2(int n, const V *__restrict__ X, const int *__restrict__ flags,
                                               double *__restrict__ partial)
{
    EW_PROLOGUE
    double d1 = 0.0;
    if (flags[c]) ew_for(tot, [&](size_t t) { return X[t]; }, [&](size_t, const V &a) { d1 += dotR(a, a); });
    block_reduce2<K>(d1, d1, partial, blk, G, true);
}

template <class V64, class V, int K>
__global__ void __launch_bounds__(256) k_from_inner(int n, V64 *__restrict__ x64, const V *__restrict__ xin,
                                                    const double *__restrict__ scale,
                                                    const int *__restrict__ oflags)
{
    EW1_PROLOGUE
    __shared__ ColPar<K, 1, 1> cp;
    const V64 xv = ldnt(x64 + tl);
    const V xi = ldnt(xin + tl);
    if (threadIdx.x < K) {
        cp.i[0][threadIdx.x] = oflags[threadIdx.x];
        cp.d[0][threadIdx.x] = scale[threadIdx.x];
    }
    __syncthreads();
    if (t >= tot || !cp.i[0][c]) return;
    stnt(x64 + t, vadd(xv, vscale(to64(xi), 1.0 / cp.d[0][c])));
}

// (cr + i ci) = (ar + i ai) / (br + i bi); false when the quotient is not finite
__device__ inline bool cdiv(double ar, double ai, double br, double bi, double &cr, double &ci)
{
    const double den = br * br + bi * bi;
    cr = (ar * br + ai * bi) / den;
    ci = (ai * br - ar * bi) / den;
    return den != 0.0 && isfinite(cr) && isfinite(ci);
}

// The scalar recurrence of one column after a reduction (r0, r1, r2: the
// summed partials); S and I hold the column's scalar and flag slots
// (I[I_NSLOT]: the outer solve's active flag).
__device__ inline void fin_column(int fk, double r0, double r1, double r2, double *S, int *I, int maxit, double tol2)
{
#define SC(slot) S[slot]
#define IC(slot) I[slot]
    switch (fk) {
    case FIN_OUTER_BNORM:
        SC(S_BNORM2) = r0;
        SC(S_SCALE) = r0 > 0.0 ? 1.0 / sqrt(r0) : 1.0;
        SC(S_BEST) = -1.0;
        IC(I_STALL) = 0;
        if (r0 == 0.0) IC(I_ACTIVE) = 0;   // zero right-hand side: the field stays at the initial guess
        break;
    case FIN_OUTER_RES: {
        if (!IC(I_ACTIVE)) break;
        SC(S_RES2) = r0;
        IC(I_ITERS) += 1;   // true-residual evaluations
        if (r0 <= tol2 * SC(S_BNORM2) || !isfinite(r0)) { IC(I_ACTIVE) = 0; break; }
        // stagnation: no 2x drop of the true residual norm over 8 evaluations
        if (SC(S_BEST) < 0.0 || r0 < 0.25 * SC(S_BEST)) { SC(S_BEST) = r0; IC(I_STALL) = 0; }
        else if (++IC(I_STALL) >= 8) IC(I_ACTIVE) = 0;
        break;
    }
    case FIN_BICG_RELIABLE:
    case FIN_CG_RELIABLE: {
        const int oact = I[I_NSLOT];
        if (!oact) { IC(I_ACTIVE) = 0; break; }
        SC(S_RES2) = r0;
        if (fk == FIN_BICG_RELIABLE) {
            // recompute beta of the coming iteration with the replaced residual
            if (!IC(I_FIRST) && SC(S_RHOP) != 0.0 && SC(S_OMEGA) != 0.0 && r1 != 0.0 && isfinite(r1))
                SC(S_BETA) = (r1 * SC(S_ALPHA)) / (SC(S_RHOP) * SC(S_OMEGA));
            else
                IC(I_FIRST) = 1;   // breakdown or not started: restart the directions from r
            SC(S_RHO) = r1;
        } else {
            // PCG: the replacement sits between the beta step and the
            // direction update d = beta d + D^-1 r, so beta is recomputed
            // from the replaced residual
            SC(S_RHO) = r1;
            if (SC(S_RHOP) != 0.0 && isfinite(r1) && !IC(I_FIRST)) SC(S_BETA) = r1 / SC(S_RHOP);
            else IC(I_FIRST) = 1;
        }
        if (IC(I_BREAK)) { IC(I_FIRST) = 1; IC(I_BREAK) = 0; }
        IC(I_ACTIVE) = 1;
        if (maxit > 0 && IC(I_ITERS) >= maxit) IC(I_ACTIVE) = 0;
        break;
    }
    case FIN_BICG_INIT:
    case FIN_CG_INIT:
        if (!IC(I_ACTIVE)) break;
        SC(S_BNORM2) = 1.0;   // the inner system is scaled by 1/||b||: absolute criterion
        SC(S_RES2) = (fk == FIN_BICG_INIT) ? r0 : r1;
        SC(S_RHO) = (fk == FIN_BICG_INIT) ? r1 : r0;
        IC(I_FIRST) = 1;
        IC(I_BREAK) = 0;
        if (SC(S_RES2) <= SC(S_TOL2)) IC(I_ACTIVE) = 0;
        break;
    case FIN_BICG_ALPHA: {
        if (!IC(I_ACTIVE)) break;
        double alpha = SC(S_RHO) / r0;
        if (r0 == 0.0 || !isfinite(alpha)) { IC(I_ACTIVE) = 0; IC(I_BREAK) = 1; alpha = 0.0; }
        SC(S_ALPHA) = alpha;
        break;
    }
    case FIN_BICG_OMEGA: {
        if (!IC(I_ACTIVE)) break;
        double om = (r1 == 0.0) ? 0.0 : r0 / r1;
        if (!isfinite(om)) om = 0.0;
        SC(S_OMEGA) = om;
        break;
    }
    case FIN_BICG_RHO: {
        if (!IC(I_ACTIVE)) break;
        IC(I_ITERS) += 1;
        IC(I_FIRST) = 0;
        SC(S_RES2) = r0;
        const double rho_new = r1;
        SC(S_RHOP) = SC(S_RHO);
        if (r0 <= SC(S_TOL2)) { IC(I_ACTIVE) = 0; break; }
        if (SC(S_OMEGA) == 0.0 || rho_new == 0.0 || !isfinite(r0)) { IC(I_ACTIVE) = 0; IC(I_BREAK) = 1; break; }
        if (maxit > 0 && IC(I_ITERS) >= maxit) { IC(I_ACTIVE) = 0; break; }
        SC(S_BETA) = (rho_new * SC(S_ALPHA)) / (SC(S_RHO) * SC(S_OMEGA));
        SC(S_RHOP) = SC(S_RHO);
        SC(S_RHO) = rho_new;
        break;
    }
    case FIN_CG_ALPHA: {
        if (!IC(I_ACTIVE)) break;
        double alpha = SC(S_RHO) / r0;
        if (alpha == 0.0 || !isfinite(alpha)) { IC(I_ACTIVE) = 0; IC(I_BREAK) = 1; alpha = 0.0; }
        SC(S_ALPHA) = alpha;
        break;
    }
    case FIN_CG_BETA: {
        if (!IC(I_ACTIVE)) break;
        IC(I_ITERS) += 1;
        IC(I_FIRST) = 0;
        const double dold = SC(S_RHO), dnew = r0;
        const double beta = dnew / dold;
        SC(S_RHOP) = dold;
        SC(S_RHO) = dnew;
        SC(S_BETA) = beta;
        SC(S_RES2) = r1;
        if (beta == 0.0 || !isfinite(beta)) { IC(I_ACTIVE) = 0; IC(I_BREAK) = 1; break; }
        if (r1 <= SC(S_TOL2)) { IC(I_ACTIVE) = 0; break; }
        if (maxit > 0 && IC(I_ITERS) >= maxit) { IC(I_ACTIVE) = 0; break; }
        break;
    }
    // COCG: rho, alpha, beta are complex (real part in the slot, imaginary
    // part in its _I slot); r0, r1 = re, im of a bilinear product, r2 = ||r||^2
    case FIN_COCG_INIT:
        if (!IC(I_ACTIVE)) break;
        SC(S_BNORM2) = 1.0;   // the inner system is scaled by 1/||b||: absolute criterion
        SC(S_RES2) = r2;
        SC(S_RHO) = r0;
        SC(S_RHO_I) = r1;
        IC(I_FIRST) = 1;
        IC(I_BREAK) = 0;
        if (r2 <= SC(S_TOL2)) IC(I_ACTIVE) = 0;
        break;
    case FIN_COCG_ALPHA: {
        // I_XPEND: the columns whose x update this iteration owes, the ones
        // the update kernel runs for (the direction update applies it when
        // the update kernel defers it)
        IC(I_XPEND) = 0;
        if (!IC(I_ACTIVE)) break;
        double ar, ai;
        if (!cdiv(SC(S_RHO), SC(S_RHO_I), r0, r1, ar, ai) || (ar == 0.0 && ai == 0.0)) {
            IC(I_ACTIVE) = 0;
            IC(I_BREAK) = 1;
            ar = ai = 0.0;
        }
        SC(S_ALPHA) = ar;
        SC(S_ALPHA_I) = ai;
        IC(I_XPEND) = IC(I_ACTIVE);
        break;
    }
    case FIN_COCG_BETA: {
        if (!IC(I_ACTIVE)) break;
        IC(I_ITERS) += 1;
        IC(I_FIRST) = 0;
        double br, bi;
        const bool ok = cdiv(r0, r1, SC(S_RHO), SC(S_RHO_I), br, bi);
        SC(S_RHOP) = SC(S_RHO);
        SC(S_RHOP_I) = SC(S_RHO_I);
        SC(S_RHO) = r0;
        SC(S_RHO_I) = r1;
        SC(S_BETA) = br;
        SC(S_BETA_I) = bi;
        SC(S_RES2) = r2;
        if (!ok || (br == 0.0 && bi == 0.0)) { IC(I_ACTIVE) = 0; IC(I_BREAK) = 1; break; }
        if (r2 <= SC(S_TOL2)) { IC(I_ACTIVE) = 0; break; }
        if (maxit > 0 && IC(I_ITERS) >= maxit) { IC(I_ACTIVE) = 0; break; }
        break;
    }
    case FIN_COCG_RELIABLE: {
        // the replacement sits between the beta step and the direction
        // update, so beta is recomputed from the replaced residual
        if (!I[I_NSLOT]) { IC(I_ACTIVE) = 0; break; }
        SC(S_RES2) = r2;
        SC(S_RHO) = r0;
        SC(S_RHO_I) = r1;
        double br, bi;
        if (!IC(I_FIRST) && cdiv(r0, r1, SC(S_RHOP), SC(S_RHOP_I), br, bi)) {
            SC(S_BETA) = br;
            SC(S_BETA_I) = bi;
        } else {
            IC(I_FIRST) = 1;
        }
        if (IC(I_BREAK)) { IC(I_FIRST) = 1; IC(I_BREAK) = 0; }
        IC(I_ACTIVE) = 1;
        if (maxit > 0 && IC(I_ITERS) >= maxit) IC(I_ACTIVE) = 0;
        break;
    }
    }
#undef SC
#undef IC
}

// --------------------------------------------------------------- finalize
// One workgroup per column: fixed-order sum of the per-block partials, then
// the scalar recurrence of that column.
__global__ void __launch_bounds__(256) k_finalize(int fk, int G, int K, const double *__restrict__ partial,
                                                  double *__restrict__ sc, int *__restrict__ ic, int maxit,
                                                  double tol2, double inner_tol2, const double *__restrict__ osc,
                                                  const int *__restrict__ oic)
{
    const int c = blockIdx.x;
    __shared__ double red[RED_SLOTS][4];
    __shared__ double sS[S_NSLOT];
    __shared__ int sI[I_NSLOT + 1];
    // the column's scalars and flags come in with one parallel load, beside
    // the partial sums, instead of one dependent load per use below
    if ((int)threadIdx.x < S_NSLOT) sS[threadIdx.x] = sc[threadIdx.x * K + c];
    else if ((int)threadIdx.x >= 64 && (int)threadIdx.x < 64 + I_NSLOT) sI[threadIdx.x - 64] = ic[(threadIdx.x - 64) * K + c];
    else if ((int)threadIdx.x == 64 + I_NSLOT && oic) sI[I_NSLOT] = oic[I_ACTIVE * K + c];
    // fixed order: per thread b = tid, tid + 256, ...; then a butterfly over
    // the wave's lanes; then the four waves in order
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nslot = (fk == FIN_COCG_INIT || fk == FIN_COCG_BETA || fk == FIN_COCG_RELIABLE) ? 3 : 2;
    // every slot's first FIN_PER partials per thread are loaded before any
    // is summed (one memory round trip instead of one per slot and chunk);
    // the sum order per thread stays b = tid, tid + 256, ...
    // (indices past G are clamped to G - 1, in bounds, and not summed)
    constexpr int FIN_PER = 8;
    double v[RED_SLOTS][FIN_PER];
#pragma unroll
    for (int slot = 0; slot < RED_SLOTS; slot++)
#pragma unroll
        for (int i = 0; i < FIN_PER; i++) {
            const int b = min((int)threadIdx.x + i * 256, G - 1);
            v[slot][i] = slot < nslot ? partial[((size_t)slot * K + c) * G + b] : 0.0;
        }
#pragma unroll
    for (int slot = 0; slot < RED_SLOTS; slot++) {
        if (slot >= nslot) break;
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < FIN_PER; i++) {
            const double t = s + v[slot][i];
            s = (int)threadIdx.x + i * 256 < G ? t : s;
        }
        for (int b = threadIdx.x + FIN_PER * 256; b < G; b += 256) s += partial[((size_t)slot * K + c) * G + b];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) red[slot][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        fin_column(fk, (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]), (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]),
                   nslot == 3 ? (red[2][0] + red[2][1]) + (red[2][2] + red[2][3]) : 0.0, sS, sI, maxit, tol2);
    __syncthreads();
    if ((int)threadIdx.x < S_NSLOT) sc[threadIdx.x * K + c] = sS[threadIdx.x];
    else if ((int)threadIdx.x >= 64 && (int)threadIdx.x < 64 + I_NSLOT) ic[(threadIdx.x - 64) * K + c] = sI[threadIdx.x - 64];
}


// ------------------------------------------------------------- utilities
template <class V>
__global__ void k_inv_diag(int n, const int *__restrict__ diagpos, const double *__restrict__ val64,
                           V *__restrict__ idiag);
template <>
__global__ void k_inv_diag<double2>(int n, const int *__restrict__ diagpos, const double *__restrict__ val64,
                                    double2 *__restrict__ idiag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = diagpos[i];
    double a = k >= 0 ? val64[2 * (size_t)k] : 0.0, b = k >= 0 ? val64[2 * (size_t)k + 1] : 0.0;
    if (a == 0.0 && b == 0.0) { idiag[i] = make_double2(1.0, 0.0); return; }
    double den = a * a + b * b;
    idiag[i] = make_double2(a / den, -b / den);
}
template <>
__global__ void k_inv_diag<float2>(int n, const int *__restrict__ diagpos, const double *__restrict__ val64,
                                   float2 *__restrict__ idiag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = diagpos[i];
    double a = k >= 0 ? val64[2 * (size_t)k] : 0.0, b = k >= 0 ? val64[2 * (size_t)k + 1] : 0.0;
    if (a == 0.0 && b == 0.0) { idiag[i] = make_float2(1.f, 0.f); return; }
    double den = a * a + b * b;
    idiag[i] = make_float2((float)(a / den), (float)(-b / den));
}
template <>
__global__ void k_inv_diag<double>(int n, const int *__restrict__ diagpos, const double *__restrict__ val64,
                                   double *__restrict__ idiag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = diagpos[i];
    double a = k >= 0 ? val64[k] : 0.0;
    idiag[i] = a != 0.0 ? 1.0 / a : 1.0;
}
template <>
__global__ void k_inv_diag<float>(int n, const int *__restrict__ diagpos, const double *__restrict__ val64,
                                  float *__restrict__ idiag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = diagpos[i];
    double a = k >= 0 ? val64[k] : 0.0;
    idiag[i] = a != 0.0 ? (float)(1.0 / a) : 1.f;
}

// right-preconditioned operator A D^-1 (column scaling), built once per Reset
template <class V>
__global__ void k_scale_cols(int nnz, const int *__restrict__ colidx, const V *__restrict__ val,
                             const V *__restrict__ idiag, V *__restrict__ out64, void *__restrict__ out32)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    V v = vmul(val[k], idiag[colidx[k]]);
    out64[k] = v;
    if (out32) ((decltype(cvt32(v)) *)out32)[k] = cvt32(v);
}

__global__ void k_f64_to_f32(const double *__restrict__ s, float *__restrict__ d, size_t cnt)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < cnt; i += (size_t)gridDim.x * blockDim.x) d[i] = (float)s[i];
}

__global__ void k_permute_vals(int cnt, int w, const int *__restrict__ slot, const double *__restrict__ v64,
                               double *__restrict__ o64, float *__restrict__ o32)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    const size_t src = (size_t)slot[k] * w;
    for (int t = 0; t < w; t++) {
        const double v = v64[src + t];
        o64[(size_t)k * w + t] = v;
        if (o32) o32[(size_t)k * w + t] = (float)v;
    }
}

template <class T>
__global__ void k_gather_tile(int n, int K, int kc, const int *__restrict__ perm, const T *__restrict__ src, long ld,
                              int q0, T *__restrict__ dst)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * K) return;
    const int r = (int)(e / K), c = (int)(e % K);
    T v{};
    if (c < kc) v = src[(size_t)perm[r] * ld + q0 + c];
    dst[e] = v;
}

template <class T>
__global__ void k_scatter_tile(int n, int K, int kc, const int *__restrict__ perm, const T *__restrict__ src,
                               T *__restrict__ dst, long ld, int q0)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * K) return;
    const int r = (int)(e / K), c = (int)(e % K);
    if (c < kc) dst[(size_t)perm[r] * ld + q0 + c] = src[e];
}

// column compaction: dst[r][j] = src[r][sel[j]] for j < nsel, 0 otherwise
template <class T>
__global__ void k_select_cols(long rows, int ks, int kd, int nsel, const int *__restrict__ sel,
                              const T *__restrict__ src, T *__restrict__ dst)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * kd) return;
    const long r = e / kd;
    const int j = (int)(e % kd);
    T v{};
    if (j < nsel) v = src[r * ks + sel[j]];
    dst[e] = v;
}

// scatter tile columns with a column map: dst[perm[r]][q0 + cmap[j]] = src[r][j]
// for every slot j with cmap[j] >= 0
template <class T>
__global__ void k_scatter_map(int n, int K, const int *__restrict__ perm, const int *__restrict__ cmap,
                              const T *__restrict__ src, T *__restrict__ dst, long ld, int q0)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * K) return;
    const int r = (int)(e / K), j = (int)(e % K);
    const int c = cmap[j];
    if (c >= 0) dst[(size_t)perm[r] * ld + q0 + c] = src[e];
}

// ================================================================ launchers
static inline void check_K(int K)
{
    TMI_CHECK(K == 1 || K == 2 || K == 4 || K == 8 || K == 16 || K == 32 || K == 64, TMI_ERR_INVALID_ARG,
              "block width %d is not a power of two <= 64", K);
}

#define DISPATCH_K(K, MACRO)          \
    switch (K) {                      \
    case 1: MACRO(1); break;          \
    case 2: MACRO(2); break;          \
    case 4: MACRO(4); break;          \
    case 8: MACRO(8); break;          \
    case 16: MACRO(16); break;        \
    case 32: MACRO(32); break;        \
    default: MACRO(64); break;        \
    }

#define DISPATCH_V(vk, ...)                                    \
    switch (vk) {                                              \
    case VK_C32: { using V = float2; __VA_ARGS__; } break;     \
    case VK_C64: { using V = double2; __VA_ARGS__; } break;    \
    case VK_R32: { using V = float; __VA_ARGS__; } break;      \
    default: { using V = double; __VA_ARGS__; } break;         \
    }

// kernels that exist for complex vectors only (COCG, per-column frequencies)
#define DISPATCH_C(vk, ...)                                                         \
    switch (vk) {                                                                   \
    case VK_C32: { using V = float2; __VA_ARGS__; } break;                          \
    case VK_C64: { using V = double2; __VA_ARGS__; } break;                         \
    default: throw Error(TMI_ERR_UNSUPPORTED, "this path needs complex vectors");  \
    }

void launch_spmm_dot(VKind vk, const BlockDesc &b, const void *val, const void *idiag, const void *X, void *Y,
                     const void *W, int want_yy, const int *iflags, double *partial, hipStream_t st,
                     const unsigned char *Wq, const double *omg)
{
    check_K(b.K);
    dim3 g(b.nblk), blk(256);
    if (omg) {
        // per-column frequencies: the row-block kernel (its coefficients
        // gathered from this operator) or the per-row kernels, MF variants
        TMI_CHECK(!idiag && !Wq && b.K >= 2, TMI_ERR_UNSUPPORTED,
                  "per-column frequencies: unpreconditioned operator, tiles of 2 or more columns");
#define MFK(KK)                                                                                                 \
    if (KK == 64 && vk == VK_C32 && b.rb.nb > 0 && b.rb.src == val)                                            \
        hipLaunchKernelGGL((k_spmm_rb<float2, 0, true>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr,     \
                           b.rb.ofs, b.rb.col, b.rb.mask, (const float2 *)b.rb.val, (const float2 *)X,         \
                           (float2 *)Y, (const float2 *)W, want_yy, iflags, partial, nullptr, omg);          \
    else if (KK == 32 && vk == VK_C32 && b.rb.nb > 0 && b.rb.src == val)                                       \
        hipLaunchKernelGGL((k_spmm_rb<float2, 0, true, 32>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr, \
                           b.rb.ofs, b.rb.col, b.rb.mask, (const float2 *)b.rb.val, (const float2 *)X,         \
                           (float2 *)Y, (const float2 *)W, want_yy, iflags, partial, nullptr, omg);          \
    else if (KK == 64)                                                                                          \
        hipLaunchKernelGGL((k_spmm64<V, 0, true>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr, b.colidx, \
                           (const V *)val, (const V *)X, (V *)Y, (const V *)W, want_yy, iflags, partial, b.rpw,   \
                           omg);                                                                                \
    else if (KK >= NK_KMIN && KK <= 32)                                                                         \
        hipLaunchKernelGGL((k_spmm_nk<V, nk_k(KK), 0, true>), g, blk, 0, st, b.n, b.rowptr, b.colidx,           \
                           (const V *)val, (const V *)X, (V *)Y, (const V *)W, want_yy, iflags, partial, omg);   \
    else                                                                                                        \
        hipLaunchKernelGGL((k_spmm_dot<V, KK, false, true>), g, blk, 0, st, b.n, b.rowptr, b.colidx,            \
                           (const V *)val, nullptr, (const V *)X, (V *)Y, (const V *)W, want_yy, iflags, partial, \
                           omg);
        DISPATCH_C(vk, DISPATCH_K(b.K, MFK))
#undef MFK
        TMI_LAUNCH_CHECK();
        return;
    }
    if (b.K == 1) {
        // lanes per row sized to the average row length
        int avg = b.n ? 0 : 1;
        (void)avg;
#define SPMV1                                                                                                   \
    hipLaunchKernelGGL((k_spmv_dot<V, 16>), g, blk, 0, st, b.n, b.rowptr, b.colidx, (const V *)val,           \
                       (const V *)idiag, (const V *)X, (V *)Y, (const V *)W, (int)want_yy, iflags, partial);
        DISPATCH_V(vk, SPMV1)
#undef SPMV1
    } else {
#define SPMMK(KK)                                                                                             \
    if ((KK == 64 || KK == 32 || KK == 16) && !idiag && vk == VK_C32 && b.rb.nb > 0 && b.rb.src == val &&     \
        b.rb.word && b.rb.mfma) {                                                                             \
        if (KK == 64 && Wq)                                                                                   \
            hipLaunchKernelGGL((k_spmm_rbm<4, true>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr, b.rb.ofs, \
                               b.rb.col, b.rb.word, (const float2 *)b.rb.val, (const float2 *)X, (float2 *)Y,  \
                               (const float2 *)W, (int)want_yy, iflags, partial, Wq);                         \
        else                                                                                                  \
            hipLaunchKernelGGL((k_spmm_rbm<(KK >= 16 ? KK / 16 : 1), false>), g, dim3(64 * SPMM64_WAVES), 0, st, \
                               b.n, b.rowptr, b.rb.ofs, b.rb.col, b.rb.word, (const float2 *)b.rb.val,         \
                               (const float2 *)X, (float2 *)Y, (const float2 *)W, (int)want_yy, iflags, partial, \
                               nullptr);                                                                      \
    } else if (KK == 64 && !idiag && vk == VK_C32 && b.rb.nb > 0 && b.rb.src == val && b.rb.word && Wq)      \
        hipLaunchKernelGGL((k_spmm_rbp<0, true>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr, b.rb.ofs, \
                           b.rb.col, b.rb.word, (const float2 *)b.rb.val, (const float2 *)X, (float2 *)Y,      \
                           (const float2 *)W, (int)want_yy, iflags, partial, Wq);                           \
    else if (KK == 64 && !idiag && vk == VK_C32 && b.rb.nb > 0 && b.rb.src == val && b.rb.word)              \
        hipLaunchKernelGGL((k_spmm_rbp<0, false>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr,         \
                           b.rb.ofs, b.rb.col, b.rb.word, (const float2 *)b.rb.val, (const float2 *)X,         \
                           (float2 *)Y, (const float2 *)W, (int)want_yy, iflags, partial, Wq);              \
    else if (KK == 64 && !idiag && vk == VK_C32 && b.rb.nb > 0 && b.rb.src == val)                           \
        hipLaunchKernelGGL((k_spmm_rb<float2, 0>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr, b.rb.ofs,   \
                           b.rb.col, b.rb.mask, (const float2 *)b.rb.val, (const float2 *)X, (float2 *)Y,      \
                           (const float2 *)W, (int)want_yy, iflags, partial, Wq);                           \
    else if (KK == 64 && !idiag && vk == VK_C32 && b.lds.nb > 0 && b.lds.dma && want_yy != DOT_BIL) {      \
        if (want_yy)                                                                                          \
            hipLaunchKernelGGL((k_spmm_dma<true>), dim3(std::min(DMA_GRID, b.nblk)), dim3(DMA_NT), 0, st, b.lds.nb,       \
                               b.lds.brow, b.lds.bu, b.lds.ucols, b.rowptr, b.lds.lcol, (const float2 *)val, \
                               (const float2 *)X, (float2 *)Y, (const float2 *)W, iflags, partial, b.nblk);  \
        else                                                                                                  \
            hipLaunchKernelGGL((k_spmm_dma<false>), dim3(std::min(DMA_GRID, b.nblk)), dim3(DMA_NT), 0, st, b.lds.nb,      \
                               b.lds.brow, b.lds.bu, b.lds.ucols, b.rowptr, b.lds.lcol, (const float2 *)val, \
                               (const float2 *)X, (float2 *)Y, (const float2 *)W, iflags, partial, b.nblk);  \
    } else if (KK == 64 && !idiag && vk == VK_C32 && b.lds.nb > 0)                                           \
        hipLaunchKernelGGL((k_spmm_lds<float2, 0, LDS_ROWS_C32, LDS_WAVES, LDS_NZ_C32>), g, dim3(64 * LDS_WAVES), \
                           0, st,                                                                             \
                           b.lds.nb, b.lds.brow, b.lds.bu,                                                    \
                           b.lds.ucols, b.rowptr, b.lds.lcol, (const float2 *)val, (const float2 *)X,          \
                           (float2 *)Y, (const float2 *)W, (int)want_yy, iflags, partial);                    \
    else if (KK == 64 && !idiag && vk == VK_C64 && b.rb.nb > 0 && b.rb.val64 && b.rb.src64 == val)          \
        hipLaunchKernelGGL((k_spmm_rb<double2, 0>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr, b.rb.ofs, \
                           b.rb.col, b.rb.mask, (const double2 *)b.rb.val64, (const double2 *)X, (double2 *)Y, \
                           (const double2 *)W, (int)want_yy, iflags, partial, nullptr);                       \
    else if (KK == 64 && !idiag && (vk == VK_C32 || vk == VK_C64))                                           \
        hipLaunchKernelGGL((k_spmm64<V, 0>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr, b.colidx, (const V *)val, (const V *)X, \
                           (V *)Y, (const V *)W, (int)want_yy, iflags, partial, b.rpw);                      \
    else if (KK == 32 && !idiag && vk == VK_C32 && b.rb.nb > 0 && b.rb.src == val)                           \
        hipLaunchKernelGGL((k_spmm_rb<float2, 0, false, 32>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr, \
                           b.rb.ofs, b.rb.col, b.rb.mask, (const float2 *)b.rb.val, (const float2 *)X,          \
                           (float2 *)Y, (const float2 *)W, (int)want_yy, iflags, partial, nullptr);         \
    else if (idiag)                                                                                                \
        hipLaunchKernelGGL((k_spmm_dot<V, KK, true>), g, blk, 0, st, b.n, b.rowptr, b.colidx, (const V *)val,  \
                           (const V *)idiag, (const V *)X, (V *)Y, (const V *)W, (int)want_yy, iflags, partial); \
    else if (KK >= NK_KMIN && KK <= 32)                                                                       \
        hipLaunchKernelGGL((k_spmm_nk<V, nk_k(KK), 0>), g, blk, 0, st, b.n, b.rowptr, b.colidx, (const V *)val,  \
                           (const V *)X, (V *)Y, (const V *)W, (int)want_yy, iflags, partial);               \
    else                                                                                                      \
        hipLaunchKernelGGL((k_spmm_dot<V, KK, false>), g, blk, 0, st, b.n, b.rowptr, b.colidx, (const V *)val, \
                           (const V *)idiag, (const V *)X, (V *)Y, (const V *)W, (int)want_yy, iflags, partial);
        DISPATCH_V(vk, DISPATCH_K(b.K, SPMMK))
#undef SPMMK
    }
    TMI_LAUNCH_CHECK();
}

void launch_resid(VKind vk, const BlockDesc &b, const void *val, const void *X, const void *B, void *R,
                  const int *iflags, double *partial, hipStream_t st, const double *omg)
{
    check_K(b.K);
    dim3 g(b.nblk), blk(256);
    if (omg) {
        TMI_CHECK(b.K >= 2, TMI_ERR_UNSUPPORTED, "per-column frequencies: tiles of 2 or more columns");
#define MFR(KK)                                                                                                 \
    if (KK == 64)                                                                                               \
        hipLaunchKernelGGL((k_spmm64<V, 1, true>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr, b.colidx, \
                           (const V *)val, (const V *)X, (V *)R, (const V *)B, 0, iflags, partial, b.rpw, omg);   \
    else if (KK >= NK_KMIN && KK <= 32)                                                                         \
        hipLaunchKernelGGL((k_spmm_nk<V, nk_k(KK), 1, true>), g, blk, 0, st, b.n, b.rowptr, b.colidx,           \
                           (const V *)val, (const V *)X, (V *)R, (const V *)B, 0, iflags, partial, omg);         \
    else                                                                                                        \
        hipLaunchKernelGGL((k_resid<V, KK, true>), g, blk, 0, st, b.n, b.rowptr, b.colidx, (const V *)val,      \
                           (const V *)X, (const V *)B, (V *)R, iflags, partial, omg);
        DISPATCH_C(vk, DISPATCH_K(b.K, MFR))
#undef MFR
        TMI_LAUNCH_CHECK();
        return;
    }
#define RESK(KK)                                                                                              \
    if (KK == 64 && vk == VK_C64 && b.rb.nb > 0 && b.rb.val64 && b.rb.src64 == val)                            \
        hipLaunchKernelGGL((k_spmm_rb<double2, 1>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr, b.rb.ofs, \
                           b.rb.col, b.rb.mask, (const double2 *)b.rb.val64, (const double2 *)X, (double2 *)R,  \
                           (const double2 *)B, 0, iflags, partial, nullptr);                                   \
    else if (KK == 64 && (vk == VK_C32 || vk == VK_C64))                                                      \
        hipLaunchKernelGGL((k_spmm64<V, 1>), g, dim3(64 * SPMM64_WAVES), 0, st, b.n, b.rowptr, b.colidx, (const V *)val, (const V *)X, \
                           (V *)R, (const V *)B, 0, iflags, partial, b.rpw);                                 \
    else if (KK >= NK_KMIN && KK <= 32)                                                                       \
        hipLaunchKernelGGL((k_spmm_nk<V, nk_k(KK), 1>), g, blk, 0, st, b.n, b.rowptr, b.colidx, (const V *)val,  \
                           (const V *)X, (V *)R, (const V *)B, 0, iflags, partial);                          \
    else                                                                                                      \
    hipLaunchKernelGGL((k_resid<V, KK>), g, blk, 0, st, b.n, b.rowptr, b.colidx, (const V *)val, (const V *)X, \
                       (const V *)B, (V *)R, iflags, partial);
    DISPATCH_V(vk, DISPATCH_K(b.K, RESK))
#undef RESK
    TMI_LAUNCH_CHECK();
}

// grid of the streaming kernels without partial sums that take EW_U
// elements per thread (ew_for): one chunk per thread, no grid-stride loop
static dim3 ew_chunk_grid(const BlockDesc &b)
{
    const size_t tot = (size_t)b.n * b.K, per = (size_t)256 * EW_U;
    TMI_CHECK(tot / per < (size_t)INT32_MAX, TMI_ERR_INVALID_ARG, "block vector of %zu elements is too large", tot);
    return dim3((unsigned)std::max<size_t>(1, (tot + per - 1) / per));
}

// grid of the one-element-per-thread kernels
static dim3 ew1_grid(const BlockDesc &b)
{
    const size_t tot = (size_t)b.n * b.K;
    TMI_CHECK(tot / 256 < (size_t)INT32_MAX, TMI_ERR_INVALID_ARG, "block vector of %zu elements is too large", tot);
    return dim3((unsigned)((tot + 255) / 256));
}

void launch_bicg_p(VKind vk, const BlockDesc &b, const void *r, void *p, const void *v, const double *sc,
                   const int *ic, hipStream_t st)
{
    dim3 g = ew1_grid(b), blk(256);
#define PK(KK) hipLaunchKernelGGL((k_bicg_p<V, KK>), g, blk, 0, st, b.n, (const V *)r, (V *)p, (const V *)v, sc, ic);
    DISPATCH_V(vk, DISPATCH_K(b.K, PK))
#undef PK
    TMI_LAUNCH_CHECK();
}

void launch_bicg_s(VKind vk, const BlockDesc &b, const void *r, const void *v, void *s, const double *sc,
                   const int *ic, hipStream_t st)
{
    dim3 g = ew1_grid(b), blk(256);
#define SK(KK) hipLaunchKernelGGL((k_bicg_s<V, KK>), g, blk, 0, st, b.n, (const V *)r, (const V *)v, (V *)s, sc, ic);
    DISPATCH_V(vk, DISPATCH_K(b.K, SK))
#undef SK
    TMI_LAUNCH_CHECK();
}

void launch_bicg_xr(VKind vk, const BlockDesc &b, const void *idiag, void *x, const void *p, const void *s,
                    const void *t, void *r, const void *rd, const double *sc, const int *ic, double *partial,
                    hipStream_t st, const unsigned char *rdq)
{
    dim3 g(b.nblk), blk(256);
    if (b.K != 64 || vk != VK_C32) rdq = nullptr;
#define XK(KK)                                                                                                \
    hipLaunchKernelGGL((k_bicg_xr<V, KK>), g, blk, 0, st, b.n, (const V *)idiag, (V *)x, (const V *)p,         \
                       (const V *)s, (const V *)t, (V *)r, (const V *)rd, sc, ic, partial, rdq);
    DISPATCH_V(vk, DISPATCH_K(b.K, XK))
#undef XK
    TMI_LAUNCH_CHECK();
}

void launch_cg_xr(VKind vk, const BlockDesc &b, const void *idiag, void *x, const void *d, const void *q, void *r,
                  const double *sc, const int *ic, double *partial, hipStream_t st, bool bil, const double *omg,
                  bool xdefer)
{
    TMI_CHECK(!xdefer || bil, TMI_ERR_UNSUPPORTED, "the deferred x update is COCG's");
    TMI_CHECK(!omg || bil, TMI_ERR_UNSUPPORTED, "per-column frequencies need COCG");
    dim3 g(b.nblk), blk(256);
#define CXK(KK)                                                                                               \
    hipLaunchKernelGGL((k_cg_xr<V, KK>), g, blk, 0, st, b.n, (const V *)idiag, (V *)x, (const V *)d,           \
                       (const V *)q, (V *)r, sc, ic, partial);
#define BXK(KK)                                                                                               \
    if (xdefer)                                                                                               \
        hipLaunchKernelGGL((k_cocg_xr<V, KK, true>), g, blk, 0, st, b.n, (const V *)idiag, (V *)x, (const V *)d, \
                           (const V *)q, (V *)r, sc, ic, partial, omg);                                       \
    else                                                                                                      \
        hipLaunchKernelGGL((k_cocg_xr<V, KK>), g, blk, 0, st, b.n, (const V *)idiag, (V *)x, (const V *)d,     \
                           (const V *)q, (V *)r, sc, ic, partial, omg);
    if (bil) {
        DISPATCH_C(vk, DISPATCH_K(b.K, BXK))
    } else {
        DISPATCH_V(vk, DISPATCH_K(b.K, CXK))
    }
#undef CXK
#undef BXK
    TMI_LAUNCH_CHECK();
}

void launch_cg_d(VKind vk, const BlockDesc &b, const void *idiag, void *d, const void *r, const double *sc,
                 const int *ic, hipStream_t st, bool bil, const double *omg, void *x)
{
    if (!x) {
        launch_cg_init(vk, b, idiag, d, r, sc, ic, nullptr, st, bil, omg);
        return;
    }
    TMI_CHECK(bil, TMI_ERR_UNSUPPORTED, "the deferred x update is COCG's");
    dim3 g = ew_chunk_grid(b), blk(256);
#define BDK(KK)                                                                                               \
    hipLaunchKernelGGL((k_cocg_d<V, KK, true>), g, blk, 0, st, b.n, (const V *)idiag, (V *)d, (const V *)r, sc, ic, \
                       nullptr, omg, (V *)x);
    DISPATCH_C(vk, DISPATCH_K(b.K, BDK))
#undef BDK
    TMI_LAUNCH_CHECK();
}

void launch_cg_init(VKind vk, const BlockDesc &b, const void *idiag, void *d, const void *r, const double *sc,
                    const int *ic, double *partial, hipStream_t st, bool bil, const double *omg)
{
    TMI_CHECK(!omg || bil, TMI_ERR_UNSUPPORTED, "per-column frequencies need COCG");
    // without partial sums (COCG's direction update) the grid is free: one
    // chunk per thread; with them, the reduction's grid
    dim3 g = (bil && !partial) ? ew_chunk_grid(b) : dim3(b.nblk), blk(256);
#define CIK(KK) \
    hipLaunchKernelGGL((k_cg_d<V, KK>), g, blk, 0, st, b.n, (const V *)idiag, (V *)d, (const V *)r, sc, ic, partial);
#define BIK(KK) \
    hipLaunchKernelGGL((k_cocg_d<V, KK>), g, blk, 0, st, b.n, (const V *)idiag, (V *)d, (const V *)r, sc, ic, partial, omg, \
                       (V *)nullptr);
    if (bil) {
        DISPATCH_C(vk, DISPATCH_K(b.K, BIK))
    } else {
        DISPATCH_V(vk, DISPATCH_K(b.K, CIK))
    }
#undef CIK
#undef BIK
    TMI_LAUNCH_CHECK();
}

void launch_finalize(FinKind fk, const BlockDesc &b, const double *partial, double *sc, int *ic, int maxit,
                     double tol2, double inner_tol2, hipStream_t st, const double *osc, const int *oic)
{
    hipLaunchKernelGGL(k_finalize, dim3(b.K), dim3(256), 0, st, (int)fk, b.nblk, b.K, partial, sc, ic, maxit, tol2,
                       inner_tol2, osc, oic);
    TMI_LAUNCH_CHECK();
}

__global__ void k_inner_reset(int K, const double *__restrict__ osc, const int *__restrict__ oic,
                              double *__restrict__ sc, int *__restrict__ ic)
{
    int c = threadIdx.x;
    if (c >= K) return;
    ic[I_ACTIVE * K + c] = oic[I_ACTIVE * K + c];
    ic[I_ITERS * K + c] = 0;
    ic[I_FIRST * K + c] = 1;
    ic[I_BREAK * K + c] = 0;
    sc[S_TOL2 * K + c] = osc[S_AUX * K + c];
    sc[S_RHOP * K + c] = 0.0;
    sc[S_BETA * K + c] = 0.0;
    sc[S_OMEGA * K + c] = 0.0;
    sc[S_ALPHA * K + c] = 0.0;
    sc[S_RHOP_I * K + c] = 0.0;
    sc[S_BETA_I * K + c] = 0.0;
    sc[S_ALPHA_I * K + c] = 0.0;
}

void launch_inner_reset(const BlockDesc &b, const double *osc, const int *oic, double *sc, int *ic, hipStream_t st)
{
    hipLaunchKernelGGL(k_inner_reset, dim3(1), dim3(64), 0, st, b.K, osc, oic, sc, ic);
    TMI_LAUNCH_CHECK();
}

void launch_pack_shadow(int n, const void *rd, unsigned char *q, hipStream_t st)
{
    const size_t cnt = (size_t)n * 32;
    hipLaunchKernelGGL(k_pack_shadow, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, n, (const float2 *)rd, q);
    TMI_LAUNCH_CHECK();
}

void launch_to_inner(VKind vk32, const BlockDesc &b, const void *r64, void *r32, void *rd32, void *x32,
                     const double *scale, const int *oflags, double *partial, hipStream_t st, bool sign_shadow)
{
    dim3 g(b.nblk), blk(256);
#define TIK(KK)                                                                                                \
    hipLaunchKernelGGL((k_to_inner<V64, V, KK>), g, blk, 0, st, b.n, (const V64 *)r64, (V *)r32, (V *)rd32,   \
                       (V *)x32, scale, oflags, partial, sign_shadow ? 1 : 0);
    switch (vk32) {
    case VK_C32: { using V = float2; using V64 = double2; DISPATCH_K(b.K, TIK) break; }
    case VK_C64: { using V = double2; using V64 = double2; DISPATCH_K(b.K, TIK) break; }
    case VK_R32: { using V = float; using V64 = double; DISPATCH_K(b.K, TIK) break; }
    default: { using V = double; using V64 = double; DISPATCH_K(b.K, TIK) break; }
    }
#undef TIK
    TMI_LAUNCH_CHECK();
}

void launch_from_inner(VKind vk32, const BlockDesc &b, void *x64, const void *x32, const double *scale,
                       const int *oflags, hipStream_t st)
{
    dim3 g = ew1_grid(b), blk(256);
#define FIK(KK)                                                                                              \
    hipLaunchKernelGGL((k_from_inner<V64, V, KK>), g, blk, 0, st, b.n, (V64 *)x64, (const V *)x32, scale,   \
                       oflags);
    switch (vk32) {
    case VK_C32: { using V = float2; using V64 = double2; DISPATCH_K(b.K, FIK) break; }
    case VK_C64: { using V = double2; using V64 = double2; DISPATCH_K(b.K, FIK) break; }
    case VK_R32: { using V = float; using V64 = double; DISPATCH_K(b.K, FIK) break; }
    default: { using V = double; using V64 = double; DISPATCH_K(b.K, FIK) break; }
    }
#undef FIK
    TMI_LAUNCH_CHECK();
}

void launch_reliable(VKind vk_in, const BlockDesc &b, const void *r64, void *rin, void *xin, const void *rd,
                     const void *idiag, const double *scale, const int *oflags, double *partial, hipStream_t st,
                     bool bil, const double *omg)
{
    dim3 g(b.nblk), blk(256);
    TMI_CHECK(!omg || bil, TMI_ERR_UNSUPPORTED, "per-column frequencies need COCG");
    if (bil && vk_in != VK_C32 && vk_in != VK_C64) throw Error(TMI_ERR_UNSUPPORTED, "COCG needs complex vectors");
#define RLK(KK)                                                                                                 \
    hipLaunchKernelGGL((k_reliable<V64, V, KK>), g, blk, 0, st, b.n, (const V64 *)r64, (V *)rin, (V *)xin,      \
                       (const V *)rd, (const V *)idiag, scale, oflags, partial, bil ? 1 : 0, omg);
    switch (vk_in) {
    case VK_C32: { using V = float2; using V64 = double2; DISPATCH_K(b.K, RLK) break; }
    case VK_C64: { using V = double2; using V64 = double2; DISPATCH_K(b.K, RLK) break; }
    case VK_R32: { using V = float; using V64 = double; DISPATCH_K(b.K, RLK) break; }
    default: { using V = double; using V64 = double; DISPATCH_K(b.K, RLK) break; }
    }
#undef RLK
    TMI_LAUNCH_CHECK();
}

void launch_norm2(VKind vk, const BlockDesc &b, const void *X, const int *flags, double *partial, hipStream_t st)
{
    dim3 g(b.nblk), blk(256);
#define NK(KK) hipLaunchKernelGGL((k_norm2<V, KK>), g, blk, 0, st, b.n, (const V *)X, flags, partial);
    DISPATCH_V(vk, DISPATCH_K(b.K, NK))
#undef NK
    TMI_LAUNCH_CHECK();
}

// the complex diagonal entries (not inverted), converted to the vector type
template <class V>
__global__ void k_get_diag(int n, const int *__restrict__ diagpos, const double *__restrict__ val64, V *__restrict__ dg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long k = diagpos[i];
    dg[i] = cvt<V>(make_double2(val64[2 * k], val64[2 * k + 1]));
}

void launch_get_diag(VKind vk, int n, const int *diagpos, const double *val64, void *dg, hipStream_t st)
{
    dim3 g(grid_for(n, 256)), blk(256);
#define GDV hipLaunchKernelGGL((k_get_diag<V>), g, blk, 0, st, n, diagpos, val64, (V *)dg);
    DISPATCH_C(vk, GDV)
#undef GDV
    TMI_LAUNCH_CHECK();
}

void launch_inv_diag(VKind vk, int n, const int *diagpos, const double *val64, void *idiag, hipStream_t st)
{
    dim3 g(grid_for(n, 256)), blk(256);
#define IDV hipLaunchKernelGGL((k_inv_diag<V>), g, blk, 0, st, n, diagpos, val64, (V *)idiag);
    DISPATCH_V(vk, IDV)
    TMI_LAUNCH_CHECK();
}

void launch_scale_cols(bool cplx, int nnz, const int *colidx, const double *val64, const void *idiag64, double *out64,
                       void *out32, hipStream_t st)
{
    dim3 g(grid_for(nnz, 256)), blk(256);
    if (cplx)
        hipLaunchKernelGGL((k_scale_cols<double2>), g, blk, 0, st, nnz, colidx, (const double2 *)val64,
                           (const double2 *)idiag64, (double2 *)out64, out32);
    else
        hipLaunchKernelGGL((k_scale_cols<double>), g, blk, 0, st, nnz, colidx, (const double *)val64,
                           (const double *)idiag64, (double *)out64, out32);
    TMI_LAUNCH_CHECK();
}

void launch_f64_to_f32(const double *src, float *dst, size_t count, hipStream_t st)
{
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(count, 256, 4096)), dim3(256), 0, st, src, dst, count);
    TMI_LAUNCH_CHECK();
}

void launch_permute_vals(bool cplx, int nnz, const int *slot, const double *val64, double *out64, float *out32,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_permute_vals, dim3(grid_for(nnz, 256)), dim3(256), 0, st, nnz, cplx ? 2 : 1, slot, val64,
                       out64, out32);
    TMI_LAUNCH_CHECK();
}

__global__ void k_gather_vals32(int cnt, int w, const int *__restrict__ slot, const float *__restrict__ in,
                                float *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    for (int t = 0; t < w; t++) out[(size_t)k * w + t] = in[(size_t)slot[k] * w + t];
}

void launch_gather_vals32(int nnz, int w, const int *slot, const float *in, float *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_gather_vals32, dim3(grid_for(nnz, 256)), dim3(256), 0, st, nnz, w, slot, in, out);
    TMI_LAUNCH_CHECK();
}

__global__ void k_gather_vals64(int cnt, int w, const int *__restrict__ slot, const double *__restrict__ in,
                                double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    for (int t = 0; t < w; t++) out[(size_t)k * w + t] = in[(size_t)slot[k] * w + t];
}

void launch_gather_vals64(int nnz, int w, const int *slot, const double *in, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_gather_vals64, dim3(grid_for(nnz, 256)), dim3(256), 0, st, nnz, w, slot, in, out);
    TMI_LAUNCH_CHECK();
}

void launch_gather_tile(int es, int n, int K, int kc, const int *perm, const void *src, long ld, int q0, void *dst,
                        hipStream_t st)
{
    const size_t cnt = (size_t)n * K;
    dim3 g((unsigned)((cnt + 255) / 256));
    if (es == 16)
        hipLaunchKernelGGL(k_gather_tile<double2>, g, dim3(256), 0, st, n, K, kc, perm, (const double2 *)src, ld, q0,
                           (double2 *)dst);
    else
        hipLaunchKernelGGL(k_gather_tile<double>, g, dim3(256), 0, st, n, K, kc, perm, (const double *)src, ld, q0,
                           (double *)dst);
    TMI_LAUNCH_CHECK();
}

void launch_scatter_tile(int es, int n, int K, int kc, const int *perm, const void *src, void *dst, long ld, int q0,
                         hipStream_t st)
{
    const size_t cnt = (size_t)n * K;
    dim3 g((unsigned)((cnt + 255) / 256));
    if (es == 16)
        hipLaunchKernelGGL(k_scatter_tile<double2>, g, dim3(256), 0, st, n, K, kc, perm, (const double2 *)src,
                           (double2 *)dst, ld, q0);
    else
        hipLaunchKernelGGL(k_scatter_tile<double>, g, dim3(256), 0, st, n, K, kc, perm, (const double *)src,
                           (double *)dst, ld, q0);
    TMI_LAUNCH_CHECK();
}

// 32 x 32 tiles: reads along a row of `in`, writes along a row of `out`
template <class T>
__global__ void __launch_bounds__(256) k_transpose(long rows, int cols, const T *__restrict__ in, T *__restrict__ out)
{
    __shared__ T tile[32][33];
    const long nct = (cols + 31) / 32;   // column tiles; blockIdx.x = row tile * nct + column tile
    const long r0 = (long)(blockIdx.x / nct) * 32;
    const int c0 = (int)(blockIdx.x % nct) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const long r = r0 + k;
        const int c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = in[r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k;
        const long r = r0 + tx;
        if (r < rows && c < cols) out[(long)c * rows + r] = tile[tx][k];
    }
}

void launch_transpose(int elem_bytes, long rows, int cols, const void *in, void *out, hipStream_t st)
{
    TMI_CHECK(rows >= 0 && cols >= 0, TMI_ERR_INVALID_ARG, "transpose of %ld x %d", rows, cols);
    if (rows == 0 || cols == 0) return;
    const long tiles = ((rows + 31) / 32) * (long)((cols + 31) / 32);
    TMI_CHECK(tiles < (1L << 31), TMI_ERR_UNSUPPORTED, "transpose of %ld x %d is too large", rows, cols);
    const dim3 g((unsigned)tiles);
    switch (elem_bytes) {
    case 16:
        hipLaunchKernelGGL(k_transpose<double2>, g, dim3(256), 0, st, rows, cols, (const double2 *)in, (double2 *)out);
        break;
    case 8:
        hipLaunchKernelGGL(k_transpose<double>, g, dim3(256), 0, st, rows, cols, (const double *)in, (double *)out);
        break;
    default:
        TMI_CHECK(false, TMI_ERR_INVALID_ARG, "transpose of %d-byte entries", elem_bytes);
    }
    TMI_LAUNCH_CHECK();
}

// thread per source: its entries added in order, so duplicate columns sum
// as the host loop over the CSR row would
__global__ void k_qvec_dense(int nq, int cw, const int *__restrict__ rp, const int *__restrict__ ci,
                             const double *__restrict__ val, double *__restrict__ B)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    for (int k = rp[q]; k < rp[q + 1]; k++) {
        const size_t at = ((size_t)ci[k] * nq + q) * cw;
        for (int c = 0; c < cw; c++) B[at + c] += val[(size_t)k * cw + c];
    }
}

void launch_qvec_dense(int nq, int cw, const int *rp, const int *ci, const double *val, double *B, hipStream_t st)
{
    if (nq <= 0) return;
    hipLaunchKernelGGL(k_qvec_dense, dim3((unsigned)((nq + 63) / 64)), dim3(64), 0, st, nq, cw, rp, ci, val, B);
    TMI_LAUNCH_CHECK();
}

void launch_select_cols(int elem_bytes, long rows, int ks, int kd, int nsel, const int *sel, const void *src,
                        void *dst, hipStream_t st)
{
    const long cnt = rows * kd;
    const dim3 g((unsigned)((cnt + 255) / 256));
    switch (elem_bytes) {
    case 16:
        hipLaunchKernelGGL(k_select_cols<double2>, g, dim3(256), 0, st, rows, ks, kd, nsel, sel, (const double2 *)src,
                           (double2 *)dst);
        break;
    case 8:
        hipLaunchKernelGGL(k_select_cols<double>, g, dim3(256), 0, st, rows, ks, kd, nsel, sel, (const double *)src,
                           (double *)dst);
        break;
    case 4:
        hipLaunchKernelGGL(k_select_cols<int>, g, dim3(256), 0, st, rows, ks, kd, nsel, sel, (const int *)src,
                           (int *)dst);
        break;
    default:
        TMI_CHECK(false, TMI_ERR_INVALID_ARG, "column selection of %d-byte entries", elem_bytes);
    }
    TMI_LAUNCH_CHECK();
}

void launch_scatter_map(int es, int n, int K, const int *perm, const int *cmap, const void *src, void *dst, long ld,
                        int q0, hipStream_t st)
{
    const size_t cnt = (size_t)n * K;
    const dim3 g((unsigned)((cnt + 255) / 256));
    if (es == 16)
        hipLaunchKernelGGL(k_scatter_map<double2>, g, dim3(256), 0, st, n, K, perm, cmap, (const double2 *)src,
                           (double2 *)dst, ld, q0);
    else
        hipLaunchKernelGGL(k_scatter_map<double>, g, dim3(256), 0, st, n, K, perm, cmap, (const double *)src,
                           (double *)dst, ld, q0);
    TMI_LAUNCH_CHECK();
}

}  // namespace tmi
