// Restarted GMRES on the GPU in FP64, complex or real: the algorithm of
// gmres() in gmres_imp.hpp:41-230 (Barrett et al., Templates, 1994), which
// IterativeSolve runs with restart 20 for ITMETHOD_GMRES
// (gsmatrix_imp.hpp:477-478, 538-539, 590-593; one right-hand side at a time).
//
// The columns of a tile of K right-hand sides (node-major [n][K], the
// solver's node order) each run their own GMRES, in lockstep: one launch
// serves the step of every column, and a column whose inner (Arnoldi) loop or
// whose restart loop has ended is skipped by the device flags below.  The
// per-column operations follow the reference's:
//   Mb = M^-1 b (left preconditioning), norm_b = ||Mb||;
//   v0 = Mb if ||x|| < tol N, else Mb - M^-1 A x;
//   while ||v0|| > tol norm_b and cycle < max_cycle (maxit, or N+1 if 0):
//     s0 = ||v0||, v0 /= s0, cycle++;
//     while |s_{j+1}| > tol norm_b, j < restart-1 and cycle < max_cycle:
//       j++; w = M^-1 A v_j;
//       modified Gram-Schmidt: H(j,k) = sum_n w_n conj(v_k,n), w -= H(j,k) v_k;
//       H(j,j+1) = ||w||, v_{j+1} = w / ||w||;
//       the rotations of the earlier steps on column j of H, then a new one
//       with r = sqrt(H(j,j)^2 + H(j,j+1)^2), c1 = H(j,j)/r, c2 = H(j,j+1)/r
//       (for complex H the reference's rotation does not conjugate:
//       c1^2 + c2^2 = 1), s_{j+1} = -c2 s_j, s_j = c1 s_j;
//     back substitution H y = s, x += y_l v_l for l = j..0;
//     v0 = Mb - M^-1 A x.
// Two guards the reference does not have end a column early: an exact
// breakdown (||w|| = 0 or r = 0) ends its cycle, and a column whose true
// residual has not halved over GM_STALL cycles stops (it is then reported
// unconverged, as the reference's loop would run on to max_cycle).
//
// Every reduction is a fixed-order two-stage sum (per-workgroup partials,
// then one thread per column), so a solve is reproducible run to run.
#include <cmath>
#include <vector>

#include "solver.hpp"

namespace tmi {

namespace {

constexpr int GM_RESTART = 20;   // the reference's restart length
constexpr int GM_STALL = 32;     // cycles without halving of the true residual before a column stops
constexpr int GM_BLOCK = 256;
constexpr int GM_MAXG = 1024;    // reduction workgroups at most

// per-column doubles [slot][K]
enum { D_NORMB = 0, D_NORMV, D_INV, D_REF, D_NSLOT };
// per-column ints [slot][K]
enum { G_OUT = 0, G_IN, G_J, G_CYC, G_STEPS, G_LASTIMP, G_XSMALL, G_NSLOT };

// ---- scalar arithmetic, complex (double2) and real (double) ----
__device__ __forceinline__ double2 mul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double mul(double a, double b) { return a * b; }
// a * conj(b)
__device__ __forceinline__ double2 mulc(double2 a, double2 b) { return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
__device__ __forceinline__ double mulc(double a, double b) { return a * b; }
__device__ __forceinline__ double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double add(double a, double b) { return a + b; }
__device__ __forceinline__ double2 sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double sub(double a, double b) { return a - b; }
__device__ __forceinline__ double2 neg(double2 a) { return make_double2(-a.x, -a.y); }
__device__ __forceinline__ double neg(double a) { return -a; }
__device__ __forceinline__ double2 scl(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double scl(double a, double s) { return a * s; }
__device__ __forceinline__ double abs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double abs2(double a) { return a * a; }
__device__ __forceinline__ double mag(double2 a) { return hypot(a.x, a.y); }
__device__ __forceinline__ double mag(double a) { return fabs(a); }
__device__ __forceinline__ bool is_zero(double2 a) { return a.x == 0.0 && a.y == 0.0; }
__device__ __forceinline__ bool is_zero(double a) { return a == 0.0; }
template <class T> __device__ __forceinline__ T from_real(double r);
template <> __device__ __forceinline__ double2 from_real<double2>(double r) { return make_double2(r, 0.0); }
template <> __device__ __forceinline__ double from_real<double>(double r) { return r; }
// a / b (Smith's method, Comm. ACM 5(8):435, 1962)
__device__ __forceinline__ double2 dvd(double2 a, double2 b)
{
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    }
    const double r = b.x / b.y, d = b.y + b.x * r;
    return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
__device__ __forceinline__ double dvd(double a, double b) { return a / b; }
// principal square root: re = sqrt((|z| + x)/2), im = y / (2 re) (for
// x >= 0; the mirrored form for x < 0 avoids the cancellation)
__device__ __forceinline__ double2 csqrt(double2 z)
{
    const double m = hypot(z.x, z.y);
    if (m == 0.0) return make_double2(0.0, 0.0);
    if (z.x >= 0.0) {
        const double t = sqrt(0.5 * (m + z.x));
        return make_double2(t, z.y / (2.0 * t));
    }
    const double t = sqrt(0.5 * (m - z.x));
    return make_double2(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}
__device__ __forceinline__ double csqrt(double z) { return sqrt(z); }

// fixed-order sum of the column's partials of every workgroup
template <class T>
__device__ __forceinline__ T col_sum(const T *part, int G, int K, int c)
{
    T s = from_real<T>(0.0);
    for (int g = 0; g < G; g++) s = add(s, part[(size_t)g * K + c]);
    return s;
}

// the workgroup's per-column sums of acc (one value per thread) into
// part[blockIdx.x][c], the rows of the workgroup added in a fixed order
template <class T>
__device__ __forceinline__ void block_partial(T acc, int K, T *part)
{
    __shared__ T red[GM_BLOCK];
    red[threadIdx.x] = acc;
    __syncthreads();
    if ((int)threadIdx.x < K) {
        T s = red[threadIdx.x];
        for (int r = 1; r < GM_BLOCK / K; r++) s = add(s, red[r * K + threadIdx.x]);
        part[(size_t)blockIdx.x * K + threadIdx.x] = s;
    }
}

// y = D^-1 A x for row i, column c (D^-1 = idiag, or none)
template <class T>
__device__ __forceinline__ T apply_row(int i, int c, int K, const int *rp, const int *ci, const T *val, const T *idiag,
                                       const T *X)
{
    T acc = from_real<T>(0.0);
    for (int k = rp[i]; k < rp[i + 1]; k++) acc = add(acc, mul(val[k], X[(size_t)ci[k] * K + c]));
    return idiag ? mul(idiag[i], acc) : acc;
}

// Mb = D^-1 b; partials [||Mb||^2, ||x||^2] (part: [G][2][K])
template <class T>
__global__ void __launch_bounds__(GM_BLOCK) k_gm_init(int n, int K, const T *B, const T *idiag, const T *X, T *Mb,
                                                      T *part)
{
    const int c = threadIdx.x % K, rpb = GM_BLOCK / K;
    T a0 = from_real<T>(0.0), a1 = from_real<T>(0.0);
    for (int i = blockIdx.x * rpb + threadIdx.x / K; i < n; i += gridDim.x * rpb) {
        const size_t e = (size_t)i * K + c;
        const T mb = idiag ? mul(idiag[i], B[e]) : B[e];
        Mb[e] = mb;
        a0 = add(a0, from_real<T>(abs2(mb)));
        a1 = add(a1, from_real<T>(abs2(X[e])));
    }
    __shared__ T red[2][GM_BLOCK];
    red[0][threadIdx.x] = a0;
    red[1][threadIdx.x] = a1;
    __syncthreads();
    if ((int)threadIdx.x < K) {
        for (int s = 0; s < 2; s++) {
            T t = red[s][threadIdx.x];
            for (int r = 1; r < rpb; r++) t = add(t, red[s][r * K + threadIdx.x]);
            part[((size_t)blockIdx.x * 2 + s) * K + threadIdx.x] = t;
        }
    }
}

template <class T>
__global__ void k_gm_fin_init(int G, int K, int n, double tol, const T *part, double *dsc, int *isc)
{
    const int c = threadIdx.x;
    if (c >= K) return;
    T sb = from_real<T>(0.0), sx = from_real<T>(0.0);
    for (int g = 0; g < G; g++) {
        sb = add(sb, part[((size_t)g * 2) * K + c]);
        sx = add(sx, part[((size_t)g * 2 + 1) * K + c]);
    }
    dsc[D_NORMB * K + c] = sqrt(mag(sb));
    dsc[D_REF * K + c] = INFINITY;
    isc[G_XSMALL * K + c] = sqrt(mag(sx)) < tol * n;
    isc[G_CYC * K + c] = 0;
    isc[G_STEPS * K + c] = 0;
    isc[G_LASTIMP * K + c] = 0;
    isc[G_IN * K + c] = 0;
    isc[G_J * K + c] = 0;
}

// v0 = Mb - D^-1 A x (v0 = Mb for a column whose initial x was taken as
// zero); partials ||v0||^2
template <class T>
__global__ void __launch_bounds__(GM_BLOCK) k_gm_resid(int n, int K, const int *rp, const int *ci, const T *val,
                                                       const T *idiag, const T *X, const T *Mb, T *V0, const int *isc,
                                                       T *part)
{
    const int c = threadIdx.x % K, rpb = GM_BLOCK / K;
    const bool on = isc[G_OUT * K + c] != 0;
    const bool xs = isc[G_XSMALL * K + c] != 0;
    T acc = from_real<T>(0.0);
    if (on) {
        for (int i = blockIdx.x * rpb + threadIdx.x / K; i < n; i += gridDim.x * rpb) {
            const size_t e = (size_t)i * K + c;
            const T v = xs ? Mb[e] : sub(Mb[e], apply_row(i, c, K, rp, ci, val, idiag, X));
            V0[e] = v;
            acc = add(acc, from_real<T>(abs2(v)));
        }
    }
    block_partial(acc, K, part);
}

// end of a restart cycle: ||v0|| against the tolerance, the cycle limit and
// the stagnation guard; a continuing column starts its next cycle
template <class T>
__global__ void k_gm_fin_outer(int G, int K, double tol, int max_cycle, const T *part, T *s, double *dsc, int *isc)
{
    const int c = threadIdx.x;
    if (c >= K || !isc[G_OUT * K + c]) return;
    const double nv = sqrt(mag(col_sum(part, G, K, c)));
    dsc[D_NORMV * K + c] = nv;
    isc[G_XSMALL * K + c] = 0;   // only the first residual treats x as zero
    const int cyc = isc[G_CYC * K + c];
    bool go = nv > tol * dsc[D_NORMB * K + c] && cyc < max_cycle;
    if (go) {
        if (nv < 0.5 * dsc[D_REF * K + c]) {
            dsc[D_REF * K + c] = nv;
            isc[G_LASTIMP * K + c] = cyc;
        } else if (cyc - isc[G_LASTIMP * K + c] >= GM_STALL) {
            go = false;
        }
    }
    if (!go) {
        isc[G_OUT * K + c] = 0;
        isc[G_IN * K + c] = 0;
        return;
    }
    isc[G_CYC * K + c] = cyc + 1;
    s[(size_t)c * (GM_RESTART + 1)] = from_real<T>(nv);
    dsc[D_INV * K + c] = 1.0 / nv;
    isc[G_IN * K + c] = cyc + 1 < max_cycle;
    isc[G_J * K + c] = 0;
}

// V *= inv[c] for the columns whose flag (slot fs) is set
template <class T>
__global__ void __launch_bounds__(GM_BLOCK) k_gm_scale(int n, int K, T *V, const double *dsc, const int *isc, int fs)
{
    const int c = threadIdx.x % K, rpb = GM_BLOCK / K;
    const int i = blockIdx.x * rpb + threadIdx.x / K;
    if (i >= n || !isc[fs * K + c]) return;
    const size_t e = (size_t)i * K + c;
    V[e] = scl(V[e], dsc[D_INV * K + c]);
}

// W = D^-1 A V for the columns in their Arnoldi loop
template <class T>
__global__ void __launch_bounds__(GM_BLOCK) k_gm_spmv(int n, int K, const int *rp, const int *ci, const T *val,
                                                      const T *idiag, const T *X, T *W, const int *isc)
{
    const int c = threadIdx.x % K, rpb = GM_BLOCK / K;
    const int i = blockIdx.x * rpb + threadIdx.x / K;
    if (i >= n || !isc[G_IN * K + c]) return;
    W[(size_t)i * K + c] = apply_row(i, c, K, rp, ci, val, idiag, X);
}

// one modified Gram-Schmidt sweep of step j: W -= H(j,k-1) v_{k-1} (k > 0),
// then the partials of <W, v_k> (Vk given) or of ||W||^2 (Vk null)
template <class T>
__global__ void __launch_bounds__(GM_BLOCK) k_gm_mgs(int n, int K, T *W, const T *Vp, const T *hp, const T *Vk,
                                                     const int *isc, T *part)
{
    const int c = threadIdx.x % K, rpb = GM_BLOCK / K;
    const bool on = isc[G_IN * K + c] != 0;
    T acc = from_real<T>(0.0);
    if (on) {
        const T h = Vp ? hp[(size_t)c * GM_RESTART * (GM_RESTART + 1)] : from_real<T>(0.0);
        for (int i = blockIdx.x * rpb + threadIdx.x / K; i < n; i += gridDim.x * rpb) {
            const size_t e = (size_t)i * K + c;
            T w = W[e];
            if (Vp) {
                w = sub(w, mul(h, Vp[e]));
                W[e] = w;
            }
            acc = add(acc, Vk ? mulc(w, Vk[e]) : from_real<T>(abs2(w)));
        }
    }
    block_partial(acc, K, part);
}

// H(j,k) = <w, v_k>
template <class T>
__global__ void k_gm_fin_h(int G, int K, const T *part, T *hp, const int *isc)
{
    const int c = threadIdx.x;
    if (c >= K || !isc[G_IN * K + c]) return;
    hp[(size_t)c * GM_RESTART * (GM_RESTART + 1)] = col_sum(part, G, K, c);
}

// end of Arnoldi step j: H(j,j+1) = ||w||, the rotations, the new rotation,
// s, and whether the column takes another step
template <class T>
__global__ void k_gm_givens(int G, int K, int j, double tol, const T *part, T *H, T *c1, T *c2, T *s, double *dsc,
                            int *isc)
{
    const int c = threadIdx.x;
    if (c >= K || !isc[G_IN * K + c]) return;
    const double nw = sqrt(mag(col_sum(part, G, K, c)));
    T *Hj = H + (size_t)c * GM_RESTART * (GM_RESTART + 1) + (size_t)j * (GM_RESTART + 1);
    T *cc1 = c1 + (size_t)c * GM_RESTART, *cc2 = c2 + (size_t)c * GM_RESTART;
    T *sc = s + (size_t)c * (GM_RESTART + 1);
    Hj[j + 1] = from_real<T>(nw);
    for (int k = 0; k < j; k++) {
        const T h1 = add(mul(cc1[k], Hj[k]), mul(cc2[k], Hj[k + 1]));
        const T h2 = add(mul(neg(cc2[k]), Hj[k]), mul(cc1[k], Hj[k + 1]));
        Hj[k] = h1;
        Hj[k + 1] = h2;
    }
    const T r = csqrt(add(mul(Hj[j], Hj[j]), mul(Hj[j + 1], Hj[j + 1])));
    isc[G_STEPS * K + c] += 1;
    if (is_zero(r)) {   // breakdown: the step does not enter the update
        isc[G_IN * K + c] = 0;
        return;
    }
    cc1[j] = dvd(Hj[j], r);
    cc2[j] = dvd(Hj[j + 1], r);
    Hj[j] = r;
    Hj[j + 1] = from_real<T>(0.0);
    sc[j + 1] = mul(neg(cc2[j]), sc[j]);
    sc[j] = mul(sc[j], cc1[j]);
    isc[G_J * K + c] = j + 1;
    dsc[D_INV * K + c] = nw > 0.0 ? 1.0 / nw : 0.0;
    isc[G_IN * K + c] = mag(sc[j + 1]) > tol * dsc[D_NORMB * K + c] && j < GM_RESTART - 1 && nw > 0.0;
}

// back substitution H y = s over the steps of the cycle
template <class T>
__global__ void k_gm_backsub(int K, T *H, T *s, T *y, const int *isc)
{
    const int c = threadIdx.x;
    if (c >= K || !isc[G_OUT * K + c]) return;
    const T *Hc = H + (size_t)c * GM_RESTART * (GM_RESTART + 1);
    T *sc = s + (size_t)c * (GM_RESTART + 1), *yc = y + (size_t)c * GM_RESTART;
    for (int l = isc[G_J * K + c] - 1; l >= 0; l--) {
        yc[l] = dvd(sc[l], Hc[l * (GM_RESTART + 1) + l]);
        for (int k = l - 1; k >= 0; k--) sc[k] = sub(sc[k], mul(Hc[l * (GM_RESTART + 1) + k], yc[l]));
    }
}

// x += y_l v_l, l = j..0
template <class T>
__global__ void __launch_bounds__(GM_BLOCK) k_gm_xupd(int n, int K, const T *V, const T *y, const int *isc, T *X)
{
    const int c = threadIdx.x % K, rpb = GM_BLOCK / K;
    const int i = blockIdx.x * rpb + threadIdx.x / K;
    if (i >= n || !isc[G_OUT * K + c]) return;
    const int jc = isc[G_J * K + c];
    if (jc == 0) return;
    const size_t e = (size_t)i * K + c, nk = (size_t)n * K;
    const T *yc = y + (size_t)c * GM_RESTART;
    T x = X[e];
    for (int l = jc - 1; l >= 0; l--) x = add(x, mul(yc[l], V[l * nk + e]));
    X[e] = x;
}

}  // namespace

// Workspace of one GMRES tile: the restart+1 basis vectors, Mb, partials
// and the per-column state.
struct GmresWork {
    int K = 0, n = 0, G = 0;
    size_t es = 0;
    DBuf<char> V, mb, H, s, c1, c2, y, part;
    DBuf<double> dsc;
    DBuf<int> isc;
    void ensure(int n_, int K_, size_t es_)
    {
        if (K == K_ && n == n_ && es == es_) return;
        K = K_, n = n_, es = es_;
        G = (int)std::min<size_t>(GM_MAXG, std::max<size_t>(1, ((size_t)n * K + 4 * GM_BLOCK - 1) / (4 * GM_BLOCK)));
        const size_t nk = (size_t)n * K;
        V.alloc(nk * (GM_RESTART + 1) * es);
        mb.alloc(nk * es);
        H.alloc((size_t)K * GM_RESTART * (GM_RESTART + 1) * es);
        s.alloc((size_t)K * (GM_RESTART + 1) * es);
        c1.alloc((size_t)K * GM_RESTART * es);
        c2.alloc((size_t)K * GM_RESTART * es);
        y.alloc((size_t)K * GM_RESTART * es);
        part.alloc((size_t)G * 2 * K * es);
        dsc.alloc((size_t)D_NSLOT * K);
        isc.alloc((size_t)G_NSLOT * K);
    }
};

namespace {

template <class T>
int gmres_tile(Solver &S, GmresWork &w, hipStream_t st, int K, int kc, const T *Bt, T *Xt, tmi_colstat *stats,
               int *h_isc, double *h_dsc)
{
    const Mesh::NodeOrder &o = S.ord();
    const int n = S.mesh->n, G = w.G, rpb = GM_BLOCK / K;
    const int *rp = o.rowptr.p, *ci = o.colidx.p;
    const T *val = (const T *)S.sval64.p;
    const T *idg = S.precon == TMI_PRECON_DIAG ? (const T *)S.idiag64.p : nullptr;
    const int max_cycle = S.maxit > 0 ? S.maxit : n + 1;
    const double tol = S.tol;
    const size_t nk = (size_t)n * K;
    T *V = (T *)w.V.p, *Mb = (T *)w.mb.p, *H = (T *)w.H.p, *s = (T *)w.s.p, *part = (T *)w.part.p;
    T *c1 = (T *)w.c1.p, *c2 = (T *)w.c2.p, *y = (T *)w.y.p;
    double *dsc = w.dsc.p;
    int *isc = w.isc.p;
    const dim3 blk(GM_BLOCK), gred(G), grow((unsigned)((n + rpb - 1) / rpb)), one(1), kth(64);

    std::vector<int> out0(K, 0);
    for (int c = 0; c < kc; c++) out0[c] = 1;
    TMI_HIP(hipMemcpyAsync(isc + G_OUT * K, out0.data(), K * sizeof(int), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL((k_gm_init<T>), gred, blk, 0, st, n, K, Bt, idg, (const T *)Xt, Mb, part);
    hipLaunchKernelGGL((k_gm_fin_init<T>), one, kth, 0, st, G, K, n, tol, (const T *)part, dsc, isc);
    auto residual = [&]() {
        hipLaunchKernelGGL((k_gm_resid<T>), gred, blk, 0, st, n, K, rp, ci, val, idg, (const T *)Xt, (const T *)Mb,
                           V, (const int *)isc, part);
        hipLaunchKernelGGL((k_gm_fin_outer<T>), one, kth, 0, st, G, K, tol, max_cycle, (const T *)part, s, dsc,
                           isc);
    };
    auto count = [&](int slot) {
        TMI_HIP(hipMemcpyAsync(h_isc, isc + slot * K, K * sizeof(int), hipMemcpyDeviceToHost, st));
        TMI_HIP(hipStreamSynchronize(st));
        int a = 0;
        for (int c = 0; c < K; c++) a += h_isc[c] != 0;
        return a;
    };
    residual();
    TMI_LAUNCH_CHECK();
    while (count(G_OUT) > 0) {
        hipLaunchKernelGGL((k_gm_scale<T>), grow, blk, 0, st, n, K, V, (const double *)dsc, (const int *)isc, (int)G_IN);
        for (int j = 0; j < GM_RESTART; j++) {
            T *Vj = V + (size_t)j * nk, *W = V + (size_t)(j + 1) * nk;
            T *Hj = H + (size_t)j * (GM_RESTART + 1);
            hipLaunchKernelGGL((k_gm_spmv<T>), grow, blk, 0, st, n, K, rp, ci, val, idg, (const T *)Vj, W,
                               (const int *)isc);
            for (int k = 0; k <= j + 1; k++) {
                const T *Vp = k > 0 ? V + (size_t)(k - 1) * nk : nullptr;
                const T *Vk = k <= j ? V + (size_t)k * nk : nullptr;
                hipLaunchKernelGGL((k_gm_mgs<T>), gred, blk, 0, st, n, K, W, Vp, (const T *)(Hj + (k > 0 ? k - 1 : 0)),
                                   Vk, (const int *)isc, part);
                if (k <= j)
                    hipLaunchKernelGGL((k_gm_fin_h<T>), one, kth, 0, st, G, K, (const T *)part, Hj + k,
                                       (const int *)isc);
            }
            hipLaunchKernelGGL((k_gm_givens<T>), one, kth, 0, st, G, K, j, tol, (const T *)part, H, c1, c2, s, dsc,
                               isc);
            TMI_LAUNCH_CHECK();
            // the columns that go on need v_{j+1} normalised
            if (j + 1 < GM_RESTART) {
                hipLaunchKernelGGL((k_gm_scale<T>), grow, blk, 0, st, n, K, W, (const double *)dsc,
                                   (const int *)isc, (int)G_IN);
                if ((j & 3) == 3 && count(G_IN) == 0) break;
            }
        }
        hipLaunchKernelGGL((k_gm_backsub<T>), one, kth, 0, st, K, H, s, y, (const int *)isc);
        hipLaunchKernelGGL((k_gm_xupd<T>), grow, blk, 0, st, n, K, (const T *)V, (const T *)y, (const int *)isc, Xt);
        residual();
        TMI_LAUNCH_CHECK();
    }
    TMI_HIP(hipMemcpyAsync(h_isc, isc, sizeof(int) * G_NSLOT * K, hipMemcpyDeviceToHost, st));
    TMI_HIP(hipMemcpyAsync(h_dsc, dsc, sizeof(double) * D_NSLOT * K, hipMemcpyDeviceToHost, st));
    TMI_HIP(hipStreamSynchronize(st));
    int notconv = 0;
    for (int c = 0; c < kc; c++) {
        const double nb = h_dsc[D_NORMB * K + c], nv = h_dsc[D_NORMV * K + c];
        const bool conv = !(nv > tol * nb);
        notconv += !conv;
        if (stats) {
            stats[c].iters = h_isc[G_CYC * K + c];
            stats[c].sweeps = h_isc[G_STEPS * K + c];
            stats[c].relres = nb > 0.0 ? nv / nb : 0.0;
            stats[c].converged = conv ? 1 : 0;
        }
    }
    return notconv;
}

}  // namespace

int Solver::gmres_on(Exec &e, int nq, const void *d_b, void *d_x, int ld, tmi_colstat *stats)
{
    if (precon == TMI_PRECON_DIAG) {
        std::lock_guard<std::mutex> lk(prep_mu);
        if (!precon_valid) {
            reset_precon();
            TMI_HIP(hipStreamSynchronize(st));
        }
    }
    if (!e.st) e.st = st;
    hipStream_t es_ = e.st;
    if (!e.gm) e.gm = new GmresWork;
    GmresWork &w = *e.gm;
    const int n = mesh->n, K = tile_for(nq, TMI_MODE_FP64_BLOCK);
    const size_t es = vbytes(vk64());
    w.ensure(n, K, es);
    DBuf<char> bt((size_t)n * K * es), xt((size_t)n * K * es);
    std::vector<int> h_isc((size_t)G_NSLOT * K);
    std::vector<double> h_dsc((size_t)D_NSLOT * K);
    const int *perm = ord().perm.p;
    hipEvent_t e0, e1;
    TMI_HIP(hipEventCreate(&e0));
    TMI_HIP(hipEventCreate(&e1));
    TMI_HIP(hipEventRecord(e0, es_));
    int notconv = 0;
    double iters = 0.0;
    for (int q0 = 0; q0 < nq; q0 += K) {
        const int kc = std::min(K, nq - q0);
        launch_gather_tile((int)es, n, K, kc, perm, d_b, ld, q0, bt.p, es_);
        launch_gather_tile((int)es, n, K, kc, perm, d_x, ld, q0, xt.p, es_);
        tmi_colstat *sq = stats ? stats + q0 : nullptr;
        if (cplx)
            notconv += gmres_tile<double2>(*this, w, es_, K, kc, (const double2 *)bt.p, (double2 *)xt.p, sq,
                                           h_isc.data(), h_dsc.data());
        else
            notconv += gmres_tile<double>(*this, w, es_, K, kc, (const double *)bt.p, (double *)xt.p, sq,
                                          h_isc.data(), h_dsc.data());
        launch_scatter_tile((int)es, n, K, kc, perm, xt.p, d_x, ld, q0, es_);
        for (int c = 0; c < kc; c++) iters += h_isc[G_STEPS * K + c];
    }
    TMI_HIP(hipEventRecord(e1, es_));
    TMI_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    TMI_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    e.last_solve_ms = ms;
    e.last_iters = nq ? iters / nq : 0.0;
    return notconv;
}

void free_gmres_work(GmresWork *w) { delete w; }

size_t gmres_work_bytes(const GmresWork *w)
{
    if (!w) return 0;
    return w->V.bytes() + w->mb.bytes() + w->H.bytes() + w->s.bytes() + w->c1.bytes() + w->c2.bytes() +
           w->y.bytes() + w->part.bytes() + w->dsc.bytes() + w->isc.bytes();
}

size_t gmres_tile_bytes(size_t n, size_t K, size_t es)
{
    // basis (restart + 1 vectors), Mb, the gathered right-hand sides and
    // solutions, and the reductions
    return n * K * es * (GM_RESTART + 4) + (size_t)GM_MAXG * 2 * K * es + K * (GM_RESTART + 2) * (GM_RESTART + 1) * es;
}

}  // namespace tmi
