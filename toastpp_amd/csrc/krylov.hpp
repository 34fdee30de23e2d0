// Block Krylov kernels: device-side declarations shared by the solver driver.
//
// Vectors of a block of K right-hand sides are stored node-major: element
// (row i, column c) at i*K + c, so one CSR row of the SpMM reads K
// consecutive values of each neighbour row (one coalesced 64-lane load for
// K = 64 complex-FP32 columns: 512 bytes).
//
// Per-column scalars live in device memory (struct of arrays below); the
// host only reads the active-column count every few iterations.
#pragma once
#include "common.hpp"

namespace tmi {

// offsets into the per-column double scalar array (each block of K doubles)
// (the _I slots hold the imaginary parts of COCG's complex scalars)
enum ScalarSlot {
    S_RHO = 0, S_ALPHA, S_OMEGA, S_BETA, S_BNORM2, S_RES2, S_TOL2, S_SCALE, S_AUX, S_RHOP, S_BEST,
    S_RHO_I, S_ALPHA_I, S_BETA_I, S_RHOP_I,
    S_FREQ,   // outer: the column's omega in a multi-frequency solve
    S_NSLOT
};
// offsets into the per-column int array
enum IntSlot { I_ACTIVE = 0, I_ITERS, I_FIRST, I_BREAK, I_STALL, I_XPEND, I_NSLOT };

// partial-sum slots written by the reduction epilogues (COCG's update
// kernel writes three: the complex <r, D^-1 r> and ||r||^2)
constexpr int RED_SLOTS = 3;

// Row blocks of the LDS-staged 64-column SpMM: consecutive rows whose
// neighbour rows (the X rows their nonzeros read) plus the rows themselves
// (their W rows) number at most LDS_ROWS, so that one workgroup of
// LDS_WAVES waves stages them in LDS once and reads them from there.
// LDS_NZ_C32 bounds a block's nonzeros (its values and positions are staged
// too) and a block holds fewer than LDS_ROWS_C32 rows.
constexpr int LDS_ROWS_C32 = 136;   // 136 x 64 x 8 B = 68 KiB (+ 6 KiB metadata): two workgroups per CU
constexpr int LDS_NZ_C32 = 576;
constexpr int LDS_WAVES = 4;
struct RowBlocks {
    int nb = 0;                    // number of blocks (0: not available)
    const int *brow = nullptr;     // nb+1 row bounds
    const int *bu = nullptr;       // nb+1 bounds into ucols
    const int *ucols = nullptr;    // neighbour rows of each block
    const unsigned short *lcol = nullptr;   // per nonzero: position of its column in the block's ucols
    bool dma = false;              // double-buffered LDS-DMA kernel instead of the register-staged one
};
// workgroups of the LDS-DMA SpMM: one per CU (each holds two 74 KiB buffers),
// at most the reduction grid (a multiple of 8)
constexpr int DMA_GRID = 256;

// rows per block of the row-block SpMM (one bit each in the column masks)
#ifndef TMI_RB_ROWS
#define TMI_RB_ROWS 8
#endif
constexpr int RB_ROWS = TMI_RB_ROWS;
static_assert(RB_ROWS <= 8 * (int)sizeof(rb_mask_t), "row-block masks hold one bit per row");
// neighbour rows gathered per batch in the row-block SpMM (MI355X, 250k-node
// tetrahedral mesh: 8 -> 0.165 ms per launch, 16 -> 0.182, 32 -> 0.23, 4 -> 0.175;
// 4, 6 or 8 rows per block within 1 % of each other at 8)
#ifndef TMI_RB_BATCH
#define TMI_RB_BATCH 8
#endif
struct RowUnions {
    int nb = 0;                          // blocks (0: not used)
    const int *ofs = nullptr;            // nb+1 bounds into col / mask
    const int *col = nullptr;            // distinct columns of each block, ascending
    const rb_mask_t *mask = nullptr;     // rows of the block holding each column
    const void *val = nullptr;           // coefficients in (column, row) order per block
    const void *src = nullptr;           // the operator (CSR values) they were taken from
    const unsigned *word = nullptr;      // packed mask | offset words: the pipelined kernel (variant 4)
    bool mfma = false;                   // matrix-core kernel (variant 5; needs word)
    const void *val64 = nullptr;         // complex FP64 coefficients in the same order (the FP64 residual)
    const void *src64 = nullptr;         // the FP64 operator they were taken from
};

struct BlockDesc {
    int n = 0;          // rows
    int K = 0;          // columns (tile width)
    int nblk = 0;       // reduction blocks of the row-parallel kernels
    const int *rowptr = nullptr, *colidx = nullptr;
    int rpw = 0;        // rows per run of the 64-column SpMM (0: default)
    RowBlocks lds;      // LDS-staged SpMM partition (complex FP32, 64 columns)
    RowUnions rb;       // row-block SpMM (complex FP32, 64 columns)
};

// precision / value-type selector
enum VKind { VK_C32 = 0, VK_C64 = 1, VK_R32 = 2, VK_R64 = 3 };

inline size_t vbytes(VKind k) { return k == VK_C64 ? 16 : (k == VK_R32 ? 4 : 8); }

// ---- launchers (krylov_kernels.hip) ----
// Y = A * D^-1 X (D^-1 = idiag, or identity if null); partials
// red[0] = <W, Y>_R and, when want_yy = 1, red[1] = <Y, Y>_R, per column;
// want_yy = DOT_BIL (complex vectors): red[0], red[1] = Re, Im of the
// unconjugated W^T Y (COCG's p^T A p).
// Wq (optional): W given as the packed shadow residual of a 64-column
// complex FP32 tile (launch_pack_shadow); used by the row-block kernel,
// the other kernels read W, which must hold the same values.
constexpr int DOT_BIL = 2;
void launch_spmm_dot(VKind vk, const BlockDesc &b, const void *val, const void *idiag, const void *X, void *Y,
                     const void *W, int want_yy, const int *iflags, double *partial, hipStream_t st,
                     const unsigned char *Wq = nullptr, const double *omg = nullptr);
// R = B - A X (FP64), red[0] = ||R||^2
// omg (complex vectors, K >= 2): per-column frequencies -- val then holds
// the operator with its imaginary part at omega = 1, and column c applies
// re + i omg[c] im (a multi-frequency block solve of one mesh)
void launch_resid(VKind vk, const BlockDesc &b, const void *val, const void *X, const void *B, void *R,
                  const int *iflags, double *partial, hipStream_t st, const double *omg = nullptr);

// BiCGSTAB element-wise steps (real inner products, real alpha/beta/omega)
void launch_bicg_p(VKind vk, const BlockDesc &b, const void *r, void *p, const void *v, const double *sc,
                   const int *ic, hipStream_t st);
void launch_bicg_s(VKind vk, const BlockDesc &b, const void *r, const void *v, void *s, const double *sc,
                   const int *ic, hipStream_t st);
void launch_bicg_xr(VKind vk, const BlockDesc &b, const void *idiag, void *x, const void *p, const void *s,
                    const void *t, void *r, const void *rd, const double *sc, const int *ic, double *partial,
                    hipStream_t st, const unsigned char *rdq = nullptr);
// packed copy (4 bits per entry) of the sign-pattern shadow residual rd of
// a 64-column complex FP32 tile: q is n x 32 bytes
void launch_pack_shadow(int n, const void *rd, unsigned char *q, hipStream_t st);
// PCG element-wise steps.  bil (complex vectors only): COCG -- complex
// alpha / beta, and the partials are the unconjugated r^T D^-1 r (re, im)
// and ||r||^2 in three slots
// omg (COCG only): per-column frequencies, idiag then holds the diagonal
// itself with its imaginary part at omega = 1
// xdefer (COCG only): leave x += alpha d to the direction update, which
// then takes x (launch_cg_d with x given) and applies it for the columns
// whose alpha step ran this iteration (I_XPEND, set by FIN_COCG_ALPHA):
// x and d are then streamed once per iteration instead of twice
void launch_cg_xr(VKind vk, const BlockDesc &b, const void *idiag, void *x, const void *d, const void *q, void *r,
                  const double *sc, const int *ic, double *partial, hipStream_t st, bool bil = false,
                  const double *omg = nullptr, bool xdefer = false);
void launch_cg_d(VKind vk, const BlockDesc &b, const void *idiag, void *d, const void *r, const double *sc,
                 const int *ic, hipStream_t st, bool bil = false, const double *omg = nullptr, void *x = nullptr);
// first PCG step: d = D^-1 r, partials [<r, D^-1 r>, <r, r>]
void launch_cg_init(VKind vk, const BlockDesc &b, const void *idiag, void *d, const void *r, const double *sc,
                    const int *ic, double *partial, hipStream_t st, bool bil = false, const double *omg = nullptr);

// scalar updates (one thread per column)
enum FinKind {
    FIN_BICG_INIT = 0,  // red: [<r,r>, <rd,r>] -> bnorm2, rho, res2
    FIN_BICG_ALPHA,     // red: [<rd,v>] -> alpha
    FIN_BICG_OMEGA,     // red: [<t,s>, <t,t>] -> omega
    FIN_BICG_RHO,       // red: [||r||^2, <rd,r>] -> res2, beta, rho, convergence
    FIN_CG_INIT,        // red: [<r,z>, <r,r>]
    FIN_CG_ALPHA,       // red: [<d,q>]
    FIN_CG_BETA,        // red: [<r,z>, ||r||^2]
    FIN_OUTER_BNORM,    // red: [<b,b>] -> outer bnorm2, scale = 1/||b||
    FIN_OUTER_RES,      // red: [||b - A x||^2] -> outer res2, converged / stalled columns leave
    FIN_BICG_RELIABLE,  // red: [<r,r>, <rd,r>] of the replaced residual -> res2, rho, beta
    FIN_CG_RELIABLE,    // red: [<r,r>, <r,D^-1 r>] of the replaced residual -> res2, rho
    // COCG (complex symmetric systems: PCG with the unconjugated bilinear
    // form r^T z and complex alpha, beta, rho); red holds complex values as
    // (re, im) in slots 0-1 and ||r||^2 in slot 2
    FIN_COCG_INIT,      // red: [r^T z, ||r||^2] -> rho, res2
    FIN_COCG_ALPHA,     // red: [p^T A p] -> alpha = rho / p^T A p
    FIN_COCG_BETA,      // red: [r^T z, ||r||^2] -> rho, beta, res2, convergence
    FIN_COCG_RELIABLE,  // red: [r^T z, ||r||^2] of the replaced residual -> rho, beta, res2
};
// maxit: iteration cap of the Krylov loop (0 = none); tol2: squared outer
// relative tolerance; inner_tol2: squared inner tolerance of the mixed mode
// (<= 0: the inner solve reproduces the outer criterion, FP64 modes)
// osc/oic: the outer (refinement) state, read by the *_RELIABLE kinds
void launch_finalize(FinKind fk, const BlockDesc &b, const double *partial, double *sc, int *ic, int maxit,
                     double tol2, double inner_tol2, hipStream_t st, const double *osc = nullptr,
                     const int *oic = nullptr);
// copy the outer active flags into the inner state and reset its counters
void launch_inner_reset(const BlockDesc &b, const double *osc, const int *oic, double *sc, int *ic, hipStream_t st);

// mixed-precision refinement helpers
// r32 = r64 * scale[c]; rd32 = r32; x32 = 0 ; partial [<r32,r32>, <rd32,r32>]
void launch_to_inner(VKind vk32, const BlockDesc &b, const void *r64, void *r32, void *rd32, void *x32,
                     const double *scale, const int *oflags, double *partial, hipStream_t st,
                     bool sign_shadow = false);
// reliable update of the inner residual: r_in = r64 * scale, x_in = 0,
// partials [<r_in, r_in>, <rd, r_in>] (rd given) or [<r_in,r_in>, <r_in, D^-1 r_in>]
void launch_reliable(VKind vk_in, const BlockDesc &b, const void *r64, void *rin, void *xin, const void *rd,
                     const void *idiag, const double *scale, const int *oflags, double *partial, hipStream_t st,
                     bool bil = false, const double *omg = nullptr);
// x64 += x32 / scale[c] for outer-active columns
void launch_from_inner(VKind vk32, const BlockDesc &b, void *x64, const void *x32, const double *scale,
                       const int *oflags, hipStream_t st);
// per-column squared norm of X: partial [<X,X>, <X,X>]
void launch_norm2(VKind vk, const BlockDesc &b, const void *X, const int *flags, double *partial, hipStream_t st);
// inverse diagonal (precon_imp.hpp:56-65): idiag[i] = 1/K[i,i] (1 where 0)
void launch_inv_diag(VKind vk, int n, const int *diagpos, const double *val64, void *idiag, hipStream_t st);
// the complex diagonal itself (vk complex)
void launch_get_diag(VKind vk, int n, const int *diagpos, const double *val64, void *dg, hipStream_t st);
// A D^-1: out64[k] = val64[k] * idiag[colidx[k]] (and an FP32 copy in out32)
void launch_scale_cols(bool cplx, int nnz, const int *colidx, const double *val64, const void *idiag64, double *out64,
                       void *out32, hipStream_t st);
// copy with conversion FP64 -> FP32 (values)
void launch_f64_to_f32(const double *src, float *dst, size_t count, hipStream_t st);
// matrix values into the solver's node order: out64[k] = val64[slot[k]]
// (complex: both parts), out32 (optional) the FP32 copy
void launch_permute_vals(bool cplx, int nnz, const int *slot, const double *val64, double *out64, float *out32,
                         hipStream_t st);
// FP32 values into the row-block SpMM's order: out[k] = in[slot[k]] (w floats each)
void launch_gather_vals32(int nnz, int w, const int *slot, const float *in, float *out, hipStream_t st);
void launch_gather_vals64(int nnz, int w, const int *slot, const double *in, double *out, hipStream_t st);
// tile of right-hand sides / solutions between the caller's [n][ld] layout
// (reference node order) and the solver's [n][K] tile (solver node order):
// gather dst[r][c] = src[perm[r]][q0 + c] for c < kc, 0 for kc <= c < K;
// scatter dst[perm[r]][q0 + c] = src[r][c] for c < kc.  es: 16 (complex) or 8.
void launch_gather_tile(int es, int n, int K, int kc, const int *perm, const void *src, long ld, int q0, void *dst,
                        hipStream_t st);
void launch_scatter_tile(int es, int n, int K, int kc, const int *perm, const void *src, void *dst, long ld, int q0,
                         hipStream_t st);
// column compaction of a [rows][ks] array into [rows][kd]: column j of dst
// is column sel[j] of src for j < nsel, zero beyond (elem_bytes 4, 8, 16)
void launch_select_cols(int elem_bytes, long rows, int ks, int kd, int nsel, const int *sel, const void *src,
                        void *dst, hipStream_t st);
// out[c][r] = in[r][c] for a rows x cols array of elem_bytes (8 or 16)
// entries (tiled through LDS)
void launch_transpose(int elem_bytes, long rows, int cols, const void *in, void *out, hipStream_t st);
// dense node-major right-hand sides from a CSR qvec (row q = source q):
// B[j][q] += val[k] for the entries k of row q, in order (B zeroed first);
// cw = 2 for complex values
void launch_qvec_dense(int nq, int cw, const int *rp, const int *ci, const double *val, double *B, hipStream_t st);
// scatter with a column map: dst[perm[r]][q0 + cmap[j]] = src[r][j] where cmap[j] >= 0
void launch_scatter_map(int es, int n, int K, const int *perm, const int *cmap, const void *src, void *dst, long ld,
                        int q0, hipStream_t st);

}  // namespace tmi
