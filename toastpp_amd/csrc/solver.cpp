// Forward-solver driver: Allocate / Assemble / Reset / CalcFields on the GPU.
//
// CalcFields runs every mode on tiles of K right-hand-side columns:
//   FP64:   r = q - K x (FP64 CSR SpMM) and the per-column convergence test
//           ||r|| <= tol ||q|| (the reference criterion, matrix_imp.hpp:350-413)
//   Krylov: block BiCGSTAB or COCG (complex) or PCG (real) -- complex/real
//           FP32 on the system scaled by 1/||q|| in TMI_MODE_MIXED_BLOCK, FP64
//           in the FP64 modes; COCG columns may each carry their own
//           frequency (the operator then assembled at omega = 1 and scaled per
//           column).  The mixed mode makes reliable updates: every few
//           iterations, and whenever every column's recursive residual meets
//           the tolerance, the FP32 correction is folded into the FP64
//           solution and the FP64 true residual replaces the FP32 recursive
//           one (search directions kept).  The FP64 modes stop on the
//           recursive residual, as the reference does.
// Columns leave the iteration individually (device-side active flags); the
// host reads the flags every few iterations only.
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace tmi {

Solver::Exec::Exec(bool new_stream)
{
    if (new_stream) {
        TMI_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        own_stream = true;
    }
}

void Solver::Exec::drop_graphs()
{
    for (Graph &g : graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
}

Solver::Exec::~Exec()
{
    if (st) (void)hipStreamSynchronize(st);
    drop_graphs();
    if (h_flags) (void)hipHostFree(h_flags);
    if (h_dbl) (void)hipHostFree(h_dbl);
    free_gmres_work(gm);
    if (own_stream) (void)hipStreamDestroy(st);
}

size_t Solver::Exec::bytes() const
{
    size_t b = 0;
    for (const Work *ws : {wk, wk_alt})
        for (int i = 0; i < 7; i++) {
            const Work &w = ws[i];
            b += w.x64.bytes() + w.b64.bytes() + w.r64.bytes() + w.xi.bytes() + w.ri.bytes() + w.rdi.bytes() +
                 w.p.bytes() + w.v.bytes() + w.s.bytes() + w.t.bytes() + w.partial.bytes() + w.sc.bytes() +
                 w.osc.bytes() + w.ic.bytes() + w.oic.bytes() + w.rdq.bytes();
        }
    return b + gmres_work_bytes(gm);
}

static int log2i(int K)
{
    int l = 0;
    while ((1 << l) < K) l++;
    return l;
}

Solver::~Solver() {}

int Solver::tile_for(int nq, int md) const
{
    if (md < 0) md = mode;
    if (md == TMI_MODE_FP64_SINGLE) return 1;
    int K = tile > 0 ? tile : 64;
    while (K > 1 && K / 2 >= nq) K /= 2;   // smallest power of two >= nq, at most the tile
    return K;
}

size_t Solver::work_bytes(int nq) const
{
    if (lsolver == TMI_LSOLVER_GMRES) return gmres_tile_bytes(mesh->n, tile_for(nq, TMI_MODE_FP64_BLOCK), vbytes(vk64()));
    // the tile and, with compaction, its narrower copies (K/2 + K/4 + ... < K)
    // and the second workspaces of the repacked widths (K and K/2)
    const size_t K = tile_for(nq), n = mesh->n, Kall = compact_min > 0 ? 2 * K + K + K / 2 : K;
    const bool mixed = mode == TMI_MODE_MIXED_BLOCK;
    const size_t s64 = vbytes(vk64()), sin = vbytes(mixed ? vk32() : vk64());
    auto tiles = [&](size_t s_inner) {
        return n * Kall * (3 * s64 + 7 * s_inner) + RED_SLOTS * 2048 * Kall * 8 + (S_NSLOT * 2 * 8 + I_NSLOT * 2 * 4) * Kall;
    };
    size_t b = tiles(sin);
    if (mixed && K == 64 && cplx) b += 2 * n * 32;   // packed shadow residuals (both 64-column workspaces)
    // the FP64 fallback of the mixed mode re-solves unconverged columns in
    // FP64: the workspaces it uses are reallocated with FP64 inner vectors,
    // and their right-hand sides and solutions are copied (at most nq each)
    if (mixed && fp64_fallback) b = std::max(b, tiles(s64)) + 2 * n * (size_t)nq * s64;
    return b;
}

void Solver::Exec::release_work()
{
    if (st) TMI_HIP(hipStreamSynchronize(st));
    drop_graphs();   // their kernel arguments point into the workspaces
    for (Work *ws : {wk, wk_alt})
        for (int i = 0; i < 7; i++) {
            Work &w = ws[i];
            for (DBuf<char> *b : {&w.x64, &w.b64, &w.r64, &w.xi, &w.ri, &w.rdi, &w.p, &w.v, &w.s, &w.t}) b->release();
            w.rdq.release();
            w.partial.release();
            w.sc.release();
            w.osc.release();
            w.ic.release();
            w.oic.release();
            w.K = w.n = w.G = 0;
        }
    free_gmres_work(gm);
    gm = nullptr;
}

void Solver::allocate()
{
    mesh->build_csr(st);
    mesh->build_order(order_kind, st);
    val64.alloc(nval());
    sval64.alloc(nval());
    sval32.alloc(nval() + 8);   // + 16-byte-chunk slack for the LDS-DMA SpMM's value staging
    allocated = true;
    assembled = false;
}

BlockDesc Solver::block(const Exec::Work &w, int K) const
{
    const Mesh::NodeOrder &o = ord();
    BlockDesc b{mesh->n, K, w.G, o.rowptr.p, o.colidx.p};
    b.rpw = rpw;
    if ((spmm_variant == 0 || spmm_variant == 2) && o.nb > 0) {
        b.lds = RowBlocks{o.nb, o.brow.p, o.bu.p, o.ucols.p, o.lcol.p};
        b.lds.dma = spmm_variant == 2;
    }
    if (spmm_variant >= 3 && o.rb_nb > 0 && rb_src) {
        b.rb = RowUnions{o.rb_nb, o.rb_ofs.p, o.rb_col.p, o.rb_mask.p, rbval32.p, rb_src};
        if (spmm_variant >= 4 && o.rb_word_ok) b.rb.word = o.rb_word.p;
        b.rb.mfma = spmm_variant == 5 && o.rb_word_ok;
        if (rb64_src) {
            b.rb.val64 = rbval64.p;
            b.rb.src64 = rb64_src;
        }
    }
    return b;
}

void Solver::assemble(const double *cmua, const double *ckap, const double *c2a, double om, bool host_ptrs,
                      bool elbasis)
{
    TMI_CHECK(allocated, TMI_ERR_STATE, "AssembleSystemMatrix called before Allocate");
    TMI_CHECK(cplx || om == 0.0, TMI_ERR_INVALID_ARG, "nonzero omega on a real (CW) solver");
    // nodal coefficients, or one per element in the element basis
    int n = elbasis ? mesh->elen : mesh->n;
    coef.alloc(3 * (size_t)n);
    if (host_ptrs) {
        TMI_HIP(hipMemcpyAsync(coef.p, cmua, n * sizeof(double), hipMemcpyHostToDevice, st));
        TMI_HIP(hipMemcpyAsync(coef.p + n, ckap, n * sizeof(double), hipMemcpyHostToDevice, st));
        TMI_HIP(hipMemcpyAsync(coef.p + 2 * n, c2a, n * sizeof(double), hipMemcpyHostToDevice, st));
    } else {
        TMI_HIP(hipMemcpyAsync(coef.p, cmua, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        TMI_HIP(hipMemcpyAsync(coef.p + n, ckap, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        TMI_HIP(hipMemcpyAsync(coef.p + 2 * n, c2a, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    omega = om;
    this->elbasis = elbasis;
    mf_valid = false;       // so does the multi-frequency operator
    precon_valid = false;   // the preconditioner and K D^-1 follow the matrix
    rb_src = nullptr;       // ... and so do the row-block coefficients
    rb64_src = nullptr;
    launch_assemble(*mesh, coef.p, coef.p + n, coef.p + 2 * n, om, cplx ? 1 : 0, val64.p, nullptr, st,
                    elbasis ? 1 : 0);
    // the solver's copy in its own node order
    launch_permute_vals(cplx, mesh->nnz, ord().slot.p, val64.p, sval64.p, sval32.p, st);
    assembled = true;
}

void Solver::reset_precon()
{
    TMI_CHECK(assembled, TMI_ERR_STATE, "preconditioner reset before assembly");
    int n = mesh->n;
    idiag64.alloc((size_t)n * vbytes(vk64()));
    idiag32.alloc((size_t)n * vbytes(vk32()));
    launch_inv_diag(vk64(), n, ord().diagpos.p, sval64.p, idiag64.p, st);
    launch_inv_diag(vk32(), n, ord().diagpos.p, sval64.p, idiag32.p, st);
    // BiCGSTAB applies M^-1 on the right (pd = M^-1 p, v = K pd): keep K D^-1
    // as its own operator so the SpMM needs no per-nonzero diagonal gather
    pval64.alloc(nval());
    pval32.alloc(nval() + 8);
    launch_scale_cols(cplx, mesh->nnz, ord().colidx.p, sval64.p, idiag64.p, pval64.p, pval32.p, st);
    precon_valid = true;
}

void Solver::prepare_mf()
{
    TMI_CHECK(assembled && cplx, TMI_ERR_STATE, "multi-frequency solve before assembly of a complex system");
    if (mf_valid) return;
    const int nc = elbasis ? mesh->elen : mesh->n, n = mesh->n;
    // the system at omega = 1: real part A, imaginary part the mass term
    DBuf<double> v64((size_t)nval());
    launch_assemble(*mesh, coef.p, coef.p + nc, coef.p + 2 * (size_t)nc, 1.0, 1, v64.p, nullptr, st,
                    elbasis ? 1 : 0);
    mfval64.alloc(nval());
    mfval32.alloc(nval() + 8);
    launch_permute_vals(true, mesh->nnz, ord().slot.p, v64.p, mfval64.p, mfval32.p, st);
    mfdiag64.alloc((size_t)n * vbytes(VK_C64));
    mfdiag32.alloc((size_t)n * vbytes(VK_C32));
    launch_get_diag(VK_C64, n, ord().diagpos.p, mfval64.p, mfdiag64.p, st);
    launch_get_diag(VK_C32, n, ord().diagpos.p, mfval64.p, mfdiag32.p, st);
    TMI_HIP(hipStreamSynchronize(st));   // v64 goes out of scope; other streams read the result next
    mf_valid = true;
}

// workgroups of a tile's block kernels and reduction partials, at most
// (2048 measured best for the 250k-node, 64-column solve)
#ifndef TMI_WORK_G
#define TMI_WORK_G 2048
#endif

Solver::Exec::Work &Solver::ensure_work(Exec &e, int K, int md, bool alt)
{
    auto &w = alt ? e.wk_alt[log2i(K)] : e.wk[log2i(K)];
    int n = mesh->n;
    VKind vin = (md == TMI_MODE_MIXED_BLOCK) ? vk32() : vk64();
    if (!e.h_flags) TMI_HIP(hipHostMalloc((void **)&e.h_flags, sizeof(int) * I_NSLOT * 64 * 2));
    if (!e.h_dbl) TMI_HIP(hipHostMalloc((void **)&e.h_dbl, sizeof(double) * S_NSLOT * 64));
    if (!e.cols.p) e.cols.alloc(128);
    if (w.K == K && w.n == n && w.vk_in == vin) return w;
    w.K = K;
    w.n = n;
    w.vk_in = vin;
    // reduction workgroups: a multiple of 8 (one slice per XCD), enough to
    // fill the GPU, few enough that the per-block partials stay small.  The
    // narrow tiles' SpMMs walk rows (a chain of dependent row-pointer,
    // column and gather loads per row), so the count follows the rows as
    // well as the elements: one workgroup per 128 rows at least (with one
    // per 1024 elements only, a single-column SpMV ran 248 workgroups and
    // took 86 us on the 250k-node mesh)
    long rows_work = std::max(((long)n * K + 1023) / 1024, ((long)n + 127) / 128);
    int G = (int)std::min<long>(TMI_WORK_G, std::max<long>(8, rows_work));
    G = (G + 7) / 8 * 8;
    w.G = G;
    size_t nb64 = (size_t)n * K * vbytes(vk64()), nbi = (size_t)n * K * vbytes(vin);
    w.x64.alloc(nb64); w.b64.alloc(nb64); w.r64.alloc(nb64);
    w.xi.alloc(nbi); w.ri.alloc(nbi); w.rdi.alloc(nbi); w.p.alloc(nbi); w.v.alloc(nbi); w.s.alloc(nbi); w.t.alloc(nbi);
    w.partial.alloc((size_t)RED_SLOTS * G * K);
    if (K == 64 && vin == VK_C32) w.rdq.alloc((size_t)n * 32);
    w.sc.alloc((size_t)S_NSLOT * K); w.osc.alloc((size_t)S_NSLOT * K);
    w.ic.alloc((size_t)I_NSLOT * K); w.oic.alloc((size_t)I_NSLOT * K);
    return w;
}

int Solver::count_active(Exec &e, const int *d_flags, int K)
{
    TMI_HIP(hipMemcpyAsync(e.h_flags, d_flags, sizeof(int) * K, hipMemcpyDeviceToHost, e.st));
    TMI_HIP(hipStreamSynchronize(e.st));
    e.n_sync++;
    int c = 0;
    for (int i = 0; i < K; i++) c += e.h_flags[i] != 0;
    return c;
}

#ifndef TMI_CHECK_EVERY
#define TMI_CHECK_EVERY 8
#endif
static constexpr int CHECK_EVERY = TMI_CHECK_EVERY;

// Repacking the running columns to the front of the tile pays when it
// frees a whole 128-byte line (16 complex FP32 columns) of every row, or
// when three or more lines are only partly active (each costs partial
// stores); flags: the running flag of each of the K columns.
static bool repack_pays(const int *flags, int K, int nact)
{
    const int LINE = 16, nl = K / LINE;
    int touched = 0, partial = 0;
    for (int l = 0; l < nl; l++) {
        int a = 0;
        for (int j = l * LINE; j < (l + 1) * LINE; j++) a += flags[j] != 0;
        touched += a > 0;
        partial += a > 0 && a < LINE;
    }
    const int packed = (nact + LINE - 1) / LINE;
    return touched > packed || partial >= 3;
}

// One tile: BiCGSTAB (complex) or PCG (real) on K columns with reliable
// updates (Sleijpen & van der Vorst): the inner iteration runs in the inner
// precision on the system scaled by 1/||q||; every `ru` iterations, and
// whenever every column's recursive residual has met the tolerance, the
// inner solution is folded into the FP64 solution, the true residual
// q - K x is recomputed in FP64 and replaces the inner residual -- the
// search directions are kept, so the iteration continues instead of
// restarting.  A column is done when its FP64 true residual meets the
// tolerance (the reference criterion ||r|| <= tol ||q||), or when it stops
// improving (no 4x drop of ||r||^2 over 8 true-residual evaluations).
int Solver::solve_tile(Exec &e, int K0, int kc, tmi_colstat *stats, void *d_x, int ld, int q0, int md,
                       const double *h_omega)
{
    hipStream_t st = e.st;
    int *h_flags = e.h_flags, *h_oflags = e.h_flags + I_NSLOT * 64;
    double *h_dbl = e.h_dbl;
    const int n = mesh->n;
    const bool mixed = (md == TMI_MODE_MIXED_BLOCK);
    // several frequencies: the operator at omega = 1 and each column's
    // omega (outer scalar slot S_FREQ, moved with the column by compaction)
    const bool mf = h_omega != nullptr;
    const void *val_in = mf ? (mixed ? (const void *)mfval32.p : (const void *)mfval64.p)
                            : (mixed ? (const void *)sval32.p : (const void *)sval64.p);
    const double *res_op = mf ? mfval64.p : sval64.p;   // the FP64 residual's operator
    const void *idiag_in = nullptr;
    // PCG (real) and COCG (complex symmetric: PCG with the unconjugated
    // bilinear form) share the launch sequence; COCG's scalars are complex
    const bool use_cg = (lsolver == TMI_LSOLVER_CG || lsolver == TMI_LSOLVER_COCG);
    const bool bil = (lsolver == TMI_LSOLVER_COCG);
    const int cg_dot = bil ? DOT_BIL : 0;
    // operator of the BiCGSTAB SpMMs: K D^-1 when diagonally preconditioned
    const void *op_in = val_in;
    if (precon == TMI_PRECON_DIAG) {
        if (mf)   // the diagonal itself: the kernels divide by re + i omega im per column
            idiag_in = mixed ? (const void *)mfdiag32.p : (const void *)mfdiag64.p;
        else
            idiag_in = mixed ? (const void *)idiag32.p : (const void *)idiag64.p;
        if (!use_cg) op_in = mixed ? (const void *)pval32.p : (const void *)pval64.p;
    }
    const int maxit_total = (mixed && mixed_maxit > 0) ? mixed_maxit : (maxit > 0 ? maxit : n + 1);
    // reliable updates only in the mixed mode: in FP64 the recursive
    // residual is the reference's own stopping test, and replacing it by the
    // computed true residual would inject eps*||K||*||x|| noise into the
    // far field, where the fields are ~1e-9 of their peak
    const int ru = mixed ? reliable_every : (1 << 30);
    const VKind vk = mixed ? vk32() : vk64();
    const int s64 = (int)vbytes(vk64()), sin = (int)vbytes(vk);
    const int *perm = ord().perm.p;

    // current workspace and the caller's tile column of each of its slots
    Exec::Work *wp = &ensure_work(e, K0, md);
    std::vector<int> cmap(K0, -1);
    for (int c = 0; c < kc; c++) cmap[c] = c;
    int K = K0;
    BlockDesc bd;
    double *sc, *osc, *part, *scale;
    const double *omg = nullptr;
    int *ic, *oic;
    const int *oact;
    // packed copy of the (sign-pattern) shadow residual, read by the
    // 64-column mixed BiCGSTAB kernels that stream it every iteration
    unsigned char *rdq = nullptr;
    auto bind = [&]() {
        K = wp->K;
        bd = block(*wp, K);
        rdq = (K == 64 && mixed && !use_cg && vk == VK_C32) ? wp->rdq.p : nullptr;
        sc = wp->sc.p;
        osc = wp->osc.p;
        part = wp->partial.p;
        ic = wp->ic.p;
        oic = wp->oic.p;
        oact = oic + I_ACTIVE * K;
        scale = osc + S_SCALE * K;
        omg = mf ? osc + S_FREQ * K : nullptr;
    };
    bind();

    std::vector<int> hic((size_t)I_NSLOT * K, 0);
    for (int c = 0; c < kc; c++) hic[I_ACTIVE * K + c] = 1;
    TMI_HIP(hipMemcpyAsync(oic, hic.data(), hic.size() * sizeof(int), hipMemcpyHostToDevice, st));
    TMI_HIP(hipMemsetAsync(osc, 0, wp->osc.bytes(), st));
    if (mf) {
        std::vector<double> hw(K, 0.0);
        for (int c = 0; c < kc; c++) hw[c] = h_omega[c];
        TMI_HIP(hipMemcpyAsync(osc + S_FREQ * K, hw.data(), K * sizeof(double), hipMemcpyHostToDevice, st));
    }
    const double tol2 = tol * tol;
    // the recursive residual aims a little below the target so that the
    // verifying FP64 residual usually passes at the first check
    std::vector<double> itol2(K, mixed ? 0.25 * tol2 : tol2);

    launch_norm2(vk64(), bd, wp->b64.p, oact, part, st);
    launch_finalize(FIN_OUTER_BNORM, bd, part, osc, oic, 0, 0, 0, st);
    launch_resid(vk64(), bd, res_op, wp->x64.p, wp->b64.p, wp->r64.p, oact, part, st, omg);
    launch_finalize(FIN_OUTER_RES, bd, part, osc, oic, 0, tol2, 0, st);
    // inner state
    launch_inner_reset(bd, osc, oic, sc, ic, st);
    TMI_HIP(hipMemcpyAsync(sc + S_TOL2 * K, itol2.data(), K * sizeof(double), hipMemcpyHostToDevice, st));
    // the mixed BiCGSTAB takes the sign pattern of r0 as its shadow residual
    // (fewer iterations); the FP64 modes keep the reference's rd = r0
    launch_to_inner(vk, bd, wp->r64.p, wp->ri.p, use_cg ? nullptr : wp->rdi.p, wp->xi.p, scale, oact, part, st,
                    mixed);
    if (rdq) launch_pack_shadow(n, wp->rdi.p, rdq, st);
    if (use_cg) {
        launch_cg_init(vk, bd, idiag_in, wp->p.p, wp->ri.p, sc, ic, part, st, bil, omg);
        launch_finalize(bil ? FIN_COCG_INIT : FIN_CG_INIT, bd, part, sc, ic, 0, 0, 0, st);
    } else {
        launch_finalize(FIN_BICG_INIT, bd, part, sc, ic, 0, 0, 0, st);
    }
    static const bool verbose = getenv("TMI_VERBOSE") != nullptr;
    int it = 0, nupd = 0, since = 0, notconv = 0, ncompact = 0;
    // the launches of a reliable update (no host work, so that a captured
    // period can hold them)
    auto reliable_launches = [&]() {
        launch_from_inner(vk, bd, wp->x64.p, wp->xi.p, scale, oact, st);
        launch_resid(vk64(), bd, res_op, wp->x64.p, wp->b64.p, wp->r64.p, oact, part, st, omg);
        launch_finalize(FIN_OUTER_RES, bd, part, osc, oic, 0, tol2, 0, st);
        launch_reliable(vk, bd, wp->r64.p, wp->ri.p, wp->xi.p, use_cg ? nullptr : wp->rdi.p, idiag_in, scale, oact,
                        part, st, bil, omg);
        launch_finalize(bil ? FIN_COCG_RELIABLE : use_cg ? FIN_CG_RELIABLE : FIN_BICG_RELIABLE, bd, part, sc, ic,
                        maxit_total, 0, 0, st, osc, oic);
    };
    auto reliable_update = [&]() {
        reliable_launches();
        nupd++;
        since = 0;
    };
    // per-column records of the slots whose caller column is in `which`,
    // from the device state; their solutions are written to d_x
    auto finish = [&](const std::vector<int> &which) {
        TMI_HIP(hipMemcpyAsync(h_flags, ic, sizeof(int) * I_NSLOT * K, hipMemcpyDeviceToHost, st));
        TMI_HIP(hipMemcpyAsync(h_oflags, oic, sizeof(int) * I_NSLOT * K, hipMemcpyDeviceToHost, st));
        TMI_HIP(hipMemcpyAsync(h_dbl, osc, sizeof(double) * S_NSLOT * K, hipMemcpyDeviceToHost, st));
        TMI_HIP(hipMemcpyAsync(e.cols.p, which.data(), sizeof(int) * K, hipMemcpyHostToDevice, st));
        launch_scatter_map(s64, n, K, perm, e.cols.p, wp->x64.p, d_x, ld, q0, st);
        TMI_HIP(hipStreamSynchronize(st));
        e.n_sync++;
        for (int j = 0; j < K; j++) {
            const int c = which[j];
            if (c < 0) continue;
            const double b2 = h_dbl[S_BNORM2 * K + j], r2 = h_dbl[S_RES2 * K + j];
            const bool conv = b2 == 0.0 || r2 <= tol2 * b2;
            if (!conv) notconv++;
            if (stats) {
                stats[c].iters = h_flags[I_ITERS * K + j];
                stats[c].sweeps = nupd;
                stats[c].relres = b2 > 0.0 ? std::sqrt(r2 / b2) : 0.0;
                stats[c].converged = conv ? 1 : 0;
            }
            e.last_iters += h_flags[I_ITERS * K + j];
        }
    };
    // narrow the tile to the columns still running: finished columns are
    // recorded and written out, the state of the others moves to the
    // workspace of the smallest power-of-two width that holds them.  With
    // repack, the width stays and the running columns move to the front of
    // the second workspace of that width: converged columns scattered over
    // the tile leave every 128-byte line of a row partly active (partial,
    // read-modify-write stores and no line skipped); packed, the lines past
    // the running columns are not touched at all.
    auto compact = [&](bool repack) -> bool {
        TMI_HIP(hipMemcpyAsync(h_oflags, oact, sizeof(int) * K, hipMemcpyDeviceToHost, st));
        TMI_HIP(hipStreamSynchronize(st));
        e.n_sync++;
        std::vector<int> done(K, -1), sel;
        for (int j = 0; j < K; j++) {
            if (cmap[j] < 0) continue;
            if (h_oflags[j]) sel.push_back(j);
            else done[j] = cmap[j];
        }
        int K2 = 1;
        while (K2 < (int)sel.size()) K2 *= 2;
        K2 = std::max(K2, compact_min);
        if (repack) K2 = K;
        else if (K2 >= K) return false;
        finish(done);
        const bool alt = repack && wp == &e.wk[log2i(K)];
        Exec::Work &w2 = ensure_work(e, K2, md, alt);
        const int ns = (int)sel.size();
        TMI_HIP(hipMemcpyAsync(e.cols.p + 64, sel.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st));
        const int *d_sel = e.cols.p + 64;
        auto mv = [&](int bytes, long rows, const void *src, void *dst) {
            launch_select_cols(bytes, rows, K, K2, ns, d_sel, src, dst, st);
        };
        // r64 is recomputed (b - A x) before every read; CG/COCG read no
        // shadow residual, and their v (= A p) is rewritten by the next
        // SpMM; BiCGSTAB's s and t are rewritten before they are read
        mv(s64, n, wp->x64.p, w2.x64.p);
        mv(s64, n, wp->b64.p, w2.b64.p);
        mv(sin, n, wp->xi.p, w2.xi.p);
        mv(sin, n, wp->ri.p, w2.ri.p);
        mv(sin, n, wp->p.p, w2.p.p);
        if (!use_cg) {
            mv(sin, n, wp->rdi.p, w2.rdi.p);
            mv(sin, n, wp->v.p, w2.v.p);
        }
        mv(8, S_NSLOT, sc, w2.sc.p);
        mv(8, S_NSLOT, osc, w2.osc.p);
        mv(4, I_NSLOT, ic, w2.ic.p);
        mv(4, I_NSLOT, oic, w2.oic.p);
        std::vector<int> cm2(K2, -1);
        for (int j = 0; j < ns; j++) cm2[j] = cmap[sel[j]];
        cmap.swap(cm2);
        wp = &w2;
        bind();
        if (rdq) launch_pack_shadow(n, wp->rdi.p, rdq, st);
        ncompact++;
        return true;
    };
    // the launch sequence of one check period (no host work inside)
    auto period_launches = [&](bool ru_due) {
        for (int k = 0; k < CHECK_EVERY; k++) {
            if (use_cg) {
                // COCG defers x += alpha p to the direction update, except in
                // the iteration whose reliable update folds x between the two
                const bool rel = k == CHECK_EVERY - 1 && ru_due, xd = bil && !rel;
                launch_spmm_dot(vk, bd, val_in, nullptr, wp->p.p, wp->v.p, wp->p.p, cg_dot, ic + I_ACTIVE * K, part,
                                st, nullptr, omg);
                launch_finalize(bil ? FIN_COCG_ALPHA : FIN_CG_ALPHA, bd, part, sc, ic, 0, 0, 0, st);
                launch_cg_xr(vk, bd, idiag_in, wp->xi.p, wp->p.p, wp->v.p, wp->ri.p, sc, ic, part, st, bil, omg, xd);
                launch_finalize(bil ? FIN_COCG_BETA : FIN_CG_BETA, bd, part, sc, ic, maxit_total, 0, 0, st);
                if (rel) reliable_launches();
                launch_cg_d(vk, bd, idiag_in, wp->p.p, wp->ri.p, sc, ic, st, bil, omg, xd ? wp->xi.p : nullptr);
            } else {
                launch_bicg_p(vk, bd, wp->ri.p, wp->p.p, wp->v.p, sc, ic, st);
                launch_spmm_dot(vk, bd, op_in, nullptr, wp->p.p, wp->v.p, wp->rdi.p, false, ic + I_ACTIVE * K, part,
                                st, rdq);
                launch_finalize(FIN_BICG_ALPHA, bd, part, sc, ic, 0, 0, 0, st);
                launch_bicg_s(vk, bd, wp->ri.p, wp->v.p, wp->s.p, sc, ic, st);
                launch_spmm_dot(vk, bd, op_in, nullptr, wp->s.p, wp->t.p, wp->s.p, true, ic + I_ACTIVE * K, part, st);
                launch_finalize(FIN_BICG_OMEGA, bd, part, sc, ic, 0, 0, 0, st);
                launch_bicg_xr(vk, bd, idiag_in, wp->xi.p, wp->p.p, wp->s.p, wp->t.p, wp->ri.p, wp->rdi.p, sc, ic,
                               part, st, rdq);
                launch_finalize(FIN_BICG_RHO, bd, part, sc, ic, maxit_total, 0, 0, st);
            }
        }
        if (!use_cg && ru_due) reliable_launches();
    };
    // run a check period from an instantiated HIP graph: captured on first
    // use for this key -- every value its kernels take as argument -- and
    // replayed afterwards (one graph launch instead of 64-104 kernel launches)
    auto run_period = [&](bool ru_due) {
        if (!use_graphs) {
            period_launches(ru_due);
            return;
        }
        typedef unsigned long long u64;
        auto P = [](const void *p) { return (u64)(uintptr_t)p; };
        u64 tol_bits = 0;
        std::memcpy(&tol_bits, &tol2, sizeof tol_bits);
        const std::vector<u64> key = {
            (u64)K, (u64)vk, (u64)lsolver, (u64)mf, (u64)ru_due, (u64)maxit_total, tol_bits, (u64)bd.nblk, (u64)bd.rpw,
            P(bd.rowptr), P(bd.colidx), (u64)bd.lds.nb, P(bd.lds.brow), P(bd.lds.bu), P(bd.lds.ucols),
            P(bd.lds.lcol), (u64)bd.lds.dma, (u64)bd.rb.nb, P(bd.rb.ofs), P(bd.rb.col), P(bd.rb.mask),
            P(bd.rb.val), P(bd.rb.src), P(bd.rb.word), P(op_in), P(val_in), P(idiag_in), P(rdq), P(sc), P(ic),
            P(osc), P(oic), P(part), P(res_op), P(wp->x64.p), P(wp->b64.p), P(wp->r64.p), P(wp->xi.p),
            P(wp->ri.p), P(wp->rdi.p), P(wp->p.p), P(wp->v.p), P(wp->s.p), P(wp->t.p), P(st)};
        hipGraphExec_t exec = nullptr;
        for (const Exec::Graph &g : e.graphs)
            if (g.key == key) { exec = g.exec; break; }
        if (!exec) {
            const long long calls = g_launch_calls;
            TMI_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            hipGraph_t graph = nullptr;
            try {
                period_launches(ru_due);
            } catch (...) {
                (void)hipStreamEndCapture(st, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            TMI_HIP(hipStreamEndCapture(st, &graph));
            const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            TMI_HIP(ie);
            g_launch_calls = calls;   // captured, not launched
            if (e.graphs.size() >= 32) {   // keep the cache small: drop the oldest
                // it may have been launched moments ago: let the stream drain first
                TMI_HIP(hipStreamSynchronize(st));
                e.n_sync++;
                (void)hipGraphExecDestroy(e.graphs.front().exec);
                e.graphs.erase(e.graphs.begin());
            }
            e.graphs.push_back(Exec::Graph{key, exec});
        }
        TMI_HIP(hipGraphLaunch(exec, st));
        e.n_graph++;
    };
    for (;;) {
        const int nact = count_active(e, oact, K);
        if (nact == 0) break;
        if (compact_min > 0 && nact <= K / 2 && K / 2 >= compact_min) compact(false);
        else if (compact_min > 0 && K >= 32 && repack_pays(e.h_flags, K, nact)) compact(true);
        if (count_active(e, ic + I_ACTIVE * K, K) == 0) {
            // every column's recursive residual met the tolerance: FP64 modes
            // stop here (as the reference does); the mixed mode verifies in FP64
            if (!mixed) break;
            reliable_update();
            if (use_cg) launch_cg_d(vk, bd, idiag_in, wp->p.p, wp->ri.p, sc, ic, st, bil, omg);
            if (count_active(e, oact, K) == 0 || it >= maxit_total) break;
            continue;
        }
        if (it >= maxit_total) break;
        // one check period: CHECK_EVERY iterations and, when it falls due,
        // the reliable update (PCG replaces the residual before the last
        // direction update, BiCGSTAB after the period)
        const bool ru_due = since + CHECK_EVERY >= ru;
        run_period(ru_due);
        it += CHECK_EVERY;
        if (ru_due) {
            nupd++;
            since = use_cg ? 1 : 0;   // PCG's last direction update follows its replacement
        } else {
            since += CHECK_EVERY;
        }
    }
    // fold the last inner correction of columns still marked active (maxit
    // exhausted) and record the final true residuals
    launch_from_inner(vk, bd, wp->x64.p, wp->xi.p, scale, oact, st);
    launch_resid(vk64(), bd, res_op, wp->x64.p, wp->b64.p, wp->r64.p, oact, part, st, omg);
    launch_finalize(FIN_OUTER_RES, bd, part, osc, oic, 0, tol2, 0, st);
    finish(cmap);
    if (verbose)
        fprintf(stderr, "[tmi] tile K=%d: %d block iterations, %d reliable updates, %d compactions (final width %d), "
                "%d not converged\n", K0, it, nupd, ncompact, K, notconv);
    return notconv;
}

void Solver::prepare_row_blocks64(const double *op)
{
    if (rb64_src == op) return;
    const Mesh::NodeOrder &o = ord();
    rbval64.alloc((size_t)2 * mesh->nnz);
    launch_gather_vals64(mesh->nnz, 2, o.rb_slot.p, op, rbval64.p, st);
    TMI_HIP(hipStreamSynchronize(st));   // other streams read it next
    rb64_src = op;
}

void Solver::prepare_row_blocks(const float *op)
{
    if (rb_src == op) return;
    const Mesh::NodeOrder &o = ord();
    rbval32.alloc(nval() + 8);
    launch_gather_vals32(mesh->nnz, cplx ? 2 : 1, o.rb_slot.p, op, rbval32.p, st);
    TMI_HIP(hipStreamSynchronize(st));   // other streams read it next
    rb_src = op;
}

int Solver::solve_on(Exec &e, int nq, const void *d_b, void *d_x, int ld, tmi_colstat *stats, int md,
                     const double *h_omega)
{
    TMI_CHECK(assembled, TMI_ERR_STATE, "CalcFields called before Reset/AssembleSystemMatrix");
    TMI_CHECK(!(cplx && lsolver == TMI_LSOLVER_CG), TMI_ERR_UNSUPPORTED,
              "CG on a complex system is not supported (the reference's complex CG returns the initial guess)");
    TMI_CHECK(!(!cplx && lsolver == TMI_LSOLVER_COCG), TMI_ERR_UNSUPPORTED,
              "COCG solves complex symmetric systems; use CG for a real one");
    if (md < 0) md = mode;
    const bool mf = h_omega != nullptr;
    if (mf) {
        TMI_CHECK(cplx && lsolver == TMI_LSOLVER_COCG, TMI_ERR_UNSUPPORTED,
                  "a multi-frequency solve needs a complex solver with COCG");
        TMI_CHECK(md == TMI_MODE_MIXED_BLOCK || md == TMI_MODE_FP64_BLOCK, TMI_ERR_UNSUPPORTED,
                  "a multi-frequency solve runs in the block modes (mixed, fp64_block)");
        for (int c = 0; c < nq; c++)
            TMI_CHECK(std::isfinite(h_omega[c]), TMI_ERR_INVALID_ARG, "frequency of column %d is not finite", c);
        std::lock_guard<std::mutex> lk(prep_mu);
        prepare_mf();
    }
    if (lsolver == TMI_LSOLVER_GMRES) return gmres_on(e, nq, d_b, d_x, ld, stats);
    if (precon == TMI_PRECON_DIAG) {
        std::lock_guard<std::mutex> lk(prep_mu);
        if (!precon_valid) {
            reset_precon();
            TMI_HIP(hipStreamSynchronize(st));   // other streams read it next
        }
    }
    if (spmm_variant >= 3 && cplx) {
        std::lock_guard<std::mutex> lk(prep_mu);
        prepare_row_blocks(mf ? mfval32.p : spmm_op32());
        if (!mf) prepare_row_blocks64(sval64.p);   // the FP64 residual's operator
    }
    if (!e.st) e.st = st;
    std::vector<tmi_colstat> own;
    if (!stats) {
        own.resize(nq);
        stats = own.data();
    }
    const int K = mf ? std::max(2, tile_for(nq, md)) : tile_for(nq, md);
    hipStream_t es_ = e.st;
    size_t es = vbytes(vk64());
    hipEvent_t e0, e1;
    TMI_HIP(hipEventCreate(&e0));
    TMI_HIP(hipEventCreate(&e1));
    TMI_HIP(hipEventRecord(e0, es_));
    e.last_iters = 0.0;
    const long long calls0 = g_launch_calls, graphs0 = e.n_graph, syncs0 = e.n_sync;
    int notconv = 0;
    for (int q0 = 0; q0 < nq; q0 += K) {
        int kc = std::min(K, nq - q0);
        // gather the tile: columns q0..q0+kc of [n][ld] -> [n][K] in the
        // solver's node order (columns kc..K-1 zero)
        auto &w = ensure_work(e, K, md);
        const int *perm = ord().perm.p;
        launch_gather_tile((int)es, mesh->n, K, kc, perm, d_b, ld, q0, w.b64.p, es_);
        launch_gather_tile((int)es, mesh->n, K, kc, perm, d_x, ld, q0, w.x64.p, es_);
        notconv += solve_tile(e, K, kc, stats + q0, d_x, ld, q0, md, mf ? h_omega + q0 : nullptr);
    }
    double iters = e.last_iters;
    // a column the mixed-precision iteration could not bring to the
    // tolerance (a BiCGSTAB breakdown in FP32) is solved again by the FP64
    // block iteration, starting from where the mixed solve left it
    if (notconv && md == TMI_MODE_MIXED_BLOCK && fp64_fallback) {
        std::vector<int> bad;
        for (int c = 0; c < nq; c++)
            if (!stats[c].converged) bad.push_back(c);
        const int nb = (int)bad.size();
        DBuf<char> tb((size_t)mesh->n * nb * es), tx((size_t)mesh->n * nb * es);
        for (int j = 0; j < nb; j++) {
            TMI_HIP(hipMemcpy2DAsync(tb.p + j * es, nb * es, (const char *)d_b + bad[j] * es, (size_t)ld * es, es,
                                     mesh->n, hipMemcpyDeviceToDevice, es_));
            TMI_HIP(hipMemcpy2DAsync(tx.p + j * es, nb * es, (const char *)d_x + bad[j] * es, (size_t)ld * es, es,
                                     mesh->n, hipMemcpyDeviceToDevice, es_));
        }
        std::vector<tmi_colstat> st2(nb);
        std::vector<double> om2;
        if (mf)
            for (int j = 0; j < nb; j++) om2.push_back(h_omega[bad[j]]);
        solve_on(e, nb, tb.p, tx.p, nb, st2.data(), TMI_MODE_FP64_BLOCK, mf ? om2.data() : nullptr);
        iters += e.last_iters * nb;
        notconv = 0;
        for (int j = 0; j < nb; j++) {
            TMI_HIP(hipMemcpy2DAsync((char *)d_x + bad[j] * es, (size_t)ld * es, tx.p + j * es, nb * es, es, mesh->n,
                                     hipMemcpyDeviceToDevice, es_));
            tmi_colstat &c = stats[bad[j]];
            c.iters += st2[j].iters;
            c.sweeps += 1;
            c.relres = st2[j].relres;
            c.converged = st2[j].converged;
            if (!c.converged) notconv++;
        }
        TMI_HIP(hipStreamSynchronize(es_));
        e.n_sync++;
    }
    TMI_HIP(hipEventRecord(e1, es_));
    TMI_HIP(hipEventSynchronize(e1));
    e.n_sync++;
    e.last_launch = g_launch_calls - calls0;
    e.last_graph = e.n_graph - graphs0;
    e.last_sync = e.n_sync - syncs0;
    float ms = 0.f;
    TMI_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    e.last_solve_ms = ms;
    e.last_iters = nq ? iters / nq : 0.0;
    return notconv;
}

void Solver::time_kernel(int which, int reps, double *avg_ms, double *bytes)
{
    auto &w = ex.wk[log2i(tile_for(64, mode))];
    hipStream_t st = ex.st ? ex.st : this->st;
    TMI_CHECK(lsolver != TMI_LSOLVER_GMRES, TMI_ERR_UNSUPPORTED, "kernel timing covers the BiCGSTAB/PCG kernels");
    TMI_CHECK(w.K > 0 && assembled, TMI_ERR_STATE, "time_kernel needs a previous CalcFields");
    TMI_CHECK(reps > 0 && which >= 0 && which <= 1, TMI_ERR_INVALID_ARG, "bad kernel timing request");
    bool mixed = (mode == TMI_MODE_MIXED_BLOCK);
    // the workspace must hold vectors of the precision the current mode's
    // operator reads: an FP64 fallback or a mode change since the last
    // CalcFields can leave FP64 vectors where FP32 ones are expected
    TMI_CHECK(w.vk_in == (mixed ? vk32() : vk64()), TMI_ERR_STATE,
              "time_kernel: the last CalcFields left this tile's workspace in another precision than the "
              "current mode; run CalcFields in this mode first");
    int K = w.K, n = mesh->n;
    BlockDesc bd = block(w, K);
    std::vector<int> ones(K, 1);
    DBuf<int> act;
    act.upload(ones.data(), K, st);
    const void *idg = nullptr;
    const void *op = mixed ? (const void *)sval32.p : (const void *)sval64.p;
    if (precon == TMI_PRECON_DIAG && !pcg()) op = mixed ? (const void *)pval32.p : (const void *)pval64.p;
    // the 64-column complex FP32 BiCGSTAB SpMM reads the packed shadow
    // residual in the solve, and so here; PCG / COCG dot with the direction
    const unsigned char *wq =
        (mixed && cplx && K == 64 && !pcg() && w.vk_in == VK_C32 && w.rdq.p) ? w.rdq.p : nullptr;
    const void *wv = pcg() ? w.p.p : w.rdi.p;
    const int dot = lsolver == TMI_LSOLVER_COCG ? DOT_BIL : 0;
    auto launch = [&]() {
        if (which == 0)
            launch_spmm_dot(w.vk_in, bd, op, idg, w.p.p, w.v.p, wv, dot, act.p, w.partial.p, st, wq);
        else
            launch_resid(vk64(), bd, sval64.p, w.x64.p, w.b64.p, w.r64.p, act.p, w.partial.p, st);
    };
    launch();   // warm
    hipEvent_t e0, e1;
    TMI_HIP(hipEventCreate(&e0));
    TMI_HIP(hipEventCreate(&e1));
    TMI_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++) launch();
    TMI_HIP(hipEventRecord(e1, st));
    TMI_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    TMI_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
    // algorithmic bytes: matrix values + column indices once, row pointers,
    // each X row read once, each Y row written once, and BiCGSTAB's shadow
    // residual for the fused dot (4 bits per entry where the packed copy is
    // read).  For PCG / COCG the dot's W is X itself, already counted (the
    // kernel reads a row's W right after gathering that row as a neighbour,
    // from the cache).
    double sv = (double)vbytes(which == 0 ? w.vk_in : vk64());
    double nnz = mesh->nnz;
    double b = nnz * (sv + 4.0) + 4.0 * (n + 1) + 2.0 * n * K * sv;
    if (which == 0 && !pcg()) b += wq ? (double)n * K / 2.0 : (double)n * K * sv;

    if (which == 1) b += (double)n * K * sv;   // the right-hand side block is read as well
    *bytes = b;
}

void Solver::autotune(int nq, const void *d_b, void *d_x, int ld, int *tile_out, int *rpw_out, int *variant_out,
                      double *ms_out)
{
    TMI_CHECK(assembled, TMI_ERR_STATE, "autotune before Reset");
    if (lsolver == TMI_LSOLVER_GMRES) {
        // GMRES has no SpMM variant or tile choice to tune: report the
        // current settings and the time of one solve
        solve(nq, d_b, d_x, ld, nullptr);
        if (tile_out) *tile_out = tile;
        if (rpw_out) *rpw_out = rpw;
        if (variant_out) *variant_out = spmm_variant;
        if (ms_out) *ms_out = ex.last_solve_ms;
        return;
    }
    const size_t es = vbytes(vk64());
    auto zero_x = [&]() { TMI_HIP(hipMemset2DAsync(d_x, (size_t)ld * es, 0, (size_t)nq * es, mesh->n, st)); };
    // 1. row schedule and SpMM variant of the 64-column kernel (mixed mode;
    // fewer than 64 right-hand sides never run a 64-column tile)
    if (mode == TMI_MODE_MIXED_BLOCK && cplx && nq >= 64) {
        const int tile0 = tile;
        const bool fb0 = fp64_fallback;
        tile = 64;
        zero_x();
        // the trial solve only fills the 64-column workspace the kernel
        // timings read: without the FP64 fallback it keeps FP32 vectors
        fp64_fallback = false;
        try {
            solve(std::min(nq, 64), d_b, d_x, ld, nullptr);
        } catch (...) {
            fp64_fallback = fb0;
            tile = tile0;
            throw;
        }
        fp64_fallback = fb0;
        struct Cand { int variant; int rpw; };
        const Cand cands[] = {{1, 0}, {1, 2}, {1, 4}, {1, 8}, {1, 16}, {0, 0}, {2, 0}, {3, 0}, {4, 0}, {5, 0}};
        double best = 1e30;
        Cand pick{spmm_variant, rpw};
        for (const Cand &c : cands) {
            if ((c.variant == 0 || c.variant == 2) && ord().nb == 0) continue;
            spmm_variant = c.variant;
            if (c.variant >= 3) {
                std::lock_guard<std::mutex> lk(prep_mu);
                prepare_row_blocks(spmm_op32());
            }
            rpw = c.rpw;
            double ms = 0, bytes = 0;
            time_kernel(0, 10, &ms, &bytes);
            if (ms < best) { best = ms; pick = c; }
        }
        // the row-block variants differ most on the narrow tiles column
        // compaction leaves (the matrix-core kernel also runs the 32- and
        // 16-column ones): among them, whole 64-column solves decide
        if (pick.variant >= 3 && ord().rb_word_ok) {
            double bs = 1e30;
            Cand sp = pick;
            fp64_fallback = false;   // as for the trial solve above
            try {
                for (const int v : {3, 5}) {
                    spmm_variant = v;
                    rpw = 0;
                    {
                        std::lock_guard<std::mutex> lk(prep_mu);
                        prepare_row_blocks(spmm_op32());
                    }
                    zero_x();
                    solve(std::min(nq, 64), d_b, d_x, ld, nullptr);
                    if (ex.last_solve_ms < bs) { bs = ex.last_solve_ms; sp = Cand{v, 0}; }
                }
            } catch (...) {
                fp64_fallback = fb0;
                tile = tile0;
                throw;
            }
            fp64_fallback = fb0;
            pick = sp;
        }
        spmm_variant = pick.variant;
        rpw = pick.rpw;
        tile = tile0;
    }
    // 2. tile width: whole solves of the caller's right-hand sides
    if (mode != TMI_MODE_FP64_SINGLE) {
        double best = 1e30;
        int pick = tile;
        for (int t = 64; t >= 8; t /= 2) {
            if (t < 64 && t >= tile_for(nq) * 2) continue;   // a wider tile than the columns need
            tile = t;
            zero_x();
            solve(nq, d_b, d_x, ld, nullptr);
            if (ex.last_solve_ms < best) { best = ex.last_solve_ms; pick = t; }
        }
        tile = pick;
        if (ms_out) *ms_out = best;
    }
    if (tile_out) *tile_out = tile;
    if (rpw_out) *rpw_out = rpw;
    if (variant_out) *variant_out = spmm_variant;
}

}  // namespace tmi
