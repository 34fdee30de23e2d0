// Forward solver: the GPU counterpart of TFwdSolver<T> (fwdsolver.h:77-508).
#pragma once
#include <cstdlib>
#include <mutex>

#include "krylov.hpp"
#include "mesh.hpp"

namespace tmi {

// workspace of the GMRES tiles (gmres.hip)
struct GmresWork;
void free_gmres_work(GmresWork *w);
size_t gmres_work_bytes(const GmresWork *w);
size_t gmres_tile_bytes(size_t n, size_t K, size_t es);

struct Solver {
    Mesh *mesh = nullptr;
    bool cplx = true;
    int lsolver = TMI_LSOLVER_BICGSTAB;
    // PCG and COCG share one launch sequence (one SpMM per iteration)
    bool pcg() const { return lsolver == TMI_LSOLVER_CG || lsolver == TMI_LSOLVER_COCG; }
    // the FP32 operator of the mixed SpMMs: K D^-1 for a diagonally
    // preconditioned BiCGSTAB, K itself for PCG / COCG
    const float *spmm_op32() const { return (precon == TMI_PRECON_DIAG && !pcg()) ? pval32.p : sval32.p; }
    double tol = 1e-10;
    int precon = TMI_PRECON_NONE;
    int maxit = 0;
    int mode = TMI_MODE_MIXED_BLOCK;
    int tile = 0;
#ifndef TMI_RELIABLE_EVERY
#define TMI_RELIABLE_EVERY 16
#endif
    int reliable_every = TMI_RELIABLE_EVERY;   // inner iterations between FP64 residual replacements (mixed mode)
    int order_kind = ORDER_CLUSTER;  // node order of the solver's copy of the system
    int rpw = 4;                     // rows per run of the 64-column SpMM (0: one contiguous share per wave)
    int spmm_variant = 5;            // 64-column SpMM: 0 LDS (register staging), 1 direct gathers, 2 LDS DMA,
                                     // 3 row blocks with shared gathers, 4 the same pipelined, 5 (default)
                                     // row blocks on the matrix cores (also the 32- and 16-column tiles)
    int compact_min = 2;             // narrowest tile the column compaction goes down to (0: off)
    bool allocated = false, assembled = false, precon_valid = false;
    void prepare_row_blocks(const float *op);   // rbval32 for the FP32 operator op (caller holds prep_mu)
    double omega = 0.0;
    hipStream_t st = nullptr;

    DBuf<double> val64;    // nnz (real) or 2 nnz (complex), reference node order
    DBuf<double> sval64;   // the same in the solver's node order (Mesh::order)
    DBuf<float> sval32;
    DBuf<char> idiag64, idiag32;
    DBuf<double> pval64;   // A D^-1 (FP64, solver order) for the right-preconditioned inner iteration
    DBuf<float> pval32;    // A D^-1 (FP32, solver order)
    DBuf<float> rbval32;   // the FP32 operator in the row-block SpMM's coefficient order
    const void *rb_src = nullptr;   // the operator rbval32 holds (null: none)
    DBuf<double> rbval64;  // the complex FP64 operator in that order (the 64-column FP64 residual)
    const void *rb64_src = nullptr;   // the operator rbval64 holds (null: none)
    void prepare_row_blocks64(const double *op);   // rbval64 for op (caller holds prep_mu)
    DBuf<double> coef;     // cmua | ckappa | c2a on the device
    bool elbasis = false;  // coef holds per-element coefficients
    // multi-frequency operator: the system assembled at omega = 1 (real part
    // A, imaginary part the mass term B), in the solver's node order, and
    // its diagonal; a column at omega w solves (A + i w B) x = b
    DBuf<double> mfval64;
    DBuf<float> mfval32;
    DBuf<char> mfdiag64, mfdiag32;
    bool mf_valid = false;
    void prepare_mf();   // caller holds prep_mu

    // per-solve state: a stream, the workspace of one tile, pinned host
    // mirrors of the per-column flags and the last solve's timing.  The
    // solver's own context runs on the library stream; a serving loop gives
    // each in-flight request its own context, so that solves of one or
    // several solvers run concurrently (the assembled matrix is read-only
    // while solving).
    struct Exec {
        hipStream_t st = nullptr;
        bool own_stream = false;
        struct Work {
            int K = 0, n = 0, G = 0;
            VKind vk_in = VK_C32;
            DBuf<char> x64, b64, r64, xi, ri, rdi, p, v, s, t;
            DBuf<unsigned char> rdq;   // packed shadow residual (64-column complex FP32 tiles)
            DBuf<double> partial, sc, osc;
            DBuf<int> ic, oic;
        };
        Work wk[7];                    // one workspace per tile width 1, 2, 4, ..., 64
        Work wk_alt[7];                // second workspace of a width, for repacking in place
        Work &w = wk[6];               // the 64-column workspace
        DBuf<int> cols;                // column maps of compaction (2 x 64)
        int *h_flags = nullptr;        // pinned host mirror of per-column ints (inner, outer)
        double *h_dbl = nullptr;       // pinned host mirror of per-column doubles
        GmresWork *gm = nullptr;       // GMRES workspace (allocated by the first GMRES solve)
        double last_solve_ms = 0.0, last_iters = 0.0;
        // instantiated graphs of the check period's launch sequence, keyed by
        // every value the captured kernels take as argument
        struct Graph {
            std::vector<unsigned long long> key;
            hipGraphExec_t exec = nullptr;
        };
        std::vector<Graph> graphs;
        void drop_graphs();
        // host-side work: graph launches and host synchronisations so far, and
        // of the last solve the kernel launch calls (outside graph capture),
        // graph launches and synchronisations
        long long n_graph = 0, n_sync = 0;
        long long last_launch = 0, last_graph = 0, last_sync = 0;
        Exec() = default;
        explicit Exec(bool new_stream);
        Exec(const Exec &) = delete;
        Exec &operator=(const Exec &) = delete;
        ~Exec();
        size_t bytes() const;          // device bytes held by the workspace
        void release_work();           // free every tile workspace (waits for the stream first)
    };
    Exec ex;                           // default context (library stream)
    std::mutex prep_mu;                // lazy preconditioner setup

    ~Solver();
    VKind vk64() const { return cplx ? VK_C64 : VK_R64; }
    VKind vk32() const { return cplx ? VK_C32 : VK_R32; }
    int nval() const { return cplx ? 2 * mesh->nnz : mesh->nnz; }
    const Mesh::NodeOrder &ord() const { return mesh->order[order_kind]; }
    void allocate();
    // elbasis: per-element coefficients (elen each, ASSEMBLE_*_EL) instead of nodal ones
    void assemble(const double *cmua, const double *ckap, const double *c2a, double omega, bool host_ptrs,
                  bool elbasis = false);
    void reset_precon();
    // Solve A X = B for nq columns.  d_b, d_x: [n][ld] node-major device
    // arrays (FP64, complex or real); d_x holds the initial guess.
    int solve(int nq, const void *d_b, void *d_x, int ld, tmi_colstat *stats) { return solve_on(ex, nq, d_b, d_x, ld, stats); }
    // md: precision / batching mode of this solve (-1: the solver's mode)
    // h_omega (host, nq values, complex COCG, block modes): column c solves
    // at omega h_omega[c] -- several frequencies in one block solve
    int solve_on(Exec &e, int nq, const void *d_b, void *d_x, int ld, tmi_colstat *stats, int md = -1,
                 const double *h_omega = nullptr);
    // device bytes a solve of nq columns needs for its workspace
    size_t work_bytes(int nq) const;
    int tile_for(int nq, int md = -1) const;
    // replay the check period's launches from captured HIP graphs (default
    // on; the environment variable TMI_GRAPHS=0 turns it off for new solvers,
    // e.g. for kernel traces under a profiler)
    bool use_graphs = graphs_default();
    static bool graphs_default()
    {
        const char *v = getenv("TMI_GRAPHS");
        return !(v && v[0] == '0');
    }
    bool fp64_fallback = true;       // mixed mode: columns that do not converge are re-solved in FP64
    int mixed_maxit = 0;             // mixed mode: iteration cap of the FP32 iteration (0: maxit)
    // device buffers of ProjectAll (link table, mvec CSR, partial sums,
    // projections), kept between calls and grown when a call needs more
    DBuf<int> pj_lk, pj_rp, pj_ci;
    DBuf<double> pj_val, pj_part, pj_out;
    // Time `reps` launches of one kernel of the last tile (which: 0 = inner
    // SpMM, 1 = FP64 residual SpMM) with HIP events on the solver stream.
    // Returns the average ms and the algorithmic bytes of one launch.
    void time_kernel(int which, int reps, double *avg_ms, double *bytes);
    // Pick the row schedule / SpMM variant (timed 64-column SpMM launches)
    // and the tile width (timed solves of the caller's right-hand sides,
    // d_x used as scratch and left holding the last solve) and keep them.
    void autotune(int nq, const void *d_b, void *d_x, int ld, int *tile_out, int *rpw_out, int *variant_out,
                  double *ms_out);

   private:
    BlockDesc block(const Exec::Work &w, int K) const;
    Exec::Work &ensure_work(Exec &e, int K, int md, bool alt = false);
    int solve_tile(Exec &e, int K, int kc, tmi_colstat *stats, void *d_x, int ld, int q0, int md,
                   const double *h_omega);
    // GMRES(20) in FP64 on tiles of the columns (gmres.hip); every mode
    int gmres_on(Exec &e, int nq, const void *d_b, void *d_x, int ld, tmi_colstat *stats);
    int count_active(Exec &e, const int *d_flags, int K);
};


}  // namespace tmi
