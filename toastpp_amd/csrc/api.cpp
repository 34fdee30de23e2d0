// C-ABI of libtoastpp_amd (include/toastpp_amd.h).  Every entry point
// converts C++ exceptions into status codes; nothing here exits the process.
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <string>

#include "solver.hpp"

#include "handles.hpp"

namespace tmi {

static thread_local std::string g_last_error;
thread_local long long g_launch_calls = 0;
void set_last_error(const std::string &msg) { g_last_error = msg; }

Context &Context::get()
{
    static Context ctx;
    static std::once_flag once;
    static int init_status = 0;
    static std::string init_msg;
    std::call_once(once, [] {
        int ndev = 0;
        hipError_t e = hipGetDeviceCount(&ndev);
        if (e != hipSuccess || ndev < 1) {
            init_status = TMI_ERR_DEVICE;
            init_msg = std::string("no HIP device available: ") + hipGetErrorString(e);
            return;
        }
        if (hipSetDevice(0) != hipSuccess || hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking) != hipSuccess) {
            init_status = TMI_ERR_DEVICE;
            init_msg = "cannot initialise HIP device 0";
            return;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess) ctx.num_cu = prop.multiProcessorCount;
    });
    if (init_status) throw Error(init_status, init_msg);
    (void)hipSetDevice(ctx.device);
    return ctx;
}

}  // namespace tmi

using namespace tmi;

#define API_BEGIN try {
#define API_END                                       \
    return TMI_OK;                                    \
    }                                                 \
    catch (const tmi::Error &e)                       \
    {                                                 \
        set_last_error(e.what());                     \
        return e.code;                                \
    }                                                 \
    catch (const std::bad_alloc &)                    \
    {                                                 \
        set_last_error("host allocation failed");     \
        return TMI_ERR_OOM;                           \
    }                                                 \
    catch (const std::exception &e)                   \
    {                                                 \
        set_last_error(e.what());                     \
        return TMI_ERR_INVALID_ARG;                   \
    }

#define NEED(p) TMI_CHECK((p) != nullptr, TMI_ERR_INVALID_ARG, "%s must not be NULL", #p)

static tmi::Mesh &M(tmi_mesh h)
{
    TMI_CHECK(h != nullptr, TMI_ERR_INVALID_ARG, "invalid mesh handle");
    return h->m;
}
static tmi::Solver &S(tmi_solver h)
{
    TMI_CHECK(h != nullptr, TMI_ERR_INVALID_ARG, "invalid solver handle");
    return h->s;
}

extern "C" {

const char *tmi_last_error(void) { return g_last_error.c_str(); }

int tmi_device_info(char *name, int name_len, double *hbm_gib)
{
    API_BEGIN
    Context &c = Context::get();
    hipDeviceProp_t prop;
    TMI_HIP(hipGetDeviceProperties(&prop, c.device));
    if (name && name_len > 0) {
        snprintf(name, name_len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    }
    if (hbm_gib) *hbm_gib = prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0);
    API_END
}

int tmi_synchronize(void)
{
    API_BEGIN
    TMI_HIP(hipStreamSynchronize(Context::get().stream));
    API_END
}

int tmi_device_alloc(unsigned long long bytes, void **ptr)
{
    API_BEGIN
    NEED(ptr);
    Context::get();
    TMI_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    API_END
}

int tmi_device_free(void *ptr)
{
    API_BEGIN
    Context::get();
    if (ptr) TMI_HIP(hipFree(ptr));
    API_END
}

int tmi_memcpy_h2d(void *dst, const void *src, unsigned long long bytes)
{
    API_BEGIN
    hipStream_t st = Context::get().stream;
    TMI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_memcpy_d2h(void *dst, const void *src, unsigned long long bytes)
{
    API_BEGIN
    hipStream_t st = Context::get().stream;
    TMI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_memcpy_d2d(void *dst, const void *src, unsigned long long bytes)
{
    API_BEGIN
    hipStream_t st = Context::get().stream;
    TMI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

// ------------------------------------------------------------------ mesh
int tmi_read_mesh_file(const char *path, int *dim, int *nlen, int *elen, int *nnode, double *nodes, int *elist,
                       int *node_region)
{
    API_BEGIN
    NEED(path);
    int d = 0;
    std::vector<double> X;
    std::vector<int> E, R;
    read_mesh_file(path, d, X, E, R);
    int n = (int)R.size(), ne = (int)(E.size() / (d + 1));
    if (dim) *dim = d;
    if (nlen) *nlen = n;
    if (elen) *elen = ne;
    if (nnode) *nnode = d + 1;
    if (nodes) memcpy(nodes, X.data(), X.size() * sizeof(double));
    if (elist) memcpy(elist, E.data(), E.size() * sizeof(int));
    if (node_region) memcpy(node_region, R.data(), R.size() * sizeof(int));
    API_END
}

static void make_mesh(int dim, int nlen, const double *nodes, int elen, const int *elist, tmi_mesh *out)
{
    NEED(nodes);
    NEED(elist);
    NEED(out);
    TMI_CHECK(dim == 2 || dim == 3, TMI_ERR_INVALID_ARG, "dim must be 2 or 3");
    TMI_CHECK(nlen > 0 && elen > 0, TMI_ERR_INVALID_ARG, "empty mesh");
    Context &c = Context::get();
    tmi_mesh h = new tmi_mesh_s;
    try {
        Mesh &m = h->m;
        m.dim = dim;
        m.n = nlen;
        m.elen = elen;
        m.nn = dim + 1;
        m.h_nodes.assign(nodes, nodes + (size_t)nlen * dim);
        m.h_elist.assign(elist, elist + (size_t)elen * (dim + 1));
        m.setup(c.stream);
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
}

int tmi_mesh_make(int dim, int nlen, const double *nodes, int elen, const int *elist, tmi_mesh *out)
{
    API_BEGIN
    make_mesh(dim, nlen, nodes, elen, elist, out);
    API_END
}

int tmi_mesh_read(const char *path, tmi_mesh *out)
{
    API_BEGIN
    NEED(path);
    int d = 0;
    std::vector<double> X;
    std::vector<int> E, R;
    read_mesh_file(path, d, X, E, R);
    make_mesh(d, (int)R.size(), X.data(), (int)(E.size() / (d + 1)), E.data(), out);
    API_END
}

int tmi_mesh_free(tmi_mesh m)
{
    API_BEGIN
    if (m) {
        (void)hipStreamSynchronize(Context::get().stream);
        delete m;
    }
    API_END
}

int tmi_mesh_dims(tmi_mesh h, int *dim, int *nlen, int *elen, int *nbnd)
{
    API_BEGIN
    Mesh &m = M(h);
    if (dim) *dim = m.dim;
    if (nlen) *nlen = m.n;
    if (elen) *elen = m.elen;
    if (nbnd) *nbnd = m.nbnd;
    API_END
}

int tmi_mesh_get_nodes(tmi_mesh h, double *nodes)
{
    API_BEGIN
    NEED(nodes);
    Mesh &m = M(h);
    memcpy(nodes, m.h_nodes.data(), m.h_nodes.size() * sizeof(double));
    API_END
}

int tmi_mesh_get_elements(tmi_mesh h, int *elist)
{
    API_BEGIN
    NEED(elist);
    Mesh &m = M(h);
    memcpy(elist, m.h_elist.data(), m.h_elist.size() * sizeof(int));
    API_END
}

int tmi_mesh_get_bnd(tmi_mesh h, int *bnd)
{
    API_BEGIN
    NEED(bnd);
    Mesh &m = M(h);
    memcpy(bnd, m.h_bnd.data(), m.h_bnd.size() * sizeof(int));
    API_END
}

int tmi_mesh_element_size(tmi_mesh h, double *size)
{
    API_BEGIN
    NEED(size);
    Mesh &m = M(h);
    hipStream_t st = Context::get().stream;
    int gs = m.geo_stride(), off = m.dim == 3 ? 16 : 9;
    TMI_HIP(hipMemcpy2DAsync(size, sizeof(double), m.geo.p + off, gs * sizeof(double), sizeof(double), m.elen,
                             hipMemcpyDeviceToHost, st));
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_mesh_sparse_structure(tmi_mesh h, int *rowptr, int *colidx, int *nnz)
{
    API_BEGIN
    Mesh &m = M(h);
    hipStream_t st = Context::get().stream;
    m.build_csr(st);
    if (nnz) *nnz = m.nnz;
    if (rowptr) m.rowptr.download(rowptr, m.n + 1, st);
    if (colidx) m.colidx.download(colidx, m.nnz, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

static void set_qm(Mesh &m, QM qm, bool pull)
{
    if (pull) {
        std::vector<double> none;
        for (int i = 0; i < qm.nm; i++) {
            double out[3];
            if (pull_to_boundary(m, none, &qm.M[(size_t)i * m.dim], out))
                for (int d = 0; d < m.dim; d++) qm.M[(size_t)i * m.dim + d] = out[d];
        }
    }
    m.qm = std::move(qm);
    m.links_valid = false;
}

int tmi_mesh_read_qm(tmi_mesh h, const char *path)
{
    API_BEGIN
    NEED(path);
    Mesh &m = M(h);
    QM qm;
    bool fixed_m = false;
    read_qm_file(path, m.dim, qm, fixed_m);
    set_qm(m, std::move(qm), !fixed_m);
    API_END
}

int tmi_mesh_set_qm(tmi_mesh h, int nq, const double *Q, int nm, const double *Mp, const int *lptr, const int *lidx,
                    int pull_m)
{
    API_BEGIN
    Mesh &m = M(h);
    TMI_CHECK(nq >= 0 && nm >= 0, TMI_ERR_INVALID_ARG, "negative source/detector count");
    NEED(lptr);
    QM qm;
    qm.nq = nq;
    qm.nm = nm;
    if (nq) { NEED(Q); qm.Q.assign(Q, Q + (size_t)nq * m.dim); }
    if (nm) { NEED(Mp); qm.M.assign(Mp, Mp + (size_t)nm * m.dim); }
    qm.lptr.assign(lptr, lptr + nq + 1);
    TMI_CHECK(qm.lptr[0] == 0, TMI_ERR_INVALID_ARG, "link pointer must start at 0");
    for (int q = 0; q < nq; q++)
        TMI_CHECK(qm.lptr[q + 1] >= qm.lptr[q], TMI_ERR_INVALID_ARG, "link pointer not ascending at %d", q);
    qm.nqm = qm.lptr[nq];
    if (qm.nqm) {
        NEED(lidx);
        qm.lidx.assign(lidx, lidx + qm.nqm);
        for (int k = 0; k < qm.nqm; k++)
            TMI_CHECK(qm.lidx[k] >= 0 && qm.lidx[k] < nm, TMI_ERR_INVALID_ARG, "link %d: detector %d out of range", k,
                      qm.lidx[k]);
    }
    set_qm(m, std::move(qm), pull_m != 0);
    API_END
}

int tmi_mesh_qm_dims(tmi_mesh h, int *nq, int *nm, int *nqm)
{
    API_BEGIN
    Mesh &m = M(h);
    if (nq) *nq = m.qm.nq;
    if (nm) *nm = m.qm.nm;
    if (nqm) *nqm = m.qm.nqm;
    API_END
}

int tmi_mesh_get_qm(tmi_mesh h, double *Q, double *Mp, int *lptr, int *lidx)
{
    API_BEGIN
    Mesh &m = M(h);
    if (Q) memcpy(Q, m.qm.Q.data(), m.qm.Q.size() * sizeof(double));
    if (Mp) memcpy(Mp, m.qm.M.data(), m.qm.M.size() * sizeof(double));
    if (lptr) memcpy(lptr, m.qm.lptr.data(), m.qm.lptr.size() * sizeof(int));
    if (lidx) memcpy(lidx, m.qm.lidx.data(), m.qm.lidx.size() * sizeof(int));
    API_END
}

int tmi_qvec_gaussian(tmi_mesh h, int npts, const double *centres, double width, int neumann, const double *scale,
                      double *out)
{
    API_BEGIN
    Mesh &m = M(h);
    NEED(out);
    TMI_CHECK(npts >= 0, TMI_ERR_INVALID_ARG, "negative point count");
    TMI_CHECK(width > 0.0, TMI_ERR_INVALID_ARG, "profile width must be positive (got %g)", width);
    if (npts == 0) return TMI_OK;
    NEED(centres);
    hipStream_t st = Context::get().stream;
    DBuf<double> dc, ds, dout((size_t)npts * m.n);
    dc.upload(centres, (size_t)npts * m.dim, st);
    if (scale) ds.upload(scale, m.n, st);
    launch_qvec_gaussian(m, npts, dc.p, width, neumann, scale ? ds.p : nullptr, dout.p, st);
    dout.download(out, (size_t)npts * m.n, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

static double a_keijzer(double n)   // param.cc:24-32
{
    double th = asin(1.0 / n);
    double costh = fabs(cos(th));
    double R0 = ((n - 1.0) * (n - 1.0)) / ((n + 1.0) * (n + 1.0));
    return (2.0 / (1.0 - R0) - 1.0 + costh * costh * costh) / (1.0 - costh * costh);
}

int tmi_qvec_point(tmi_mesh h, int npts, const double *centres, int neumann, double *out)
{
    API_BEGIN
    NEED(centres); NEED(out);
    Mesh &m = M(h);
    for (int i = 0; i < npts; i++) qvec_point(m, centres + (size_t)i * m.dim, neumann != 0, out + (size_t)i * m.n);
    API_END
}

int tmi_optical_to_coeffs(int nlen, const double *mua, const double *mus, const double *ref, double *cmua,
                          double *ckappa, double *c2a)
{
    API_BEGIN
    NEED(mua); NEED(mus); NEED(ref);
    const double c0 = 0.3;
    for (int i = 0; i < nlen; i++) {
        // A_Keijzer is defined for n >= 1 (n = 1, the index-matched boundary, gives A = 1)
        TMI_CHECK(std::isfinite(mua[i]) && std::isfinite(mus[i]) && std::isfinite(ref[i]), TMI_ERR_INVALID_ARG,
                  "non-finite optical parameter at node %d (mua %g, mus %g, ref %g)", i, mua[i], mus[i], ref[i]);
        TMI_CHECK(ref[i] >= 1.0, TMI_ERR_INVALID_ARG, "refractive index at node %d must be >= 1 (got %g)", i, ref[i]);
        if (cmua) cmua[i] = mua[i] * c0 / ref[i];
        if (ckappa) ckappa[i] = c0 / (3.0 * ref[i] * (mua[i] + mus[i]));
        if (c2a) c2a[i] = c0 / (2.0 * ref[i] * a_keijzer(ref[i]));
    }
    API_END
}

// --------------------------------------------------------------- solver
int tmi_solver_create(tmi_mesh mh, int complex_valued, int lsolver, double tol, tmi_solver *out)
{
    API_BEGIN
    Mesh &m = M(mh);
    NEED(out);
    TMI_CHECK(lsolver >= TMI_LSOLVER_DIRECT && lsolver <= TMI_LSOLVER_COCG, TMI_ERR_INVALID_ARG,
              "unknown linear solver %d", lsolver);
    tmi_solver h = new tmi_solver_s;
    Solver &s = h->s;
    s.mesh = &m;
    s.cplx = complex_valued != 0;
    s.st = Context::get().stream;
    if (lsolver == TMI_LSOLVER_DIRECT) {
        s.lsolver = s.cplx ? TMI_LSOLVER_BICGSTAB : TMI_LSOLVER_CG;
        s.tol = 1e-14;
        s.precon = TMI_PRECON_DIAG;
    } else {
        s.lsolver = lsolver;
        s.tol = tol;
    }
    if (s.cplx && s.lsolver == TMI_LSOLVER_CG) {
        delete h;
        fail(TMI_ERR_UNSUPPORTED, "CG on a complex system is not supported (the reference's complex CG is a stub)");
    }
    if (!s.cplx && s.lsolver == TMI_LSOLVER_COCG) {
        delete h;
        fail(TMI_ERR_UNSUPPORTED, "COCG solves complex symmetric systems; use CG for a real one");
    }
    *out = h;
    API_END
}

int tmi_solver_free(tmi_solver h)
{
    API_BEGIN
    if (h) {
        (void)hipStreamSynchronize(Context::get().stream);
        delete h;
    }
    API_END
}

int tmi_solver_set_linsolver(tmi_solver h, const char *name, double tol)
{
    API_BEGIN
    Solver &s = S(h);
    NEED(name);
    if (!strcasecmp(name, "DIRECT")) {
        s.lsolver = s.cplx ? TMI_LSOLVER_BICGSTAB : TMI_LSOLVER_CG;
        s.tol = 1e-14;
        s.precon = TMI_PRECON_DIAG;
        return TMI_OK;
    }
    int ls;
    if (!strcasecmp(name, "CG")) ls = TMI_LSOLVER_CG;
    else if (!strcasecmp(name, "BICGSTAB")) ls = TMI_LSOLVER_BICGSTAB;
    else if (!strcasecmp(name, "GMRES")) ls = TMI_LSOLVER_GMRES;
    else if (!strcasecmp(name, "COCG")) ls = TMI_LSOLVER_COCG;
    else fail(TMI_ERR_UNSUPPORTED, "linear solver '%s' is not supported (DIRECT, CG, BICGSTAB, GMRES, COCG)", name);
    TMI_CHECK(!(s.cplx && ls == TMI_LSOLVER_CG), TMI_ERR_UNSUPPORTED,
              "CG on a complex system is not supported (the reference's complex CG is a stub)");
    TMI_CHECK(!(!s.cplx && ls == TMI_LSOLVER_COCG), TMI_ERR_UNSUPPORTED,
              "COCG solves complex symmetric systems; use CG for a real one");
    TMI_CHECK(tol > 0.0, TMI_ERR_INVALID_ARG, "tolerance must be positive");
    s.lsolver = ls;
    s.tol = tol;
    API_END
}

int tmi_solver_set_precon(tmi_solver h, int precon)
{
    API_BEGIN
    Solver &s = S(h);
    TMI_CHECK(precon == TMI_PRECON_NONE || precon == TMI_PRECON_DIAG, TMI_ERR_UNSUPPORTED,
              "preconditioner %d not supported (NONE, DIAG)", precon);
    s.precon = precon;
    if (precon == TMI_PRECON_DIAG && s.assembled) s.reset_precon();
    API_END
}

int tmi_solver_set_maxit(tmi_solver h, int maxit)
{
    API_BEGIN
    TMI_CHECK(maxit >= 0, TMI_ERR_INVALID_ARG, "maxit must be >= 0");
    S(h).maxit = maxit;
    API_END
}

int tmi_solver_set_mode(tmi_solver h, int mode)
{
    API_BEGIN
    TMI_CHECK(mode == TMI_MODE_FP64_SINGLE || mode == TMI_MODE_MIXED_BLOCK || mode == TMI_MODE_FP64_BLOCK,
              TMI_ERR_INVALID_ARG, "unknown mode %d", mode);
    S(h).mode = mode;
    API_END
}

int tmi_solver_set_tile(tmi_solver h, int tile)
{
    API_BEGIN
    TMI_CHECK(tile == 0 || tile == 1 || tile == 2 || tile == 4 || tile == 8 || tile == 16 || tile == 32 || tile == 64,
              TMI_ERR_INVALID_ARG, "tile must be 0 (auto) or a power of two <= 64");
    S(h).tile = tile;
    API_END
}

int tmi_solver_set_order(tmi_solver h, int kind)
{
    API_BEGIN
    TMI_CHECK(kind == TMI_ORDER_NATURAL || kind == TMI_ORDER_MORTON || kind == TMI_ORDER_CLUSTER, TMI_ERR_INVALID_ARG,
              "unknown node order %d", kind);
    Solver &s = S(h);
    if (s.order_kind != kind) {
        s.order_kind = kind;
        s.allocated = s.assembled = s.precon_valid = false;   // Allocate builds the solver's copy anew
    }
    API_END
}

int tmi_node_order(int kind, int dim, int n, const double *nodes, const int *rowptr, const int *colidx, int *perm)
{
    API_BEGIN
    TMI_CHECK(kind == TMI_ORDER_NATURAL || kind == TMI_ORDER_MORTON || kind == TMI_ORDER_CLUSTER, TMI_ERR_INVALID_ARG,
              "unknown node order %d", kind);
    TMI_CHECK(n > 0 && (dim == 2 || dim == 3) && perm && rowptr && colidx && (nodes || kind == TMI_ORDER_NATURAL),
              TMI_ERR_INVALID_ARG, "bad node order request");
    TMI_CHECK(rowptr[0] == 0, TMI_ERR_INVALID_ARG, "row pointers must start at 0");
    for (int i = 0; i < n; i++)
        TMI_CHECK(rowptr[i + 1] >= rowptr[i], TMI_ERR_INVALID_ARG, "row pointers must not decrease");
    const std::vector<int> rp(rowptr, rowptr + n + 1), ci(colidx, colidx + rowptr[n]);
    for (int c : ci) TMI_CHECK(c >= 0 && c < n, TMI_ERR_INVALID_ARG, "column index %d outside 0..%d", c, n - 1);
    std::vector<int> p;
    if (kind == TMI_ORDER_NATURAL) {
        p.resize(n);
        for (int i = 0; i < n; i++) p[i] = i;
    } else {
        p = morton_order(dim, n, nodes);
        if (kind == TMI_ORDER_CLUSTER) p = cluster_order(n, rp, ci, p, RB_ROWS);
    }
    std::copy(p.begin(), p.end(), perm);
    API_END
}

int tmi_solver_set_rows_per_wave(tmi_solver h, int rows)
{
    API_BEGIN
    TMI_CHECK(rows >= 0 && rows <= 63, TMI_ERR_INVALID_ARG, "rows per wave must be 0 .. 63");
    S(h).rpw = rows;
    API_END
}

int tmi_solver_set_spmm_variant(tmi_solver h, int variant)
{
    API_BEGIN
    TMI_CHECK(variant >= 0 && variant <= 5, TMI_ERR_INVALID_ARG,
              "SpMM variant must be 0 (LDS, register staging), 1 (direct), 2 (LDS DMA, double-buffered), 3 "
              "(row blocks with shared gathers), 4 (row blocks, software-pipelined) or 5 (row blocks on the "
              "matrix cores)");
    S(h).spmm_variant = variant;
    API_END
}

int tmi_solver_autotune(tmi_solver h, int nq, const void *d_rhs, void *d_phi, int ld, int *tile, int *rows_per_wave,
                        int *spmm_variant, double *solve_ms)
{
    API_BEGIN
    TMI_CHECK(d_rhs && d_phi && nq > 0 && ld >= nq, TMI_ERR_INVALID_ARG, "invalid autotune request");
    S(h).autotune(nq, d_rhs, d_phi, ld, tile, rows_per_wave, spmm_variant, solve_ms);
    API_END
}

int tmi_solver_set_compaction(tmi_solver h, int min_width)
{
    API_BEGIN
    TMI_CHECK(min_width == 0 || min_width == 1 || min_width == 2 || min_width == 4 || min_width == 8 ||
                  min_width == 16 || min_width == 32,
              TMI_ERR_INVALID_ARG, "compaction width must be 0 (off) or a power of two <= 32");
    S(h).compact_min = min_width;
    API_END
}

int tmi_solver_set_fallback(tmi_solver h, int enable, int mixed_maxit)
{
    API_BEGIN
    TMI_CHECK(mixed_maxit >= 0, TMI_ERR_INVALID_ARG, "mixed_maxit must be >= 0");
    Solver &s = S(h);
    s.fp64_fallback = enable != 0;
    s.mixed_maxit = mixed_maxit;
    API_END
}

int tmi_solver_work_bytes(tmi_solver h, int nq, double *bytes)
{
    API_BEGIN
    TMI_CHECK(bytes && nq > 0, TMI_ERR_INVALID_ARG, "invalid argument");
    Solver &s = S(h);
    TMI_CHECK(s.allocated, TMI_ERR_STATE, "work_bytes before Allocate");
    *bytes = (double)s.work_bytes(nq);
    API_END
}

int tmi_solver_allocate(tmi_solver h)
{
    API_BEGIN
    S(h).allocate();
    API_END
}

// FEM coefficients and omega must be finite: a NaN or Inf would reach every
// entry of the rows it touches and every iterate of the solve
static void check_coeffs(const Solver &s, const double *cmua, const double *ckappa, const double *c2a, double omega,
                         bool elbasis = false)
{
    TMI_CHECK(std::isfinite(omega), TMI_ERR_INVALID_ARG, "non-finite omega");
    const int n = elbasis ? s.mesh->elen : s.mesh->n;
    for (int i = 0; i < n; i++)
        TMI_CHECK(std::isfinite(cmua[i]) && std::isfinite(ckappa[i]) && std::isfinite(c2a[i]), TMI_ERR_INVALID_ARG,
                  "non-finite coefficient at %s %d (cmua %g, ckappa %g, c2a %g)", elbasis ? "element" : "node", i,
                  cmua[i], ckappa[i], c2a[i]);
}

int tmi_solver_assemble_basis(tmi_solver h, const double *cmua, const double *ckappa, const double *c2a,
                              double omega, int elbasis)
{
    API_BEGIN
    NEED(cmua); NEED(ckappa); NEED(c2a);
    Solver &s = S(h);
    check_coeffs(s, cmua, ckappa, c2a, omega, elbasis != 0);
    s.assemble(cmua, ckappa, c2a, omega, true, elbasis != 0);
    TMI_HIP(hipStreamSynchronize(s.st));
    API_END
}

int tmi_solver_reset_basis(tmi_solver h, const double *cmua, const double *ckappa, const double *c2a, double omega,
                           int elbasis)
{
    API_BEGIN
    NEED(cmua); NEED(ckappa); NEED(c2a);
    Solver &s = S(h);
    check_coeffs(s, cmua, ckappa, c2a, omega, elbasis != 0);
    s.assemble(cmua, ckappa, c2a, omega, true, elbasis != 0);
    s.reset_precon();
    TMI_HIP(hipStreamSynchronize(s.st));
    API_END
}

int tmi_solver_assemble(tmi_solver h, const double *cmua, const double *ckappa, const double *c2a, double omega)
{
    API_BEGIN
    NEED(cmua); NEED(ckappa); NEED(c2a);
    Solver &s = S(h);
    check_coeffs(s, cmua, ckappa, c2a, omega);
    s.assemble(cmua, ckappa, c2a, omega, true);
    TMI_HIP(hipStreamSynchronize(s.st));
    API_END
}

int tmi_solver_reset(tmi_solver h, const double *cmua, const double *ckappa, const double *c2a, double omega)
{
    API_BEGIN
    NEED(cmua); NEED(ckappa); NEED(c2a);
    Solver &s = S(h);
    check_coeffs(s, cmua, ckappa, c2a, omega);
    s.assemble(cmua, ckappa, c2a, omega, true);
    s.reset_precon();
    TMI_HIP(hipStreamSynchronize(s.st));
    API_END
}

int tmi_solver_nnz(tmi_solver h, int *nnz)
{
    API_BEGIN
    NEED(nnz);
    Solver &s = S(h);
    TMI_CHECK(s.allocated, TMI_ERR_STATE, "solver not allocated");
    *nnz = s.mesh->nnz;
    API_END
}

int tmi_solver_get_sysmat(tmi_solver h, int *rowptr, int *colidx, double *val)
{
    API_BEGIN
    Solver &s = S(h);
    TMI_CHECK(s.assembled, TMI_ERR_STATE, "system matrix not assembled");
    Mesh &m = *s.mesh;
    if (rowptr) m.rowptr.download(rowptr, m.n + 1, s.st);
    if (colidx) m.colidx.download(colidx, m.nnz, s.st);
    if (val) s.val64.download(val, s.nval(), s.st);
    TMI_HIP(hipStreamSynchronize(s.st));
    API_END
}

int tmi_solver_calc_fields(tmi_solver h, int nq, const int *qrp, const int *qci, const double *qval, double *phi,
                           tmi_colstat *stats)
{
    API_BEGIN
    Solver &s = S(h);
    NEED(qrp); NEED(qci); NEED(qval); NEED(phi);
    TMI_CHECK(nq >= 0, TMI_ERR_INVALID_ARG, "negative source count");
    if (nq == 0) return TMI_OK;
    int n = s.mesh->n, cw = s.cplx ? 2 : 1;
    TMI_CHECK(qrp[0] == 0, TMI_ERR_INVALID_ARG, "qvec row pointer must start at 0");
    for (int q = 0; q < nq; q++) {
        TMI_CHECK(qrp[q + 1] >= qrp[q], TMI_ERR_INVALID_ARG, "qvec row pointer not ascending at %d", q);
        for (int k = qrp[q]; k < qrp[q + 1]; k++)
            TMI_CHECK(qci[k] >= 0 && qci[k] < n, TMI_ERR_INVALID_ARG, "qvec column %d out of range", qci[k]);
    }
    // the caller's arrays go over as they are (qvec CSR, phi source-major);
    // the dense node-major right-hand sides and initial guesses are formed,
    // and the fields transposed back, on the GPU
    const size_t tot = (size_t)n * nq * cw, qnnz = (size_t)qrp[nq];
    const int es = 8 * cw;
    DBuf<double> dB(tot), dX(tot), dT(tot), dqv;
    DBuf<int> dqr, dqc;
    dqr.upload(qrp, nq + 1, s.st);
    dqc.upload(qci, qnnz, s.st);
    dqv.upload(qval, qnnz * cw, s.st);
    TMI_HIP(hipMemsetAsync(dB.p, 0, tot * sizeof(double), s.st));
    launch_qvec_dense(nq, cw, dqr.p, dqc.p, dqv.p, dB.p, s.st);
    TMI_HIP(hipMemcpyAsync(dT.p, phi, tot * sizeof(double), hipMemcpyHostToDevice, s.st));
    launch_transpose(es, nq, n, dT.p, dX.p, s.st);   // [nq][n] -> [n][nq]
    int notconv = s.solve(nq, dB.p, dX.p, nq, stats);
    launch_transpose(es, n, nq, dX.p, dT.p, s.st);   // [n][nq] -> [nq][n]
    TMI_HIP(hipMemcpyAsync(phi, dT.p, tot * sizeof(double), hipMemcpyDeviceToHost, s.st));
    TMI_HIP(hipStreamSynchronize(s.st));
    if (notconv) {
        set_last_error(std::to_string(notconv) + " right-hand side(s) did not reach the tolerance");
        return TMI_ERR_NOT_CONVERGED;
    }
    API_END
}

int tmi_solver_calc_fields_device(tmi_solver h, int nq, const void *d_rhs, void *d_phi, tmi_colstat *stats)
{
    API_BEGIN
    Solver &s = S(h);
    NEED(d_rhs); NEED(d_phi);
    TMI_CHECK(nq > 0, TMI_ERR_INVALID_ARG, "source count must be positive");
    size_t bytes = (size_t)s.mesh->n * nq * (s.cplx ? 16 : 8);
    TMI_HIP(hipMemsetAsync(d_phi, 0, bytes, s.st));
    int notconv = s.solve(nq, d_rhs, d_phi, nq, stats);
    if (notconv) {
        set_last_error(std::to_string(notconv) + " right-hand side(s) did not reach the tolerance");
        return TMI_ERR_NOT_CONVERGED;
    }
    API_END
}

int tmi_solver_calc_fields_freqs_device(tmi_solver h, int nq, const void *d_rhs, void *d_phi, const double *omega,
                                        tmi_colstat *stats)
{
    API_BEGIN
    Solver &s = S(h);
    NEED(d_rhs); NEED(d_phi); NEED(omega);
    TMI_CHECK(nq > 0, TMI_ERR_INVALID_ARG, "source count must be positive");
    size_t bytes = (size_t)s.mesh->n * nq * (s.cplx ? 16 : 8);
    TMI_HIP(hipMemsetAsync(d_phi, 0, bytes, s.st));
    int notconv = s.solve_on(s.ex, nq, d_rhs, d_phi, nq, stats, -1, omega);
    if (notconv) {
        set_last_error(std::to_string(notconv) + " right-hand side(s) did not reach the tolerance");
        return TMI_ERR_NOT_CONVERGED;
    }
    API_END
}

// ProjectAll on fields held either on the host ([nq][n], source-major as the
// reference's phi array) or on the device ([n][ld] node-major, block-solve
// layout): sq / sj are the strides of source and node in complex entries
static void project_all(Solver &s, int nm, const int *mrp, const int *mci, const double *mval, const double *d_phi,
                        long sq, long sj, int dscale, double *proj)
{
    Mesh &m = *s.mesh;
    const QM &qm = m.qm;
    TMI_CHECK(nm == qm.nm, TMI_ERR_INVALID_ARG, "mvec has %d rows, the QM description %d detectors", nm, qm.nm);
    int n = m.n, cw = s.cplx ? 2 : 1, nqm = qm.nqm;
    if (nqm == 0) return;
    // link index of every (detector, source) pair, -1 where not linked
    std::vector<int> lk((size_t)nm * qm.nq, -1);
    for (int q = 0; q < qm.nq; q++)
        for (int k = qm.lptr[q]; k < qm.lptr[q + 1]; k++) lk[(size_t)qm.lidx[k] * qm.nq + q] = k;
    TMI_CHECK(mrp[0] == 0, TMI_ERR_INVALID_ARG, "mvec row pointer must start at 0");
    int maxrow = 0;
    for (int i = 0; i < nm; i++) {
        TMI_CHECK(mrp[i + 1] >= mrp[i], TMI_ERR_INVALID_ARG, "mvec row pointer not ascending");
        maxrow = std::max(maxrow, mrp[i + 1] - mrp[i]);
    }
    int mnnz = mrp[nm];
    for (int k = 0; k < mnnz; k++) TMI_CHECK(mci[k] >= 0 && mci[k] < n, TMI_ERR_INVALID_ARG, "mvec column out of range");
    const ProjectPlan pl = project_plan(qm.nq, nm, maxrow, s.cplx ? 1 : 0);
    s.pj_lk.upload(lk.data(), lk.size(), s.st);
    s.pj_rp.upload(mrp, nm + 1, s.st);
    s.pj_ci.upload(mci, mnnz, s.st);
    s.pj_val.upload(mval, (size_t)mnnz * cw, s.st);
    if (s.pj_part.n < pl.partial_doubles) s.pj_part.alloc(pl.partial_doubles);
    if (s.pj_out.n < (size_t)nqm * cw) s.pj_out.alloc((size_t)nqm * cw);
    launch_project(pl, qm.nq, nm, s.pj_lk.p, s.pj_rp.p, s.pj_ci.p, s.pj_val.p, d_phi, sq, sj, s.cplx ? 1 : 0,
                   s.pj_part.p, s.pj_out.p, s.st);
    s.pj_out.download(proj, (size_t)nqm * cw, s.st);
    TMI_HIP(hipStreamSynchronize(s.st));
    if (dscale == TMI_DATA_LOG) {   // complex log: Re = ln|y|, Im = arg y
        for (int k = 0; k < nqm; k++) {
            if (s.cplx) {
                std::complex<double> z = std::log(std::complex<double>(proj[2 * k], proj[2 * k + 1]));
                proj[2 * k] = z.real();
                proj[2 * k + 1] = z.imag();
            } else {
                proj[k] = std::log(proj[k]);
            }
        }
    }
}

int tmi_solver_project_all(tmi_solver h, int nm, const int *mrp, const int *mci, const double *mval,
                           const double *phi, int dscale, double *proj)
{
    API_BEGIN
    Solver &s = S(h);
    NEED(mrp); NEED(mci); NEED(mval); NEED(phi); NEED(proj);
    Mesh &m = *s.mesh;
    DBuf<double> dphi;
    dphi.upload(phi, (size_t)m.qm.nq * m.n * (s.cplx ? 2 : 1), s.st);
    project_all(s, nm, mrp, mci, mval, dphi.p, m.n, 1, dscale, proj);
    API_END
}

int tmi_solver_project_all_device(tmi_solver h, int nm, const int *mrp, const int *mci, const double *mval,
                                  const void *d_phi, int ld, int dscale, double *proj)
{
    API_BEGIN
    Solver &s = S(h);
    NEED(mrp); NEED(mci); NEED(mval); NEED(d_phi); NEED(proj);
    TMI_CHECK(ld >= s.mesh->qm.nq, TMI_ERR_INVALID_ARG, "field row stride %d below the source count %d", ld,
              s.mesh->qm.nq);
    project_all(s, nm, mrp, mci, mval, (const double *)d_phi, 1, ld, dscale, proj);
    API_END
}

int tmi_solver_time_kernel(tmi_solver h, int which, int reps, double *avg_ms, double *bytes)
{
    API_BEGIN
    NEED(avg_ms); NEED(bytes);
    S(h).time_kernel(which, reps, avg_ms, bytes);
    API_END
}

int tmi_solver_last_timing(tmi_solver h, double *solve_ms, double *iters_per_col)
{
    API_BEGIN
    Solver &s = S(h);
    if (solve_ms) *solve_ms = s.ex.last_solve_ms;
    if (iters_per_col) *iters_per_col = s.ex.last_iters;
    API_END
}

int tmi_solver_last_counts(tmi_solver h, long long *kernel_launches, long long *graph_launches,
                           long long *host_syncs)
{
    API_BEGIN
    Solver &s = S(h);
    if (kernel_launches) *kernel_launches = s.ex.last_launch;
    if (graph_launches) *graph_launches = s.ex.last_graph;
    if (host_syncs) *host_syncs = s.ex.last_sync;
    API_END
}

int tmi_solver_set_graphs(tmi_solver h, int enable)
{
    API_BEGIN
    Solver &s = S(h);
    s.use_graphs = enable != 0;
    if (!s.use_graphs) s.ex.drop_graphs();
    API_END
}

}  // extern "C"
