// Jacobian entry points of the C-ABI: GenerateJacobian for the nodal basis
// (jacobian.cc:109-126 with RASTER_NDBASIS) and the Raster_Pixel basis.
#include <climits>
#include <cstring>
#include <vector>

#include <algorithm>

#include "jacobian.hpp"
#include "raster.hpp"
#include "solver.hpp"

#include "handles.hpp"

struct tmi_raster_s {
    tmi::Raster r;
};

namespace tmi {
int mesh_dmax(Mesh &m, hipStream_t st);
}

using namespace tmi;

#define API_BEGIN try {
#define API_END                                   \
    return TMI_OK;                                \
    }                                             \
    catch (const tmi::Error &e)                   \
    {                                             \
        set_last_error(e.what());                 \
        return e.code;                            \
    }                                             \
    catch (const std::exception &e)               \
    {                                             \
        set_last_error(e.what());                 \
        return TMI_ERR_INVALID_ARG;               \
    }

namespace tmi {

// longest CSR row (neighbour count) of the mesh, cached
int mesh_dmax(Mesh &m, hipStream_t st)
{
    m.build_csr(st);
    if (m.dmax > 0) return m.dmax;
    std::vector<int> rp = m.rowptr.to_host(st);
    int d = 0;
    for (int i = 0; i < m.n; i++) d = std::max(d, rp[i + 1] - rp[i]);
    m.dmax = d;
    return d;
}

// device copy of the link list: row pointer, source and detector per link
static void ensure_links(Mesh &m, hipStream_t st)
{
    if (m.links_valid) return;
    const QM &qm = m.qm;
    std::vector<int> lq(qm.nqm), lm(qm.nqm), lk((size_t)qm.nq * qm.nm, -1);
    for (int q = 0; q < qm.nq; q++)
        for (int k = qm.lptr[q]; k < qm.lptr[q + 1]; k++) {
            lq[k] = q;
            lm[k] = qm.lidx[k];
            lk[(size_t)q * qm.nm + qm.lidx[k]] = k;
        }
    m.d_lptr.upload(qm.lptr.data(), qm.lptr.size(), st);
    m.d_lq.upload(lq.data(), lq.size(), st);
    m.d_lm.upload(lm.data(), lm.size(), st);
    m.d_lk.upload(lk.data(), lk.size(), st);
    TMI_HIP(hipStreamSynchronize(st));   // the host vectors go out of scope
    m.links_valid = true;
}

static void jacobian_nodal_dev(Mesh &m, int nq, int nm, const double *d_dphi, int ldd, const double *d_aphi,
                               int lda, const double *d_proj, int dscale, double *d_J, hipStream_t st,
                               bool cw = false, double *d_Jkap = nullptr)
{
    const QM &qm = m.qm;
    TMI_CHECK(qm.nq == nq && qm.nm == nm, TMI_ERR_INVALID_ARG,
              "field counts (%d sources, %d detectors) do not match the QM description (%d, %d)", nq, nm, qm.nq,
              qm.nm);
    TMI_CHECK(qm.nqm > 0, TMI_ERR_INVALID_ARG, "the QM description links no measurements");
    TMI_CHECK(ldd >= nq && lda >= nm, TMI_ERR_INVALID_ARG, "field row strides (%d, %d) below the field counts (%d, %d)",
              ldd, lda, nq, nm);
    TMI_CHECK(dscale == TMI_DATA_LIN || dscale == TMI_DATA_LOG, TMI_ERR_INVALID_ARG, "unknown data scale %d", dscale);
    const int dmax = mesh_dmax(m, st);
    ensure_links(m, st);
    // the node-major intermediate of the two-pass variant
    const size_t ws = jacobian_two_pass() ? (size_t)m.n * qm.nqm * 4 : 0;
    if (m.jac_ws.n < ws) m.jac_ws.alloc(ws);
    launch_jacobian_nodal(m, nq, nm, d_dphi, ldd, d_aphi, lda, qm.nqm, d_proj, dscale == TMI_DATA_LOG ? 1 : 0,
                          m.jac_ws.p, d_J, dmax, st, cw, d_Jkap);
}

// CW inputs on the host: real node-major fields and projections, uploaded and
// widened to the complex layout of the kernels (imaginary parts zero)
struct CwInputs {
    DBuf<double> rd, ra, cd, ca, dp;
    CwInputs(size_t n, int nq, int nm, const double *dphi, const double *aphi, const double *proj, size_t nqm,
             hipStream_t st)
    {
        rd.upload(dphi, n * nq, st);
        ra.upload(aphi, n * nm, st);
        cd.alloc(n * nq * 2);
        ca.alloc(n * nm * 2);
        launch_real_to_cplx((long)n, nq, rd.p, nq, cd.p, st);
        launch_real_to_cplx((long)n, nm, ra.p, nm, ca.p, st);
        if (proj) dp.upload(proj, nqm, st);
    }
};

}  // namespace tmi

extern "C" {

int tmi_jacobian_nodal(tmi_mesh h, int nq, int nm, const double *dphi, const double *aphi, const double *proj,
                       int dscale, double *J)
{
    API_BEGIN
    TMI_CHECK(h && dphi && aphi && proj && J, TMI_ERR_INVALID_ARG, "NULL argument");
    Mesh &m = h->m;
    hipStream_t st = Context::get().stream;
    size_t n = m.n, nqm = m.qm.nqm;
    DBuf<double> dd, da, dp, dJ(4 * nqm * n);
    dd.upload(dphi, n * nq * 2, st);
    da.upload(aphi, n * nm * 2, st);
    dp.upload(proj, nqm * 2, st);
    jacobian_nodal_dev(m, nq, nm, dd.p, nq, da.p, nm, dp.p, dscale, dJ.p, st);
    dJ.download(J, 4 * nqm * n, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_jacobian_nodal_device(tmi_mesh h, int nq, int nm, const void *d_dphi, int ldd, const void *d_aphi, int lda,
                              const void *d_proj, int dscale, void *d_J)
{
    API_BEGIN
    TMI_CHECK(h && d_dphi && d_aphi && d_proj && d_J, TMI_ERR_INVALID_ARG, "NULL argument");
    hipStream_t st = Context::get().stream;
    jacobian_nodal_dev(h->m, nq, nm, (const double *)d_dphi, ldd, (const double *)d_aphi, lda,
                       (const double *)d_proj, dscale, (double *)d_J, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_jacobian_cw(tmi_mesh h, int nq, int nm, const double *dphi, const double *aphi, const double *proj,
                    int dscale, double *Jmua, double *Jkap)
{
    API_BEGIN
    TMI_CHECK(h && dphi && aphi && (Jmua || Jkap), TMI_ERR_INVALID_ARG, "NULL argument");
    TMI_CHECK(dscale == TMI_DATA_LIN || proj, TMI_ERR_INVALID_ARG, "log data scaling needs the projections");
    Mesh &m = h->m;
    hipStream_t st = Context::get().stream;
    const size_t n = m.n, nqm = m.qm.nqm;
    CwInputs in(n, nq, nm, dphi, aphi, proj, nqm, st);
    DBuf<double> jm(Jmua ? nqm * n : 0), jk(Jkap ? nqm * n : 0);
    jacobian_nodal_dev(m, nq, nm, in.cd.p, nq, in.ca.p, nm, in.dp.p, dscale, jm.p, st, true, jk.p);
    if (Jmua) jm.download(Jmua, nqm * n, st);
    if (Jkap) jk.download(Jkap, nqm * n, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

static void raster_create(tmi_mesh h, const int *bdim, const int *gdim, const double *bb, tmi_raster *out)
{
    TMI_CHECK(h && bdim && gdim && out, TMI_ERR_INVALID_ARG, "NULL argument");
    Mesh &m = h->m;
    hipStream_t st = Context::get().stream;
    tmi_raster r = new tmi_raster_s;
    try {
        Raster &R = r->r;
        R.mesh = &m;
        R.dim = m.dim;
        R.glen = 1;
        R.blen = 1;
        for (int d = 0; d < m.dim; d++) {
            TMI_CHECK(gdim[d] >= 2 && bdim[d] >= 2, TMI_ERR_INVALID_ARG, "grid dimensions must be >= 2 (got %d, %d)",
                      gdim[d], bdim[d]);
            R.gdim[d] = gdim[d];
            R.bdim[d] = bdim[d];
            R.glen *= gdim[d];
            R.blen *= bdim[d];
            if (bdim[d] != gdim[d]) R.has_g = true;
        }
        TMI_CHECK(R.glen < INT_MAX && R.blen < INT_MAX, TMI_ERR_UNSUPPORTED, "grid too large");
        if (bb) {
            for (int d = 0; d < m.dim; d++) { R.bbmin[d] = bb[d]; R.bbmax[d] = bb[m.dim + d]; }
        } else {   // Mesh::BoundingBox with pad 0
            for (int d = 0; d < m.dim; d++) { R.bbmin[d] = 1e300; R.bbmax[d] = -1e300; }
            for (int i = 0; i < m.n; i++)
                for (int d = 0; d < m.dim; d++) {
                    const double v = m.h_nodes[(size_t)i * m.dim + d];
                    R.bbmin[d] = std::min(R.bbmin[d], v);
                    R.bbmax[d] = std::max(R.bbmax[d], v);
                }
        }
        for (int d = 0; d < m.dim; d++) {
            TMI_CHECK(R.bbmax[d] > R.bbmin[d], TMI_ERR_INVALID_ARG, "invalid bounding box");
            R.dW[d] = (R.bbmax[d] - R.bbmin[d]) / (double)(R.gdim[d] - 1);
        }
        R.build(st);
    } catch (...) {
        delete r;
        throw;
    }
    *out = r;
}

int tmi_raster_create(tmi_mesh h, const int *gdim, const double *bb, tmi_raster *out)
{
    API_BEGIN
    raster_create(h, gdim, gdim, bb, out);
    API_END
}

int tmi_raster_create_basis(tmi_mesh h, const int *bdim, const int *gdim, const double *bb, tmi_raster *out)
{
    API_BEGIN
    raster_create(h, bdim, gdim, bb, out);
    API_END
}

int tmi_raster_basis_len(tmi_raster r, int *blen, int *has_g)
{
    API_BEGIN
    TMI_CHECK(r, TMI_ERR_INVALID_ARG, "invalid raster handle");
    if (blen) *blen = (int)r->r.blen;
    if (has_g) *has_g = r->r.has_g ? 1 : 0;
    API_END
}

// CSR export: rowptr (rows+1) always; colidx / val when given
static void export_csr(const DBuf<int> &rp, const DBuf<int> &ci, const DBuf<double> &val, long rows, int *rowptr,
                       int *colidx, double *v, int *nnz, hipStream_t st)
{
    int nz = 0;
    TMI_HIP(hipMemcpyAsync(&nz, rp.p + rows, sizeof(int), hipMemcpyDeviceToHost, st));
    TMI_HIP(hipStreamSynchronize(st));
    if (nnz) *nnz = nz;
    if (rowptr) rp.download(rowptr, rows + 1, st);
    if (colidx) ci.download(colidx, nz, st);
    if (v) val.download(v, nz, st);
    TMI_HIP(hipStreamSynchronize(st));
}

int tmi_raster_get_gmap(tmi_raster r, int *rowptr, int *colidx, double *val, int *nnz)
{
    API_BEGIN
    TMI_CHECK(r, TMI_ERR_INVALID_ARG, "invalid raster handle");
    Raster &R = r->r;
    TMI_CHECK(R.has_g, TMI_ERR_UNSUPPORTED, "basis grid equals the intermediate grid: G is the identity");
    export_csr(R.g_rp, R.g_ci, R.g_val, R.blen, rowptr, colidx, val, nnz, Context::get().stream);
    API_END
}

int tmi_raster_get_gimap(tmi_raster r, int *rowptr, int *colidx, double *val, int *nnz)
{
    API_BEGIN
    TMI_CHECK(r, TMI_ERR_INVALID_ARG, "invalid raster handle");
    Raster &R = r->r;
    TMI_CHECK(R.has_g, TMI_ERR_UNSUPPORTED, "basis grid equals the intermediate grid: GI is the identity");
    hipStream_t st = Context::get().stream;
    R.build_gimap(st);
    export_csr(R.gi_rp, R.gi_ci, R.gi_val, R.glen, rowptr, colidx, val, nnz, st);
    API_END
}

int tmi_raster_map_grid_to_basis(tmi_raster r, int ncols, const double *gvec, double *bvec)
{
    API_BEGIN
    TMI_CHECK(r && gvec && bvec && ncols > 0, TMI_ERR_INVALID_ARG, "invalid argument");
    hipStream_t st = Context::get().stream;
    Raster &R = r->r;
    DBuf<double> dg, db((size_t)R.blen * ncols * 2);
    dg.upload(gvec, (size_t)R.glen * ncols * 2, st);
    launch_grid_to_basis(R, ncols, dg.p, db.p, st);
    db.download(bvec, (size_t)R.blen * ncols * 2, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_raster_map_basis_to_grid(tmi_raster r, int ncols, const double *bvec, double *gvec)
{
    API_BEGIN
    TMI_CHECK(r && gvec && bvec && ncols > 0, TMI_ERR_INVALID_ARG, "invalid argument");
    hipStream_t st = Context::get().stream;
    Raster &R = r->r;
    DBuf<double> db, dg((size_t)R.glen * ncols * 2);
    db.upload(bvec, (size_t)R.blen * ncols * 2, st);
    launch_basis_to_grid(R, ncols, db.p, dg.p, st);
    dg.download(gvec, (size_t)R.glen * ncols * 2, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_raster_free(tmi_raster r)
{
    API_BEGIN
    if (r) {
        (void)hipStreamSynchronize(Context::get().stream);
        delete r;
    }
    API_END
}

int tmi_raster_dims(tmi_raster r, int *glen, int *slen)
{
    API_BEGIN
    TMI_CHECK(r, TMI_ERR_INVALID_ARG, "invalid raster handle");
    if (glen) *glen = (int)r->r.glen;
    if (slen) *slen = r->r.slen;
    API_END
}

int tmi_raster_bbox(tmi_raster r, double *bbmin, double *bbmax)
{
    API_BEGIN
    TMI_CHECK(r, TMI_ERR_INVALID_ARG, "invalid raster handle");
    for (int d = 0; d < r->r.dim; d++) {
        if (bbmin) bbmin[d] = r->r.bbmin[d];
        if (bbmax) bbmax[d] = r->r.bbmax[d];
    }
    API_END
}

int tmi_raster_get_elref(tmi_raster r, int *elref, int *sol2basis)
{
    API_BEGIN
    TMI_CHECK(r, TMI_ERR_INVALID_ARG, "invalid raster handle");
    hipStream_t st = Context::get().stream;
    if (elref) r->r.elref.download(elref, r->r.glen, st);
    if (sol2basis) r->r.sol2basis.download(sol2basis, r->r.slen, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_raster_get_bmap(tmi_raster r, int *cols, double *w)
{
    API_BEGIN
    TMI_CHECK(r, TMI_ERR_INVALID_ARG, "invalid raster handle");
    hipStream_t st = Context::get().stream;
    const size_t cnt = (size_t)r->r.glen * (r->r.dim + 1);
    if (cols) r->r.bcols.download(cols, cnt, st);
    if (w) r->r.bw.download(w, cnt, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_raster_get_nodemap(tmi_raster r, int *cols, double *w)
{
    API_BEGIN
    TMI_CHECK(r, TMI_ERR_INVALID_ARG, "invalid raster handle");
    hipStream_t st = Context::get().stream;
    Raster &R = r->r;
    R.build_nodemap(st);
    const size_t cnt = (size_t)R.mesh->n * (1 << R.dim);
    if (cols) R.ncols.download(cols, cnt, st);
    if (w) R.nw.download(w, cnt, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_raster_map_grid_to_mesh(tmi_raster r, int ncols, const double *gvec, double *mvec)
{
    API_BEGIN
    TMI_CHECK(r && gvec && mvec && ncols > 0, TMI_ERR_INVALID_ARG, "invalid argument");
    hipStream_t st = Context::get().stream;
    Raster &R = r->r;
    DBuf<double> dg, dm((size_t)R.mesh->n * ncols * 2);
    dg.upload(gvec, (size_t)R.glen * ncols * 2, st);
    launch_grid_to_mesh(R, ncols, dg.p, dm.p, st);
    dm.download(mvec, (size_t)R.mesh->n * ncols * 2, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_raster_map_mesh_to_grid(tmi_raster r, int ncols, const double *mvec, double *gvec)
{
    API_BEGIN
    TMI_CHECK(r && mvec && gvec && ncols > 0, TMI_ERR_INVALID_ARG, "invalid argument");
    hipStream_t st = Context::get().stream;
    const Raster &R = r->r;
    DBuf<double> df, dg((size_t)R.glen * ncols * 2);
    df.upload(mvec, (size_t)R.mesh->n * ncols * 2, st);
    launch_raster_map(R, ncols, df.p, ncols, dg.p, st);
    dg.download(gvec, (size_t)R.glen * ncols * 2, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

static void jacobian_raster_dev(Raster &R, int nq, int nm, const double *d_dphi, int ldd, const double *d_aphi,
                                int lda, const double *d_proj, int dscale, double *d_J, hipStream_t st,
                                bool cw = false, double *d_Jkap = nullptr)
{
    Mesh &m = *R.mesh;
    const QM &qm = m.qm;
    TMI_CHECK(qm.nq == nq && qm.nm == nm, TMI_ERR_INVALID_ARG,
              "field counts (%d sources, %d detectors) do not match the QM description (%d, %d)", nq, nm, qm.nq,
              qm.nm);
    TMI_CHECK(qm.nqm > 0, TMI_ERR_INVALID_ARG, "the QM description links no measurements");
    TMI_CHECK(ldd >= nq && lda >= nm, TMI_ERR_INVALID_ARG, "field row strides below the field counts");
    TMI_CHECK(dscale == TMI_DATA_LIN || dscale == TMI_DATA_LOG, TMI_ERR_INVALID_ARG, "unknown data scale %d", dscale);
    ensure_links(m, st);
    launch_jacobian_grid(R, nq, nm, d_dphi, ldd, d_aphi, lda, m.d_lk.p, qm.nqm, d_proj,
                         dscale == TMI_DATA_LOG ? 1 : 0, d_J, st, cw, d_Jkap);
}

int tmi_jacobian_raster(tmi_raster r, int nq, int nm, const double *dphi, const double *aphi, const double *proj,
                        int dscale, double *J)
{
    API_BEGIN
    TMI_CHECK(r && dphi && aphi && proj && J, TMI_ERR_INVALID_ARG, "NULL argument");
    Raster &R = r->r;
    hipStream_t st = Context::get().stream;
    const size_t n = R.mesh->n, nqm = R.mesh->qm.nqm;
    DBuf<double> dd, da, dp, dJ(4 * nqm * (size_t)R.slen);
    dd.upload(dphi, n * nq * 2, st);
    da.upload(aphi, n * nm * 2, st);
    dp.upload(proj, nqm * 2, st);
    jacobian_raster_dev(R, nq, nm, dd.p, nq, da.p, nm, dp.p, dscale, dJ.p, st);
    dJ.download(J, 4 * nqm * (size_t)R.slen, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_jacobian_raster_cw(tmi_raster r, int nq, int nm, const double *dphi, const double *aphi, const double *proj,
                           int dscale, double *Jmua, double *Jkap)
{
    API_BEGIN
    TMI_CHECK(r && dphi && aphi && (Jmua || Jkap), TMI_ERR_INVALID_ARG, "NULL argument");
    TMI_CHECK(dscale == TMI_DATA_LIN || proj, TMI_ERR_INVALID_ARG, "log data scaling needs the projections");
    Raster &R = r->r;
    hipStream_t st = Context::get().stream;
    const size_t n = R.mesh->n, nqm = R.mesh->qm.nqm, slen = (size_t)R.slen;
    CwInputs in(n, nq, nm, dphi, aphi, proj, nqm, st);
    DBuf<double> jm(Jmua ? nqm * slen : 0), jk(Jkap ? nqm * slen : 0);
    jacobian_raster_dev(R, nq, nm, in.cd.p, nq, in.ca.p, nm, in.dp.p, dscale, jm.p, st, true, jk.p);
    if (Jmua) jm.download(Jmua, nqm * slen, st);
    if (Jkap) jk.download(Jkap, nqm * slen, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_raster_data_gradient(tmi_raster r, int nq, const double *dphi, const double *aphi, double *grad)
{
    API_BEGIN
    TMI_CHECK(r && dphi && aphi && grad && nq > 0, TMI_ERR_INVALID_ARG, "NULL argument");
    Raster &R = r->r;
    hipStream_t st = Context::get().stream;
    const size_t n = R.mesh->n, slen = (size_t)R.slen;
    DBuf<double> dd, da, dg;
    dd.upload(dphi, n * nq * 2, st);
    da.upload(aphi, n * nq * 2, st);
    dg.upload(grad, 2 * slen, st);
    launch_data_gradient(R, nq, dd.p, nq, da.p, nq, dg.p, st);
    dg.download(grad, 2 * slen, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_raster_data_gradient_device(tmi_raster r, int nq, const void *d_dphi, int ldd, const void *d_aphi, int lda,
                                    void *d_grad)
{
    API_BEGIN
    TMI_CHECK(r && d_dphi && d_aphi && d_grad && nq > 0 && ldd >= nq && lda >= nq, TMI_ERR_INVALID_ARG,
              "invalid argument");
    hipStream_t st = Context::get().stream;
    launch_data_gradient(r->r, nq, (const double *)d_dphi, ldd, (const double *)d_aphi, lda, (double *)d_grad, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_jacobian_raster_device(tmi_raster r, int nq, int nm, const void *d_dphi, int ldd, const void *d_aphi, int lda,
                               const void *d_proj, int dscale, void *d_J)
{
    API_BEGIN
    TMI_CHECK(r && d_dphi && d_aphi && d_proj && d_J, TMI_ERR_INVALID_ARG, "NULL argument");
    hipStream_t st = Context::get().stream;
    jacobian_raster_dev(r->r, nq, nm, (const double *)d_dphi, ldd, (const double *)d_aphi, lda,
                        (const double *)d_proj, dscale, (double *)d_J, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

// ------------------------------------------------------------ element basis
static void jacobian_el_dev(Mesh &m, int nq, int nm, const double *d_dphi, int ldd, const double *d_aphi, int lda,
                            const double *d_proj, int dscale, double *d_J, hipStream_t st, bool cw = false,
                            double *d_Jkap = nullptr, int mvec_n = 0)
{
    const QM &qm = m.qm;
    TMI_CHECK(qm.nq == nq && qm.nm == nm, TMI_ERR_INVALID_ARG,
              "field counts (%d sources, %d detectors) do not match the QM description (%d, %d)", nq, nm, qm.nq,
              qm.nm);
    TMI_CHECK(qm.nqm > 0, TMI_ERR_INVALID_ARG, "the QM description links no measurements");
    TMI_CHECK(ldd >= nq && lda >= nm, TMI_ERR_INVALID_ARG, "field row strides (%d, %d) below the field counts (%d, %d)",
              ldd, lda, nq, nm);
    TMI_CHECK(dscale == TMI_DATA_LIN || dscale == TMI_DATA_LOG, TMI_ERR_INVALID_ARG, "unknown data scale %d", dscale);
    TMI_CHECK(nq <= 65535, TMI_ERR_UNSUPPORTED, "element-basis Jacobian: at most 65535 sources");
    ensure_links(m, st);
    launch_jacobian_el(m, nq, d_dphi, ldd, d_aphi, lda, m.d_lptr.p, m.d_lm.p, qm.nqm, d_proj,
                       dscale == TMI_DATA_LOG ? 1 : 0, mvec_n, d_J, st, cw, d_Jkap);
}

// linear projections from the measurement vectors (ProjectSingle / ProjectAll
// with DATA_LIN, fwdsolver.cc:1415-1447): d_phi [n][nq] node-major (complex
// or real), mvec nm x n CSR on the host; d_proj gets nqm values in link order
static void project_from_mvec(Mesh &m, int nm, const int *mrp, const int *mci, const double *mval,
                              const double *d_phi, int cplx, DBuf<double> &d_proj, hipStream_t st)
{
    const QM &qm = m.qm;
    TMI_CHECK(nm == qm.nm, TMI_ERR_INVALID_ARG, "mvec has %d rows, the QM description %d detectors", nm, qm.nm);
    TMI_CHECK(mrp[0] == 0, TMI_ERR_INVALID_ARG, "mvec row pointer must start at 0");
    int maxrow = 0;
    for (int i = 0; i < nm; i++) {
        TMI_CHECK(mrp[i + 1] >= mrp[i], TMI_ERR_INVALID_ARG, "mvec row pointer not ascending");
        maxrow = std::max(maxrow, mrp[i + 1] - mrp[i]);
    }
    const int mnnz = mrp[nm];
    for (int k = 0; k < mnnz; k++)
        TMI_CHECK(mci[k] >= 0 && mci[k] < m.n, TMI_ERR_INVALID_ARG, "mvec column out of range");
    std::vector<int> lk((size_t)nm * qm.nq, -1);   // link of (detector, source), -1 if not linked
    for (int q = 0; q < qm.nq; q++)
        for (int k = qm.lptr[q]; k < qm.lptr[q + 1]; k++) lk[(size_t)qm.lidx[k] * qm.nq + q] = k;
    const int cw = cplx ? 2 : 1;
    const ProjectPlan pl = project_plan(qm.nq, nm, maxrow, cplx);
    DBuf<int> dlk, drp, dci;
    DBuf<double> dval, dpart(pl.partial_doubles);
    dlk.upload(lk.data(), lk.size(), st);
    drp.upload(mrp, nm + 1, st);
    dci.upload(mci, mnnz, st);
    dval.upload(mval, (size_t)mnnz * cw, st);
    d_proj.alloc((size_t)qm.nqm * cw);
    launch_project(pl, qm.nq, nm, dlk.p, drp.p, dci.p, dval.p, d_phi, 1, qm.nq, cplx, dpart.p, d_proj.p, st);
    TMI_HIP(hipStreamSynchronize(st));   // the uploads above go out of scope
}

static size_t basis_len(Mesh &m, tmi_raster r, int basis)
{
    switch (basis) {
    case TMI_BASIS_NODAL: return m.n;
    case TMI_BASIS_ELEMENT: return m.elen;
    case TMI_BASIS_RASTER:
        TMI_CHECK(r && r->r.mesh == &m, TMI_ERR_INVALID_ARG, "the raster basis needs a raster of this mesh");
        return (size_t)r->r.slen;
    }
    fail(TMI_ERR_INVALID_ARG, "unknown basis %d", basis);
}

int tmi_jacobian_element(tmi_mesh h, int nq, int nm, const double *dphi, const double *aphi, const double *proj,
                         int dscale, double *J)
{
    API_BEGIN
    TMI_CHECK(h && dphi && aphi && proj && J, TMI_ERR_INVALID_ARG, "NULL argument");
    Mesh &m = h->m;
    hipStream_t st = Context::get().stream;
    size_t n = m.n, nqm = m.qm.nqm, E = m.elen;
    DBuf<double> dd, da, dp, dJ(4 * nqm * E);
    dd.upload(dphi, n * nq * 2, st);
    da.upload(aphi, n * nm * 2, st);
    dp.upload(proj, nqm * 2, st);
    jacobian_el_dev(m, nq, nm, dd.p, nq, da.p, nm, dp.p, dscale, dJ.p, st);
    dJ.download(J, 4 * nqm * E, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_jacobian_element_device(tmi_mesh h, int nq, int nm, const void *d_dphi, int ldd, const void *d_aphi, int lda,
                                const void *d_proj, int dscale, void *d_J)
{
    API_BEGIN
    TMI_CHECK(h && d_dphi && d_aphi && d_proj && d_J, TMI_ERR_INVALID_ARG, "NULL argument");
    hipStream_t st = Context::get().stream;
    jacobian_el_dev(h->m, nq, nm, (const double *)d_dphi, ldd, (const double *)d_aphi, lda, (const double *)d_proj,
                    dscale, (double *)d_J, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_jacobian_cw_element(tmi_mesh h, int nq, int nm, const double *dphi, const double *aphi, const double *proj,
                            int dscale, double *Jmua, double *Jkap)
{
    API_BEGIN
    TMI_CHECK(h && dphi && aphi && (Jmua || Jkap), TMI_ERR_INVALID_ARG, "NULL argument");
    TMI_CHECK(dscale == TMI_DATA_LIN || proj, TMI_ERR_INVALID_ARG, "log data scaling needs the projections");
    Mesh &m = h->m;
    hipStream_t st = Context::get().stream;
    const size_t n = m.n, nqm = m.qm.nqm, E = m.elen;
    CwInputs in(n, nq, nm, dphi, aphi, proj, nqm, st);
    DBuf<double> jm(Jmua ? nqm * E : 0), jk(Jkap ? nqm * E : 0);
    jacobian_el_dev(m, nq, nm, in.cd.p, nq, in.ca.p, nm, in.dp.p, dscale, jm.p, st, true, jk.p);
    if (Jmua) jm.download(Jmua, nqm * E, st);
    if (Jkap) jk.download(Jkap, nqm * E, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_jacobian_mvec(tmi_mesh h, tmi_raster r, int basis, int nq, int nm, const int *mrp, const int *mci,
                      const double *mval, const double *dphi, const double *aphi, int dscale, double *J)
{
    API_BEGIN
    TMI_CHECK(h && mrp && mci && mval && dphi && aphi && J, TMI_ERR_INVALID_ARG, "NULL argument");
    Mesh &m = h->m;
    hipStream_t st = Context::get().stream;
    const size_t n = m.n, nqm = m.qm.nqm, len = basis_len(m, r, basis);
    TMI_CHECK(m.qm.nq == nq, TMI_ERR_INVALID_ARG, "%d sources, the QM description %d", nq, m.qm.nq);
    DBuf<double> dd, da, dp, dJ(4 * nqm * len);
    dd.upload(dphi, n * nq * 2, st);
    da.upload(aphi, n * nm * 2, st);
    project_from_mvec(m, nm, mrp, mci, mval, dd.p, 1, dp, st);
    if (basis == TMI_BASIS_NODAL) jacobian_nodal_dev(m, nq, nm, dd.p, nq, da.p, nm, dp.p, dscale, dJ.p, st);
    else if (basis == TMI_BASIS_ELEMENT)
        jacobian_el_dev(m, nq, nm, dd.p, nq, da.p, nm, dp.p, dscale, dJ.p, st, false, nullptr, m.elen > m.n ? m.n : 0);
    else jacobian_raster_dev(r->r, nq, nm, dd.p, nq, da.p, nm, dp.p, dscale, dJ.p, st);
    dJ.download(J, 4 * nqm * len, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

int tmi_jacobian_cw_mvec(tmi_mesh h, tmi_raster r, int basis, int nq, int nm, const int *mrp, const int *mci,
                         const double *mval, const double *dphi, const double *aphi, int dscale, double *Jmua,
                         double *Jkap)
{
    API_BEGIN
    TMI_CHECK(h && mrp && mci && mval && dphi && aphi && (Jmua || Jkap), TMI_ERR_INVALID_ARG, "NULL argument");
    Mesh &m = h->m;
    hipStream_t st = Context::get().stream;
    const size_t n = m.n, nqm = m.qm.nqm, len = basis_len(m, r, basis);
    TMI_CHECK(m.qm.nq == nq, TMI_ERR_INVALID_ARG, "%d sources, the QM description %d", nq, m.qm.nq);
    CwInputs in(n, nq, nm, dphi, aphi, nullptr, nqm, st);
    DBuf<double> dp;
    project_from_mvec(m, nm, mrp, mci, mval, in.rd.p, 0, dp, st);
    DBuf<double> jm(Jmua ? nqm * len : 0), jk(Jkap ? nqm * len : 0);
    if (basis == TMI_BASIS_NODAL)
        jacobian_nodal_dev(m, nq, nm, in.cd.p, nq, in.ca.p, nm, dp.p, dscale, jm.p, st, true, jk.p);
    else if (basis == TMI_BASIS_ELEMENT)
        jacobian_el_dev(m, nq, nm, in.cd.p, nq, in.ca.p, nm, dp.p, dscale, jm.p, st, true, jk.p);
    else jacobian_raster_dev(r->r, nq, nm, in.cd.p, nq, in.ca.p, nm, dp.p, dscale, jm.p, st, true, jk.p);
    if (Jmua) jm.download(Jmua, nqm * len, st);
    if (Jkap) jk.download(Jkap, nqm * len, st);
    TMI_HIP(hipStreamSynchronize(st));
    API_END
}

}  // extern "C"
