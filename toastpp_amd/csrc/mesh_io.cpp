// Host-side mesh and QM file handling.
//   "MeshData 5.0" reader: mesh.cc:1787-1825, ndlist.cc:146-159,
//       node.cc:130-162, ellist.cc:154-190, element.cc:429-437
//   QM reader: QMMesh::LoadQM (qmmesh.cc:319-466)
//   measurement pulling: QMMesh::InitM (qmmesh.cc:105-145) ->
//       Mesh::PullToBoundary (mesh.cc:1052-1111)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "fe_integrals.hpp"
#include "mesh.hpp"

namespace tmi {

void read_mesh_file(const char *path, int &dim, std::vector<double> &nodes, std::vector<int> &elist,
                    std::vector<int> &region)
{
    std::ifstream f(path);
    TMI_CHECK(f.good(), TMI_ERR_IO, "cannot open mesh file '%s'", path);
    std::string line;
    bool found = false;
    while (std::getline(f, line))
        if (!line.compare(0, 12, "MeshData 5.0")) { found = true; break; }
    TMI_CHECK(found, TMI_ERR_IO, "'%s' is not a MeshData 5.0 file", path);
    int n = -1;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string tag;
        is >> tag;
        if (tag == "NodeList") { is >> n; break; }
    }
    TMI_CHECK(n > 0, TMI_ERR_IO, "'%s': NodeList not found", path);
    nodes.clear();
    region.assign(n, -1);
    dim = 0;
    std::vector<double> crd;
    for (int i = 0; i < n; i++) {
        TMI_CHECK((bool)std::getline(f, line), TMI_ERR_IO, "'%s': truncated node list", path);
        size_t lb = line.find('['), rb = line.find(']');
        TMI_CHECK(lb != std::string::npos && rb != std::string::npos, TMI_ERR_IO, "'%s': bad node line %d", path, i);
        std::istringstream is(line.substr(lb + 1, rb - lb - 1));
        double v;
        int d = 0;
        double c3[3];
        while (d < 3 && is >> v) c3[d++] = v;
        if (i == 0) dim = d;
        TMI_CHECK(d == dim && (dim == 2 || dim == 3), TMI_ERR_IO, "'%s': node %d has %d coordinates", path, i, d);
        for (int k = 0; k < dim; k++) nodes.push_back(c3[k]);
        if (rb + 1 < line.size() && line[rb + 1] == 'R') region[i] = atoi(line.c_str() + rb + 2);
    }
    int elen = -1;
    while (std::getline(f, line))
        if (!strncasecmp(line.c_str(), "ElementList", 11)) { elen = atoi(line.c_str() + 11); break; }
    TMI_CHECK(elen > 0, TMI_ERR_IO, "'%s': ElementList not found", path);
    int nn = dim + 1;
    elist.assign((size_t)elen * nn, 0);
    int want = dim == 3 ? TMI_ELID_TET4 : TMI_ELID_TRI3;
    for (int e = 0; e < elen; e++) {
        char c = ' ';
        do { f >> c; } while (f && c == ' ');
        TMI_CHECK((bool)f, TMI_ERR_IO, "'%s': truncated element list", path);
        int id = (c >= 'a' && c <= 'z') ? 1 + (c - 'a') : 1 + (c - '1');
        TMI_CHECK(id == want, TMI_ERR_UNSUPPORTED, "'%s': element %d has type %d; only Tri3 (2D) and Tet4 (3D) are supported",
                  path, e, id);
        for (int k = 0; k < nn; k++) {
            int v;
            f >> v;
            elist[(size_t)e * nn + k] = v - 1;
        }
        std::getline(f, line);
    }
}

void read_qm_file(const char *path, int dim, QM &qm, bool &fixed_m)
{
    std::ifstream f(path);
    TMI_CHECK(f.good(), TMI_ERR_IO, "cannot open QM file '%s'", path);
    std::string line;
    std::getline(f, line);
    TMI_CHECK(!line.compare(0, 7, "QM file"), TMI_ERR_IO, "'%s' is not a QM file", path);
    std::getline(f, line);
    TMI_CHECK(!line.compare(0, 9, "Dimension"), TMI_ERR_IO, "'%s': unknown QM format", path);
    int qdim = atoi(line.c_str() + 9);
    TMI_CHECK(qdim == dim, TMI_ERR_INVALID_ARG, "QM dimension %d does not match mesh dimension %d", qdim, dim);
    auto read_list = [&](const char *tag, int &count, std::vector<double> &pts, bool &fixed) {
        size_t tl = strlen(tag);
        bool ok = false;
        while (std::getline(f, line))
            if (!line.compare(0, tl, tag)) { ok = true; break; }
        TMI_CHECK(ok, TMI_ERR_IO, "'%s': %s not found", path, tag);
        std::istringstream is(line.substr(tl));
        std::string flag;
        is >> count >> flag;
        fixed = !strcasecmp(flag.c_str(), "fixed");
        pts.assign((size_t)count * dim, 0.0);
        for (int i = 0; i < count; i++) {
            TMI_CHECK((bool)std::getline(f, line), TMI_ERR_IO, "'%s': truncated %s", path, tag);
            std::istringstream ps(line);
            for (int d = 0; d < dim; d++) ps >> pts[(size_t)i * dim + d];
        }
    };
    bool fixed_q = false;
    read_list("SourceList", qm.nq, qm.Q, fixed_q);
    read_list("MeasurementList", qm.nm, qm.M, fixed_m);
    bool ok = false;
    while (std::getline(f, line))
        if (!line.compare(0, 8, "LinkList")) { ok = true; break; }
    TMI_CHECK(ok, TMI_ERR_IO, "'%s': LinkList not found", path);
    qm.lptr.assign(qm.nq + 1, 0);
    qm.lidx.clear();
    for (int q = 0; q < qm.nq; q++) {
        int cnt;
        f >> cnt;
        char c;
        do { f.get(c); } while (f && c != ':');
        for (int j = 0; j < cnt; j++) {
            int m;
            f >> m;
            TMI_CHECK(m >= 0 && m < qm.nm, TMI_ERR_IO, "'%s': link %d of source %d out of range", path, m, q);
            qm.lidx.push_back(m);
        }
        qm.lptr[q + 1] = qm.lptr[q] + cnt;
    }
    qm.nqm = qm.lptr[qm.nq];
}

// Local coordinates of a global point in element e (tet4.cc:186-197,
// tri3.cc:134-144); g: the element's geometry record.
static void el_local(int dim, const double *g, const double *p, double *loc)
{
    if (dim == 3) {
        const double *a = g, *b = g + 4, *c = g + 8, *d = g + 12;
        double scale = 1.0 / (a[0] + a[1] + a[2] + a[3]);
        loc[0] = (a[1] + b[1] * p[0] + c[1] * p[1] + d[1] * p[2]) * scale;
        loc[1] = (a[2] + b[2] * p[0] + c[2] * p[1] + d[2] * p[2]) * scale;
        loc[2] = (a[3] + b[3] * p[0] + c[3] * p[1] + d[3] * p[2]) * scale;
    } else {
        const double *a = g, *b = g + 3, *c = g + 6;
        double scale = 1.0 / (a[0] + a[1] + a[2]);
        loc[0] = (a[1] + b[1] * p[0] + c[1] * p[1]) * scale;
        loc[1] = (a[2] + b[2] * p[0] + c[2] * p[1]) * scale;
    }
}

bool pull_to_boundary(const Mesh &m, const std::vector<double> &, const double *p, double *out)
{
    // local side normals, including the reference's truncated literals
    // (tet4.cc:244-256, tri3.cc:165-176)
    static const double lnm3[4][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {0.577350269, 0.577350269, 0.577350269}};
    static const double lnm2[3][2] = {{0, -1}, {0.7071067814, 0.7071067814}, {-1, 0}};
    int dim = m.dim, nn = m.nn, nmin = -1;
    double dmin = 1e10;
    const double *X = m.h_nodes.data();
    for (int i = 0; i < m.n; i++) {
        if (!m.h_bnd[i]) continue;
        double d2 = 0.0;
        for (int k = 0; k < dim; k++) d2 += (p[k] - X[(size_t)i * dim + k]) * (p[k] - X[(size_t)i * dim + k]);
        double d = sqrt(d2);
        if (d < dmin) { dmin = d; nmin = i; }
    }
    if (nmin < 0) return false;
    for (int e = 0; e < m.elen; e++) {
        const int *nd = m.h_elist.data() + (size_t)e * nn;
        bool has = false;
        for (int k = 0; k < nn; k++) has = has || nd[k] == nmin;
        if (!has) continue;
        int bsd[4], nb = 0;
        for (int s = 0; s < nn; s++) {   // Element::BndSideList (element.cc:338-348)
            bool all = true;
            for (int k = 0; k < dim; k++)
                all = all && m.h_bnd[nd[dim == 3 ? tet_side_node(s, k) : tri_side_node(s, k)]];
            if (all) bsd[nb++] = s;
        }
        if (!nb) continue;
        double g[GEO3], P[4][3];
        for (int k = 0; k < nn; k++)
            for (int d = 0; d < dim; d++) P[k][d] = X[(size_t)nd[k] * dim + d];
        if (dim == 3) tet_geometry(P[0], P[1], P[2], P[3], g);
        else tri_geometry(P[0], P[1], P[2], g);
        double ploc[3], nloc[3], pnew[3];
        el_local(dim, g, p, ploc);
        for (int t = 0; t < nb; t++) {
            int s = bsd[t];
            const double *lnm = dim == 3 ? lnm3[s] : lnm2[s];
            int n0 = nd[dim == 3 ? tet_side_node(s, 0) : tri_side_node(s, 0)];
            el_local(dim, g, X + (size_t)n0 * dim, nloc);
            double pd = 0.0;
            for (int k = 0; k < dim; k++) pd += (ploc[k] - nloc[k]) * lnm[k];
            for (int k = 0; k < dim; k++) pnew[k] = ploc[k] - lnm[k] * pd;
            const double EPS = 1e-8;   // LContains with padding
            bool inside = dim == 3 ? (pnew[0] + EPS >= 0.0 && pnew[1] + EPS >= 0.0 && pnew[2] + EPS >= 0.0 &&
                                      pnew[0] + pnew[1] + pnew[2] - EPS <= 1.0)
                                   : (pnew[0] + EPS >= 0.0 && pnew[1] + EPS >= 0.0 && pnew[0] + pnew[1] - EPS <= 1.0);
            if (!inside) continue;
            // Element::Global (element.cc:164-176)
            double u[4];
            if (dim == 3) { u[0] = 1.0 - pnew[0] - pnew[1] - pnew[2]; u[1] = pnew[0]; u[2] = pnew[1]; u[3] = pnew[2]; }
            else { u[0] = 1.0 - pnew[0] - pnew[1]; u[1] = pnew[0]; u[2] = pnew[1]; }
            for (int d = 0; d < dim; d++) {
                double acc = 0.0;
                for (int k = 0; k < nn; k++) acc += u[k] * P[k][d];
                out[d] = acc;
            }
            return true;
        }
    }
    // PullToBoundary_old (mesh.cc:1114-1172): project p onto the plane of the
    // boundary side through nmin whose centre is nearest to p
    const double EPS = 1e-10;
    double wmax = 0.0;
    bool found = false;
    for (int e = 0; e < m.elen; e++) {
        const int *nd = m.h_elist.data() + (size_t)e * nn;
        bool has = false;
        for (int k = 0; k < nn; k++) has = has || nd[k] == nmin;
        if (!has) continue;
        double g[GEO3], P[4][3];
        for (int k = 0; k < nn; k++)
            for (int d = 0; d < dim; d++) P[k][d] = X[(size_t)nd[k] * dim + d];
        if (dim == 3) tet_geometry(P[0], P[1], P[2], P[3], g);
        else tri_geometry(P[0], P[1], P[2], g);
        for (int s = 0; s < nn; s++) {
            bool containsmin = false, isbnd = true;
            for (int k = 0; k < dim; k++) {
                int v = nd[dim == 3 ? tet_side_node(s, k) : tri_side_node(s, k)];
                if (v == nmin) containsmin = true;
                else if (!m.h_bnd[v]) { isbnd = false; break; }
            }
            if (!containsmin || !isbnd) continue;
            double gc[3] = {0, 0, 0}, dc[3] = {0, 0, 0};
            for (int k = 0; k < dim; k++) {
                int lk = dim == 3 ? tet_side_node(s, k) : tri_side_node(s, k);
                for (int d = 0; d < dim; d++) gc[d] += P[lk][d] / dim;
            }
            // outward direction cosine (tri3.cc:155-163, tet4.cc:230-241):
            // minus the gradient of the shape function of the opposite node
            static const int opp3[4] = {3, 2, 1, 0}, opp2[3] = {2, 0, 1};
            int o = dim == 3 ? opp3[s] : opp2[s];
            const double *b = g + (dim == 3 ? 4 : 3), *c = g + (dim == 3 ? 8 : 6), *dd = g + 12;
            dc[0] = -b[o];
            dc[1] = -c[o];
            if (dim == 3) dc[2] = -dd[o];
            double len = 0.0;
            for (int d = 0; d < dim; d++) len += dc[d] * dc[d];
            len = sqrt(len);
            for (int d = 0; d < dim; d++) dc[d] /= len;
            double d2 = 0.0;
            for (int d = 0; d < dim; d++) d2 += (p[d] - gc[d]) * (p[d] - gc[d]);
            double dist = sqrt(d2);
            if (dist < EPS) dist = EPS;
            double w = 1.0 / dist;
            if (w > wmax) {
                wmax = w;
                double pd = 0.0;
                for (int d = 0; d < dim; d++) pd += (p[d] - gc[d]) * dc[d];
                for (int d = 0; d < dim; d++) out[d] = p[d] - dc[d] * pd;
                found = true;
            }
        }
    }
    return found;
}

// QVec_Point (source.cc:157-196): shape functions of the element holding
// p at p; with Neumann mode only on boundary nodes.  Elements are searched
// in ascending order (the reference's Mesh::ElFind starts next to the last
// element found; on a shared face both elements give the same values).  A
// point outside the mesh takes the element of the nearest node pair, as the
// reference does.
void qvec_point(const Mesh &m, const double *p, bool neumann, double *out)
{
    const int dim = m.dim, nn = m.nn;
    const double *X = m.h_nodes.data();
    int el = -1;
    double g[GEO3], P[4][3], loc[3];
    auto geom = [&](int e) {
        const int *nd = m.h_elist.data() + (size_t)e * nn;
        for (int k = 0; k < nn; k++)
            for (int d = 0; d < dim; d++) P[k][d] = X[(size_t)nd[k] * dim + d];
        if (dim == 3) tet_geometry(P[0], P[1], P[2], P[3], g);
        else tri_geometry(P[0], P[1], P[2], g);
    };
    for (int e = 0; e < m.elen && el < 0; e++) {
        geom(e);
        bool in = true;   // Element_Unstructured::GContains: bounding box (EPS 1e-8) then LContains
        for (int d = 0; d < dim && in; d++) {
            double lo = P[0][d], hi = P[0][d];
            for (int k = 1; k < nn; k++) { lo = std::min(lo, P[k][d]); hi = std::max(hi, P[k][d]); }
            if (p[d] < lo - 1e-8 || p[d] > hi + 1e-8) in = false;
        }
        if (!in) continue;
        el_local(dim, g, p, loc);
        const double EPS = 1e-8;
        in = dim == 3 ? (loc[0] + EPS >= 0.0 && loc[1] + EPS >= 0.0 && loc[2] + EPS >= 0.0 &&
                         loc[0] + loc[1] + loc[2] - EPS <= 1.0)
                      : (loc[0] + EPS >= 0.0 && loc[1] + EPS >= 0.0 && loc[0] + loc[1] - EPS <= 1.0);
        if (in) el = e;
    }
    if (el < 0) {
        double dmin = 1e10, dmin2 = 1e10;
        const double eps = 1e-8;
        for (int e = 0; e < m.elen; e++) {
            const int *nd = m.h_elist.data() + (size_t)e * nn;
            for (int j = 0; j < nn; j++) {
                double d = 0.0;
                for (int k = 0; k < dim; k++) d += (X[(size_t)nd[j] * dim + k] - p[k]) * (X[(size_t)nd[j] * dim + k] - p[k]);
                d = sqrt(d);
                if (d < dmin + eps) {
                    for (int kk = 0; kk < nn; kk++) {
                        if (kk == j) continue;
                        double d2 = 0.0;
                        for (int k = 0; k < dim; k++)
                            d2 += (X[(size_t)nd[kk] * dim + k] - p[k]) * (X[(size_t)nd[kk] * dim + k] - p[k]);
                        d2 = sqrt(d2);
                        if (d2 < dmin2) { dmin2 = d2; dmin = d; el = e; }
                    }
                }
            }
        }
        TMI_CHECK(el >= 0, TMI_ERR_INVALID_ARG, "no element found for the point source");
        geom(el);
        el_local(dim, g, p, loc);
    }
    for (int i = 0; i < m.n; i++) out[i] = 0.0;
    double fun[4];
    if (dim == 3) { fun[0] = 1.0 - loc[0] - loc[1] - loc[2]; fun[1] = loc[0]; fun[2] = loc[1]; fun[3] = loc[2]; }
    else { fun[0] = 1.0 - loc[0] - loc[1]; fun[1] = loc[0]; fun[2] = loc[1]; }
    const int *nd = m.h_elist.data() + (size_t)el * nn;
    for (int i = 0; i < nn; i++)
        if (m.h_bnd[nd[i]] || !neumann) out[nd[i]] = fun[i];
}

}  // namespace tmi
