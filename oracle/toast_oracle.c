/*
 * toast_oracle.c -- FP64 CPU restatement of the Toast++ diffusion forward
 * solve, projection and Jacobian path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the GPU library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in the product (toastpp_amd/) links or calls it.
 *
 * Every function restates the reference algorithm it names (file:line in the
 * toastpp/toastpp tree).  Element integrals, assembly order and solver steps
 * follow the reference operation by operation so that FP64 results agree to
 * the last bit where the reference is deterministic.
 *
 * Pinning (how each part is checked; see tests/test_oracle.py):
 *   - mesh reader / Setup / MarkBoundary / CSR / assembly / QVec_Gaussian /
 *     QM loading with PullToBoundary / PCG / BiCGSTAB / ProjectAll + log:
 *     end-to-end against the reference's own golden vectors
 *     test/2D/fwdfem/testdata.zip and test/3D/fwdfem/testdata.zip, plus
 *     closed forms (element integrals vs barycentric monomial integrals,
 *     row sums of the mass matrix, zero row sums of the stiffness matrix).
 *   - nodal Jacobian: finite differences of the projections with respect to
 *     nodal cmua / ckappa (the check of test/2D/matlab/jacobian_test.m).
 *   - raster (pixel) basis: the element of every grid point against a
 *     brute-force barycentric search (EPS and shared-face rules included),
 *     the grid map B by exact reproduction of linear fields, and every
 *     entry of the grid Jacobian (mua and kappa blocks, complex with the
 *     masked adjoint gradient and CW with the unmasked one) against its
 *     closed form on linear fields; ImageGradient's central, forward,
 *     backward and zero branches by value on a quadratic image under a
 *     random mask (tests/test_oracle_raster.py).  The reference's tests
 *     hold no raster output, so these closed forms are the pin.
 *   - NodeMapMatrix: weights summing to 1 at every node, exact
 *     reproduction of linear fields where all grid corners are inside.
 *   - linear pixel maps (Grid2LinPixMatrix, LinPix2GridMatrix,
 *     SubsampleLinPixel): brute-force tent weights, exact interpolation of
 *     linear fields, the solution voxels as the support of the inside mask;
 *     the Jacobians and the data gradient in a coarser basis equal the
 *     grid-basis ones mapped through G.
 *   - CW Jacobians and the data gradient: identities with the pinned
 *     complex Jacobians (real parts, the adjoint-source linearity of
 *     AddDataGradient).
 *   - QVec_Point: a unit vector at a node, partition of unity and linear
 *     reproduction inside an element, boundary nodes only for Neumann.
 *   - GMRES(20) (numpy, oracle.py gmres): the golden vectors of the 2D
 *     GMRES targets without and with DIAG (fmod/farg_cplx_gmres[_diag].fem).
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OM_PI 3.1415926535897932384626433832795
#define OM_C0 0.3

typedef struct {
    int dim, n, elen, nn, nsd, nsn;   /* nn nodes/element, nsd sides, nsn nodes/side */
    double *x;                        /* n*dim */
    int *el;                          /* elen*nn, 0-based */
    int *nreg;                        /* node region (-1 if none) */
    int *bnd;                         /* node boundary flag after MarkBoundary */
    double *ga, *gb, *gc, *gd;        /* elen*nn geometry coefficients */
    double *size;                     /* elen signed size */
    double *sside;                    /* elen*4 side areas (Tet4) */
    unsigned char *bndside;           /* elen*nsd */
    unsigned char *bndel;             /* elen */
    double *intdd;                    /* elen*nn*nn (full) */
    double *intbff;                   /* elen*nn*nn (full) */
    int *csr_rp, *csr_ci, csr_nnz;    /* SparseRowStructure, built on first use */
} om_mesh;

static const int TET_SN[4][3] = {{0,1,2},{0,3,1},{0,2,3},{1,3,2}};   /* tet4.cc:177 */
static const int TRI_SN[3][2] = {{0,1},{1,2},{2,0}};                 /* tri3.cc:121-127 */

static void *xcalloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s); if (!p) { fprintf(stderr, "oracle: out of memory\n"); exit(2); } return p; }

/* ------------------------------------------------------------------------
 * Mesh reader: "MeshData 5.0" (mesh.cc:1787-1825, ndlist.cc:146-159,
 * node.cc:130-162, ellist.cc:154-190, element.cc:429-437).
 * Only Tri3 ('o', ELID 15) and Tet4 ('c', ELID 3) are accepted.
 * ------------------------------------------------------------------------ */
static om_mesh *mesh_alloc(int dim, int n, int elen)
{
    om_mesh *m = (om_mesh*)xcalloc(1, sizeof(om_mesh));
    m->dim = dim; m->n = n; m->elen = elen;
    m->nn = dim + 1; m->nsd = dim + 1; m->nsn = dim;
    m->x = (double*)xcalloc((size_t)n*dim, sizeof(double));
    m->el = (int*)xcalloc((size_t)elen*m->nn, sizeof(int));
    m->nreg = (int*)xcalloc(n, sizeof(int));
    m->bnd = (int*)xcalloc(n, sizeof(int));
    return m;
}

om_mesh *om_mesh_load(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return NULL;
    char line[1024];
    int found = 0;
    while (fgets(line, sizeof line, f)) if (!strncmp(line, "MeshData 5.0", 12)) { found = 1; break; }
    if (!found) { fclose(f); return NULL; }
    int n = 0, dof = 0;
    while (fgets(line, sizeof line, f)) if (sscanf(line, " NodeList %d %d", &n, &dof) >= 1) break;
    double *xs = (double*)xcalloc((size_t)n*3, sizeof(double));
    int *reg = (int*)xcalloc(n, sizeof(int));
    int dim = 0;
    for (int i = 0; i < n; i++) {
        if (!fgets(line, sizeof line, f)) { fclose(f); return NULL; }
        char *p = strchr(line, '[');
        if (!p) { fclose(f); return NULL; }
        p++;
        int d = 0; char *end;
        while (d < 3) {
            double v = strtod(p, &end);
            if (end == p) break;
            xs[i*3+d] = v; d++; p = end;
        }
        if (i == 0) dim = d;
        char *q = strchr(p, ']');
        reg[i] = -1;
        if (q && q[1] == 'R') reg[i] = atoi(q+2);
    }
    int elen = 0;
    while (fgets(line, sizeof line, f)) if (!strncasecmp(line, "ElementList", 11)) { elen = atoi(line+11); break; }
    om_mesh *m = mesh_alloc(dim, n, elen);
    for (int i = 0; i < n; i++) { for (int d = 0; d < dim; d++) m->x[i*dim+d] = xs[i*3+d]; m->nreg[i] = reg[i]; }
    free(xs); free(reg);
    for (int e = 0; e < elen; e++) {
        char c;
        do { if (fscanf(f, " %c", &c) != 1) { fclose(f); return NULL; } } while (c == ' ');
        int id = (c >= 'a' && c <= 'z') ? 1 + (c - 'a') : 1 + (c - '1');
        if ((dim == 2 && id != 15) || (dim == 3 && id != 3)) { fclose(f); fprintf(stderr, "oracle: unsupported element type %d\n", id); return NULL; }
        for (int k = 0; k < m->nn; k++) { int v; if (fscanf(f, "%d", &v) != 1) { fclose(f); return NULL; } m->el[e*m->nn+k] = v - 1; }
        if (!fgets(line, sizeof line, f)) break;   /* rest of line */
    }
    fclose(f);
    return m;
}

om_mesh *om_mesh_make(int dim, int n, const double *x, int elen, const int *el)
{
    om_mesh *m = mesh_alloc(dim, n, elen);
    memcpy(m->x, x, sizeof(double)*(size_t)n*dim);
    memcpy(m->el, el, sizeof(int)*(size_t)elen*m->nn);
    for (int i = 0; i < n; i++) m->nreg[i] = -1;
    return m;
}

void om_mesh_free(om_mesh *m)
{
    if (!m) return;
    free(m->x); free(m->el); free(m->nreg); free(m->bnd);
    free(m->ga); free(m->gb); free(m->gc); free(m->gd); free(m->size); free(m->sside);
    free(m->bndside); free(m->bndel); free(m->intdd); free(m->intbff);
    free(m->csr_rp); free(m->csr_ci);
    free(m);
}

int om_dim(const om_mesh *m) { return m->dim; }
int om_nlen(const om_mesh *m) { return m->n; }
int om_elen(const om_mesh *m) { return m->elen; }
void om_get_nodes(const om_mesh *m, double *x) { memcpy(x, m->x, sizeof(double)*(size_t)m->n*m->dim); }
void om_get_elements(const om_mesh *m, int *el) { memcpy(el, m->el, sizeof(int)*(size_t)m->elen*m->nn); }
void om_get_bnd(const om_mesh *m, int *b) { memcpy(b, m->bnd, sizeof(int)*(size_t)m->n); }
void om_get_regions(const om_mesh *m, int *r) { memcpy(r, m->nreg, sizeof(int)*(size_t)m->n); }
void om_get_size(const om_mesh *m, double *s) { memcpy(s, m->size, sizeof(double)*(size_t)m->elen); }
void om_get_bndside(const om_mesh *m, unsigned char *b) { memcpy(b, m->bndside, (size_t)m->elen*m->nsd); }
void om_get_intbff(const om_mesh *m, double *b) { memcpy(b, m->intbff, sizeof(double)*(size_t)m->elen*m->nn*m->nn); }
void om_get_intdd(const om_mesh *m, double *b) { memcpy(b, m->intdd, sizeof(double)*(size_t)m->elen*m->nn*m->nn); }

/* ------------------------------------------------------------------------
 * Element geometry.  Tet4: tet4.cc:113-171 (coefficients), :1011-1014 (size),
 * tet4.cc:19-29 (side areas by Heron); Tri3: tri3.cc:79-90, :184-189.
 * ------------------------------------------------------------------------ */
static double pdist(const double *p, const double *q, int dim)   /* point.cc:33-41 */
{
    double d2 = 0.0;
    for (int i = 0; i < dim; i++) d2 += (p[i]-q[i]) * (p[i]-q[i]);
    return sqrt(d2);
}

static double tri_area3(const double *p1, const double *p2, const double *p3)  /* tet4.cc:19-29 */
{
    double a = pdist(p1, p2, 3), b = pdist(p2, p3, 3), c = pdist(p3, p1, 3);
    double s = 0.5 * (a+b+c);
    return sqrt(s * (s-a) * (s-b) * (s-c));
}

static void geometry(om_mesh *m)
{
    int E = m->elen, nn = m->nn;
    m->ga = (double*)xcalloc((size_t)E*nn, sizeof(double));
    m->gb = (double*)xcalloc((size_t)E*nn, sizeof(double));
    m->gc = (double*)xcalloc((size_t)E*nn, sizeof(double));
    m->gd = (double*)xcalloc((size_t)E*nn, sizeof(double));
    m->size = (double*)xcalloc(E, sizeof(double));
    m->sside = (double*)xcalloc((size_t)E*4, sizeof(double));
    for (int e = 0; e < E; e++) {
        const int *nd = m->el + e*nn;
        double *a = m->ga + e*nn, *b = m->gb + e*nn, *c = m->gc + e*nn, *d = m->gd + e*nn;
        if (m->dim == 3) {
            const double *P0 = m->x + 3*nd[0], *P1 = m->x + 3*nd[1], *P2 = m->x + 3*nd[2], *P3 = m->x + 3*nd[3];
            double x0 = P0[0], y0 = P0[1], z0 = P0[2];
            double x1 = P1[0], y1 = P1[1], z1 = P1[2];
            double x2 = P2[0], y2 = P2[1], z2 = P2[2];
            double x3 = P3[0], y3 = P3[1], z3 = P3[2];
            /* cofactor expansion, same operand order as tet4.cc:125-140 */
            a[0] = x1*(y2*z3-y3*z2) - x2*(y1*z3-y3*z1) + x3*(y1*z2-y2*z1);
            a[1] = x0*(y3*z2-y2*z3) - x2*(y3*z0-y0*z3) + x3*(y2*z0-y0*z2);
            a[2] = x0*(y1*z3-y3*z1) - x1*(y0*z3-y3*z0) + x3*(y0*z1-y1*z0);
            a[3] = x0*(y2*z1-y1*z2) - x1*(y2*z0-y0*z2) + x2*(y1*z0-y0*z1);
            b[0] = y1*(z3-z2) + y2*(z1-z3) + y3*(z2-z1);
            b[1] = y0*(z2-z3) + y2*(z3-z0) + y3*(z0-z2);
            b[2] = y0*(z3-z1) + y1*(z0-z3) + y3*(z1-z0);
            b[3] = y0*(z1-z2) + y1*(z2-z0) + y2*(z0-z1);
            c[0] = x1*(z2-z3) + x2*(z3-z1) + x3*(z1-z2);
            c[1] = x0*(z3-z2) + x2*(z0-z3) + x3*(z2-z0);
            c[2] = x0*(z1-z3) + x1*(z3-z0) + x3*(z0-z1);
            c[3] = x0*(z2-z1) + x1*(z0-z2) + x2*(z1-z0);
            d[0] = x1*(y3-y2) + x2*(y1-y3) + x3*(y2-y1);
            d[1] = x0*(y2-y3) + x2*(y3-y0) + x3*(y0-y2);
            d[2] = x0*(y3-y1) + x1*(y0-y3) + x3*(y1-y0);
            d[3] = x0*(y1-y2) + x1*(y2-y0) + x2*(y0-y1);
            for (int sd = 0; sd < 4; sd++)
                m->sside[e*4+sd] = tri_area3(m->x + 3*nd[TET_SN[sd][0]], m->x + 3*nd[TET_SN[sd][1]], m->x + 3*nd[TET_SN[sd][2]]);
            m->size[e] = (1.0/6.0) * (a[0]+a[1]+a[2]+a[3]);
        } else {
            const double *P0 = m->x + 2*nd[0], *P1 = m->x + 2*nd[1], *P2 = m->x + 2*nd[2];
            double x0 = P0[0], y0 = P0[1], x1 = P1[0], y1 = P1[1], x2 = P2[0], y2 = P2[1];
            a[0] = x1*y2 - x2*y1;  b[0] = y1-y2;  c[0] = x2-x1;
            a[1] = x2*y0 - x0*y2;  b[1] = y2-y0;  c[1] = x0-x2;
            a[2] = x0*y1 - x1*y0;  b[2] = y0-y1;  c[2] = x1-x0;
            m->size[e] = 0.5 * (a[0]+a[1]+a[2]);
        }
    }
}

/* ------------------------------------------------------------------------
 * MarkBoundary (mesh.cc:1614-1703): sides sorted by node tuple, a side that
 * equals neither sorted neighbour is a boundary side; its nodes become
 * boundary nodes.
 * ------------------------------------------------------------------------ */
typedef struct { int el, sd, nd[3]; } side_rec;
static int g_nsn;
static int side_cmp(const void *A, const void *B)
{
    const side_rec *a = (const side_rec*)A, *b = (const side_rec*)B;
    for (int k = 0; k < g_nsn; k++) { if (a->nd[k] < b->nd[k]) return -1; if (a->nd[k] > b->nd[k]) return 1; }
    /* ties are broken by element/side so the order is deterministic */
    if (a->el != b->el) return a->el < b->el ? -1 : 1;
    return a->sd < b->sd ? -1 : (a->sd > b->sd);
}

static void mark_boundary(om_mesh *m)
{
    int E = m->elen, nsd = m->nsd, nsn = m->nsn;
    size_t S = (size_t)E*nsd;
    side_rec *sd = (side_rec*)xcalloc(S, sizeof(side_rec));
    for (int e = 0; e < E; e++)
        for (int s = 0; s < nsd; s++) {
            side_rec *r = sd + (size_t)e*nsd + s;
            r->el = e; r->sd = s;
            for (int k = 0; k < nsn; k++) r->nd[k] = m->el[e*m->nn + (m->dim == 3 ? TET_SN[s][k] : TRI_SN[s][k])];
            for (int j = 0; j < nsn-1; j++) for (int k = j+1; k < nsn; k++)
                if (r->nd[j] > r->nd[k]) { int t = r->nd[j]; r->nd[j] = r->nd[k]; r->nd[k] = t; }
        }
    g_nsn = nsn;
    qsort(sd, S, sizeof(side_rec), side_cmp);
    m->bndside = (unsigned char*)xcalloc(S, 1);
    m->bndel = (unsigned char*)xcalloc(E, 1);
    for (int i = 0; i < m->n; i++) m->bnd[i] = 0;
    for (size_t i = 0; i < S; i++) {
        int isbnd = 1;
        if (i && !memcmp(sd[i].nd, sd[i-1].nd, sizeof(int)*nsn)) isbnd = 0;
        if (i+1 < S && isbnd && !memcmp(sd[i].nd, sd[i+1].nd, sizeof(int)*nsn)) isbnd = 0;
        if (isbnd) {
            m->bndside[(size_t)sd[i].el*nsd + sd[i].sd] = 1;
            m->bndel[sd[i].el] = 1;
            for (int k = 0; k < nsn; k++) m->bnd[sd[i].nd[k]] = 1;
        }
    }
    free(sd);
}

/* ------------------------------------------------------------------------
 * PostInitialisation: intdd (tet4.cc:1016-1031, tri3.cc:490-506) and
 * intbff (tet4.cc:1109-1122 with sym_bndintff :1039-1068; tri3.cc:561-582)
 * ------------------------------------------------------------------------ */
static int in_side_tet(int sd, int i) { return i == TET_SN[sd][0] || i == TET_SN[sd][1] || i == TET_SN[sd][2]; }

static void post_init(om_mesh *m)
{
    int E = m->elen, nn = m->nn;
    m->intdd = (double*)xcalloc((size_t)E*nn*nn, sizeof(double));
    m->intbff = (double*)xcalloc((size_t)E*nn*nn, sizeof(double));
    for (int e = 0; e < E; e++) {
        const double *b = m->gb + e*nn, *c = m->gc + e*nn, *d = m->gd + e*nn;
        double *dd = m->intdd + (size_t)e*nn*nn, *bf = m->intbff + (size_t)e*nn*nn;
        if (m->dim == 3) {
            double scale = 1.0/(36.0*m->size[e]);
            for (int i = 0; i < 4; i++) for (int j = 0; j <= i; j++)
                dd[i*4+j] = dd[j*4+i] = scale * (b[j]*b[i]+c[j]*c[i]+d[j]*d[i]);
            if (m->bndel[e]) {
                const double f = 2.0/24.0;
                for (int sd = 0; sd < 4; sd++) {
                    if (!m->bndside[e*4+sd]) continue;
                    for (int i = 0; i < 4; i++) for (int j = 0; j <= i; j++) {
                        double v = (in_side_tet(sd,i) && in_side_tet(sd,j)) ? (i == j ? 2.0 : 1.0) * f : 0.0;
                        double t = v * m->sside[e*4+sd];
                        bf[i*4+j] += t;
                        if (i != j) bf[j*4+i] = bf[i*4+j];
                    }
                }
            }
        } else {
            double scale = 0.25/m->size[e];
            for (int i = 0; i < 3; i++) for (int j = 0; j <= i; j++)
                dd[i*3+j] = dd[j*3+i] = scale * (b[i]*b[j] + c[i]*c[j]);
            if (m->bndel[e]) {
                for (int s = 0; s < 3; s++) {
                    if (!m->bndside[e*3+s]) continue;
                    int n0 = TRI_SN[s][0], n1 = TRI_SN[s][1];
                    const double *p0 = m->x + 2*m->el[e*3+n0], *p1 = m->x + 2*m->el[e*3+n1];
                    double dx = p0[0]-p1[0], dy = p0[1]-p1[1];
                    double dl = hypot(dx, dy);
                    for (int i = 0; i < 3; i++) {
                        if (i != n0 && i != n1) continue;
                        bf[i*3+i] += dl/3.0;
                        for (int j = 0; j < i; j++)
                            if (j == n0 || j == n1) { bf[i*3+j] += dl/6.0; bf[j*3+i] = bf[i*3+j]; }
                    }
                }
            }
        }
    }
}

/* Mesh::Setup (mesh.cc:96-158): Initialise, MarkBoundary, PostInitialisation */
void om_setup(om_mesh *m)
{
    geometry(m);
    mark_boundary(m);
    post_init(m);
}

/* ------------------------------------------------------------------------
 * Element integrals (reference tables restated as integer numerators)
 * ------------------------------------------------------------------------ */
static int mult3(int i, int j, int k)      /* 6 / 2 / 1 numerators of int l_i l_j l_k */
{
    if (i == j && j == k) return 6;
    if (i == j || j == k || i == k) return 2;
    return 1;
}

static double el_intff(const om_mesh *m, int e, int i, int j)   /* tet4.cc:348-355, tri3.cc:285-292 */
{
    if (m->dim == 3) return ((i == j ? 2.0 : 1.0) * (1.0/20.0)) * m->size[e];
    return ((i == j ? 2.0 : 1.0) * (1.0/12.0)) * m->size[e];
}

static double el_intfff(const om_mesh *m, int e, int i, int j, int k)  /* tet4.cc:357-361, tri3.cc:294-297 */
{
    if (m->dim == 3) return ((double)mult3(i,j,k) / 120.0) * m->size[e];
    static const double sc[3] = {1.0/60.0, 1.0/30.0, 1.0/10.0};
    int idx = (i == j && j == k) ? 2 : ((i == j || j == k || i == k) ? 1 : 0);
    return m->size[e] * sc[idx];
}

static double el_intfdd(const om_mesh *m, int e, int i, int j, int k)   /* tet4.cc:380-412, tri3.h:196-197 */
{
    (void)i;
    int nn = m->nn;
    if (m->dim == 3) {
        if (j > k) { int t = j; j = k; k = t; }
        const double *b = m->gb + e*nn, *c = m->gc + e*nn, *d = m->gd + e*nn;
        double scale = 1.0/(144.0*m->size[e]);
        return (b[j]*b[k]+c[j]*c[k]+d[j]*d[k]) * scale;
    }
    return m->intdd[(size_t)e*nn*nn + j*nn + k] * 0.3333333333;
}

static double el_intpff(const om_mesh *m, int e, int i, int j, const double *P)  /* tet4.cc:371-378, tri3.cc:325-353 */
{
    const int *nd = m->el + e*m->nn;
    if (m->dim == 3) {
        return ((double)mult3(0,i,j)/120.0 * P[nd[0]] +
                (double)mult3(1,i,j)/120.0 * P[nd[1]] +
                (double)mult3(2,i,j)/120.0 * P[nd[2]] +
                (double)mult3(3,i,j)/120.0 * P[nd[3]]) * m->size[e];
    }
    if (i < j) { int t = i; i = j; j = t; }
    const double i30 = 1.0/30.0;
    double p0 = P[nd[0]], p1 = P[nd[1]], p2 = P[nd[2]];
    double s;
    if (i == 0) s = 3.0*p0 + p1 + p2;
    else if (i == 1) s = (j == 0) ? p0 + p1 + 0.5*p2 : p0 + 3.0*p1 + p2;
    else s = (j == 0) ? p0 + 0.5*p1 + p2 : (j == 1) ? 0.5*p0 + p1 + p2 : p0 + p1 + 3.0*p2;
    return m->size[e]*i30 * s;
}

static double el_intpdd(const om_mesh *m, int e, int i, int j, const double *P)  /* tet4.cc:414-419, tri3.h:208-209 */
{
    const int *nd = m->el + e*m->nn;
    if (m->dim == 3) {
        double res = 0.0;
        for (int k = 0; k < 4; k++) res += P[nd[k]];
        return res * el_intfdd(m, e, 0, i, j);
    }
    return m->intdd[(size_t)e*9 + i*3 + j] * ((P[nd[0]]+P[nd[1]]+P[nd[2]])/3.0);
}

static double el_bndintpff(const om_mesh *m, int e, int i, int j, const double *P)  /* tet4.cc:819-832, tri3.h:217-218 */
{
    const int *nd = m->el + e*m->nn;
    if (m->dim == 3) {
        if (!m->bndel[e]) return 0.0;
        double res = 0.0;
        for (int sd = 0; sd < 4; sd++) {
            if (!m->bndside[e*4+sd]) continue;
            double sres = 0.0;
            for (int k = 0; k < 4; k++) {
                double t = (in_side_tet(sd,i) && in_side_tet(sd,j) && in_side_tet(sd,k)) ? (double)mult3(k,i,j) * (2.0/120.0) : 0.0;
                sres += P[nd[k]] * t;
            }
            res += sres * m->sside[e*4+sd];
        }
        return res;
    }
    return m->intbff[(size_t)e*9 + i*3 + j] * ((P[nd[0]]+P[nd[1]]+P[nd[2]])/3.0);
}

/* ------------------------------------------------------------------------
 * CSR structure (mesh.cc:422-549): sorted unique (row, col) element pairs,
 * diagonal included.
 * ------------------------------------------------------------------------ */
static int pair_cmp(const void *A, const void *B)
{
    const int *a = (const int*)A, *b = (const int*)B;
    if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
    return a[1] < b[1] ? -1 : (a[1] > b[1]);
}

static void build_csr(om_mesh *m)
{
    if (m->csr_rp) return;
    size_t P = (size_t)m->elen*m->nn*m->nn;
    int *pr = (int*)xcalloc(P*2, sizeof(int));
    size_t p = 0;
    for (int e = 0; e < m->elen; e++)
        for (int i = 0; i < m->nn; i++) for (int j = 0; j < m->nn; j++) {
            pr[2*p] = m->el[e*m->nn+i]; pr[2*p+1] = m->el[e*m->nn+j]; p++;
        }
    qsort(pr, P, 2*sizeof(int), pair_cmp);
    int nz = 0;
    for (size_t i = 0; i < P; i++)
        if (!i || pr[2*i] != pr[2*i-2] || pr[2*i+1] != pr[2*i-1]) { pr[2*nz] = pr[2*i]; pr[2*nz+1] = pr[2*i+1]; nz++; }
    int *rp = (int*)xcalloc(m->n+1, sizeof(int));
    int *ci = (int*)xcalloc(nz, sizeof(int));
    for (int i = 0; i < nz; i++) { rp[pr[2*i]+1]++; ci[i] = pr[2*i+1]; }
    for (int i = 0; i < m->n; i++) rp[i+1] += rp[i];
    m->csr_rp = rp; m->csr_ci = ci; m->csr_nnz = nz;
    free(pr);
}

int om_csr_nnz(om_mesh *m) { build_csr(m); return m->csr_nnz; }
void om_csr(om_mesh *m, int *rowptr, int *colidx)
{
    build_csr(m);
    memcpy(rowptr, m->csr_rp, sizeof(int)*(size_t)(m->n+1));
    memcpy(colidx, m->csr_ci, sizeof(int)*(size_t)m->csr_nnz);
}

/* TCompRowMatrix::operator() row search (crmatrix_imp.hpp:756-767) */
static int slot(const int *rp, const int *ci, int r, int c)
{
    for (int k = rp[r]; k < rp[r+1]; k++) if (ci[k] == c) return k;
    fprintf(stderr, "oracle: entry (%d,%d) not in pattern\n", r, c); exit(3);
}

/* ------------------------------------------------------------------------
 * AssembleSystemMatrix (fwdsolver.cc:405-424 complex, :338-357 real):
 * PFF(cmua), PDD(ckappa), BNDPFF(c2a) then iCFF(omega); elements ascending,
 * local i then j (mesh.cc:2085-2129, :2252-2282).
 * val: complex interleaved [re, im] per nonzero.
 * ------------------------------------------------------------------------ */
void om_assemble(const om_mesh *m, const int *rp, const int *ci, const double *cmua, const double *ckappa,
                 const double *c2a, double omega, double *val)
{
    int nz = rp[m->n], nn = m->nn;
    memset(val, 0, sizeof(double)*2*(size_t)nz);
    for (int pass = 0; pass < 4; pass++)
        for (int e = 0; e < m->elen; e++) {
            const int *nd = m->el + e*nn;
            for (int i = 0; i < nn; i++) for (int j = 0; j < nn; j++) {
                int k = slot(rp, ci, nd[i], nd[j]);
                switch (pass) {
                case 0: val[2*k] += el_intpff(m, e, i, j, cmua); break;
                case 1: val[2*k] += el_intpdd(m, e, i, j, ckappa); break;
                case 2: val[2*k] += el_bndintpff(m, e, i, j, c2a); break;
                case 3: val[2*k+1] += el_intff(m, e, i, j) * omega; break;
                }
            }
        }
}

/* ------------------------------------------------------------------------
 * Element-basis assembly: AssembleSystemMatrix(sol, omega, elbasis = true)
 * (fwdsolver.cc:338-357, :405-424) with ASSEMBLE_PFF_EL / PDD_EL /
 * BNDPFF_EL (AddToElMatrix, mesh.cc:1985-1993; complex twin :2112-2120):
 * the element's coefficient times IntFF, IntDD (the stored intdd,
 * tet4.cc:1016-1031, tri3.cc:490-506) and BndIntFF (the stored intbff,
 * tet4.cc:1109-1122, tri3.cc:561-582; zero off the boundary), pass by pass
 * over ascending elements, then iCFF(omega) as in the nodal assembly.
 * cmua, ckappa, c2a: elen values each.  val: complex interleaved.
 * ------------------------------------------------------------------------ */
static void cmul(const double *a, const double *b, double *o);
static double el_intdd(const om_mesh *m, int e, int i, int j) { return m->intdd[(size_t)e*m->nn*m->nn + i*m->nn + j]; }
static double el_bndintff(const om_mesh *m, int e, int i, int j) { return m->intbff[(size_t)e*m->nn*m->nn + i*m->nn + j]; }

void om_assemble_el(const om_mesh *m, const int *rp, const int *ci, const double *cmua, const double *ckappa,
                    const double *c2a, double omega, double *val)
{
    int nz = rp[m->n], nn = m->nn;
    memset(val, 0, sizeof(double)*2*(size_t)nz);
    for (int pass = 0; pass < 4; pass++)
        for (int e = 0; e < m->elen; e++) {
            const int *nd = m->el + e*nn;
            for (int i = 0; i < nn; i++) for (int j = 0; j < nn; j++) {
                int k = slot(rp, ci, nd[i], nd[j]);
                switch (pass) {
                case 0: val[2*k] += el_intff(m, e, i, j) * cmua[e]; break;
                case 1: val[2*k] += el_intdd(m, e, i, j) * ckappa[e]; break;
                case 2: val[2*k] += el_bndintff(m, e, i, j) * c2a[e]; break;
                case 3: val[2*k+1] += el_intff(m, e, i, j) * omega; break;
                }
            }
        }
}

/* ------------------------------------------------------------------------
 * Element-basis Jacobian (GenerateJacobian with RASTER_ELBASIS,
 * jacobian.cc:95-96,117-119 -> GenerateJacobian_mesh :441-497 with
 * elbasis = true): per element
 *   pmdf_mua[el] = sum_i sum_j (f[n_i] g[n_j]) IntFF(i,j)   (IntFG_el, :879-906)
 *   pmdf_kap[el] = sum_i sum_j (f[n_i] g[n_j]) IntDD(i,j)   (IntGradFGradG_el, :911-939)
 * then /proj for logscale (PMDF_log, :74-79) and J = -Re / -Im in rows k and
 * k + nQM, mua block in columns [0, elen), kappa block in [elen, 2 elen).
 * dphi: nq x n, aphi: nm x n complex; proj: nQM complex (link order).
 * mvec_offset != 0 reproduces the mvec overload's view of the kappa PMDF
 * at offset n instead of elen (:397): mua columns n..elen-1 then hold the
 * kappa PMDF of elements 0..elen-n-1.
 * ------------------------------------------------------------------------ */
void om_jacobian_el(const om_mesh *m, int nq, int nm, const double *dphi, const double *aphi,
                    const int *lptr, const int *lidx, const double *proj, int logscale, int mvec_offset, double *J)
{
    (void)nm;
    int n = m->n, nn = m->nn, E = m->elen, nqm = lptr[nq];
    size_t rowlen = 2*(size_t)E;
    for (int q = 0; q < nq; q++)
        for (int k = lptr[q]; k < lptr[q+1]; k++) {
            const double *f = dphi + 2*(size_t)q*n, *g = aphi + 2*(size_t)lidx[k]*n;
            double pr = proj[2*k], pi = proj[2*k+1], den = pr*pr + pi*pi;
            double *R0 = J + (size_t)k*rowlen, *R1 = J + (size_t)(k+nqm)*rowlen;
            for (int e = 0; e < E; e++) {
                const int *nd = m->el + e*nn;
                double sm[2] = {0.0, 0.0}, sk[2] = {0.0, 0.0};
                for (int i = 0; i < nn; i++)
                    for (int j = 0; j < nn; j++) {
                        double fg[2];
                        cmul(f + 2*nd[i], g + 2*nd[j], fg);
                        double w = el_intff(m, e, i, j), wd = el_intdd(m, e, i, j);
                        sm[0] += fg[0]*w; sm[1] += fg[1]*w;
                        sk[0] += fg[0]*wd; sk[1] += fg[1]*wd;
                    }
                double a[2] = {sm[0], sm[1]}, b[2] = {sk[0], sk[1]};
                if (logscale) {   /* complex division by proj */
                    double ar = (a[0]*pr + a[1]*pi)/den, ai = (a[1]*pr - a[0]*pi)/den;
                    double br = (b[0]*pr + b[1]*pi)/den, bi = (b[1]*pr - b[0]*pi)/den;
                    a[0] = ar; a[1] = ai; b[0] = br; b[1] = bi;
                }
                R0[e] = -a[0]; R1[e] = -a[1]; R0[e+E] = -b[0]; R1[e+E] = -b[1];
            }
            if (mvec_offset)
                for (int e = n; e < E; e++) { R0[e] = R0[E + e - n]; R1[e] = R1[E + e - n]; }
        }
}

/* Mass matrix FF (ASSEMBLE_FF) used by closed-form checks */
void om_assemble_ff(const om_mesh *m, const int *rp, const int *ci, double *val)
{
    memset(val, 0, sizeof(double)*(size_t)rp[m->n]);
    for (int e = 0; e < m->elen; e++) {
        const int *nd = m->el + e*m->nn;
        for (int i = 0; i < m->nn; i++) for (int j = 0; j < m->nn; j++)
            val[slot(rp, ci, nd[i], nd[j])] += el_intff(m, e, i, j);
    }
}

/* Element-integral probes for closed-form tests */
double om_el_intfff(const om_mesh *m, int e, int i, int j, int k) { return el_intfff(m, e, i, j, k); }
double om_el_intfdd(const om_mesh *m, int e, int i, int j, int k) { return el_intfdd(m, e, i, j, k); }
double om_el_intpff(const om_mesh *m, int e, int i, int j, const double *P) { return el_intpff(m, e, i, j, P); }
double om_el_bndintpff(const om_mesh *m, int e, int i, int j, const double *P) { return el_bndintpff(m, e, i, j, P); }

/* A_Keijzer (param.cc:24-32) */
double om_a_keijzer(double n)
{
    double th = asin(1.0/n);
    double costh = fabs(cos(th));
    double R0 = ((n-1.0)*(n-1.0)) / ((n+1.0)*(n+1.0));
    return (2.0/(1.0-R0) - 1.0 + costh*costh*costh) / (1.0 - costh*costh);
}

/* Solution coefficients (toastmodule.cc:1308-1318, fwdfem.cc:737-742) */
void om_coeffs(int n, const double *mua, const double *mus, const double *ref, double *cmua, double *ckap, double *c2a)
{
    for (int i = 0; i < n; i++) {
        cmua[i] = mua[i]*OM_C0/ref[i];
        ckap[i] = OM_C0/(3.0*ref[i]*(mua[i]+mus[i]));
        c2a[i] = OM_C0/(2.0*ref[i]*om_a_keijzer(ref[i]));
    }
}

/* ------------------------------------------------------------------------
 * QVec_Gaussian (source.cc:4-41).  neumann != 0: boundary source distributed
 * with BndIntFF; otherwise isotropic volume source with IntFF.
 * ------------------------------------------------------------------------ */
void om_qvec_gaussian(const om_mesh *m, const double *cnt, double w, int neumann, double *q)
{
    int n = m->n, nn = m->nn;
    double w2 = w*w;
    double fac1 = 1.0 / (sqrt(2.0*OM_PI) * w);
    if (m->dim == 3) fac1 /= (sqrt(2.0*OM_PI) * w);
    if (!neumann) fac1 /= (sqrt(2.0*OM_PI) * w);
    double fac2 = -0.5/w2;
    memset(q, 0, sizeof(double)*(size_t)n);
    for (int e = 0; e < m->elen; e++) {
        const int *nd = m->el + e*nn;
        for (int i = 0; i < nn; i++) {
            int is = nd[i];
            if (neumann && !m->bnd[is]) continue;
            double d = pdist(cnt, m->x + (size_t)is*m->dim, m->dim);
            double qq = exp(d*d*fac2) * fac1;
            for (int j = 0; j < nn; j++) {
                int js = nd[j];
                if (neumann) { if (m->bnd[js]) q[js] += qq * m->intbff[(size_t)e*nn*nn + i*nn + j]; }
                else q[js] += qq * el_intff(m, e, i, j);
            }
        }
    }
}

/* ------------------------------------------------------------------------
 * QM file (qmmesh.cc:319-466) with measurement pulling (InitM
 * qmmesh.cc:105-145 -> Mesh::PullToBoundary mesh.cc:1052-1111).
 * ------------------------------------------------------------------------ */
typedef struct { int dim, nq, nm, nqm; double *Q, *M; int *lptr, *lidx; } om_qm;

static void el_local(const om_mesh *m, int e, const double *g, double *loc)  /* tet4.cc:186-197, tri3.cc:134-144 */
{
    int nn = m->nn;
    const double *a = m->ga + e*nn, *b = m->gb + e*nn, *c = m->gc + e*nn, *d = m->gd + e*nn;
    if (m->dim == 3) {
        double scale = 1.0/(a[0]+a[1]+a[2]+a[3]);
        loc[0] = (a[1] + b[1]*g[0] + c[1]*g[1] + d[1]*g[2]) * scale;
        loc[1] = (a[2] + b[2]*g[0] + c[2]*g[1] + d[2]*g[2]) * scale;
        loc[2] = (a[3] + b[3]*g[0] + c[3]*g[1] + d[3]*g[2]) * scale;
    } else {
        double scale = 1.0/(a[0]+a[1]+a[2]);
        loc[0] = (a[1] + b[1]*g[0] + c[1]*g[1]) * scale;
        loc[1] = (a[2] + b[2]*g[0] + c[2]*g[1]) * scale;
    }
}

static int el_lcontains(int dim, const double *l)   /* tet4.cc:268-280, tri3.cc:191-201 (pad = true) */
{
    const double EPS = 1e-8;
    if (dim == 3) return l[0]+EPS >= 0.0 && l[1]+EPS >= 0.0 && l[2]+EPS >= 0.0 && l[0]+l[1]+l[2]-EPS <= 1.0;
    return l[0]+EPS >= 0.0 && l[1]+EPS >= 0.0 && l[0]+l[1]-EPS <= 1.0;
}

static void el_global(const om_mesh *m, int e, const double *l, double *g)   /* element.cc:164-176 */
{
    double u[4];
    if (m->dim == 3) { u[0] = 1.0-l[0]-l[1]-l[2]; u[1] = l[0]; u[2] = l[1]; u[3] = l[2]; }
    else { u[0] = 1.0-l[0]-l[1]; u[1] = l[0]; u[2] = l[1]; }
    for (int i = 0; i < m->dim; i++) {
        g[i] = 0.0;
        for (int j = 0; j < m->nn; j++) g[i] += u[j] * m->x[(size_t)m->el[e*m->nn+j]*m->dim + i];
    }
}

static int pull_to_boundary(const om_mesh *m, const double *p, double *pshift)
{
    static const double lnm3[4][3] = {{0,0,-1},{0,-1,0},{-1,0,0},{0.577350269,0.577350269,0.577350269}};  /* tet4.cc:244-256 */
    static const double lnm2[3][2] = {{0,-1},{0.7071067814,0.7071067814},{-1,0}};                          /* tri3.cc:165-176 */
    int dim = m->dim, nn = m->nn, nmin = -1;
    double dmin = 1e10;
    for (int n = 0; n < m->n; n++) {
        if (!m->bnd[n]) continue;
        double d = pdist(p, m->x + (size_t)n*dim, dim);
        if (d < dmin) { dmin = d; nmin = n; }
    }
    for (int e = 0; e < m->elen; e++) {
        const int *nd = m->el + e*nn;
        int has = 0;
        for (int k = 0; k < nn; k++) if (nd[k] == nmin) has = 1;
        if (!has) continue;
        int bsd[4], nb = 0;
        for (int s = 0; s < m->nsd; s++) {          /* Element::BndSideList (element.cc:338-348) */
            int ok = 1;
            for (int k = 0; k < m->nsn; k++) if (!m->bnd[nd[dim == 3 ? TET_SN[s][k] : TRI_SN[s][k]]]) { ok = 0; break; }
            if (ok) bsd[nb++] = s;
        }
        if (!nb) continue;
        double ploc[3], nloc[3], pnew[3];
        el_local(m, e, p, ploc);
        for (int t = 0; t < nb; t++) {
            int s = bsd[t];
            const double *lnm = (dim == 3) ? lnm3[s] : lnm2[s];
            int n0 = nd[dim == 3 ? TET_SN[s][0] : TRI_SN[s][0]];
            el_local(m, e, m->x + (size_t)n0*dim, nloc);
            double pd = 0.0;
            for (int k = 0; k < dim; k++) pd += (ploc[k]-nloc[k]) * lnm[k];
            for (int k = 0; k < dim; k++) pnew[k] = ploc[k] - lnm[k]*pd;
            if (el_lcontains(dim, pnew)) { el_global(m, e, pnew, pshift); return 1; }
        }
    }
    /* PullToBoundary_old (mesh.cc:1114-1172): project onto the plane of the
     * boundary side through nmin whose centre is nearest to p */
    {
        const double EPS = 1e-10;
        double wmax = 0.0;
        int found = 0;
        for (int e = 0; e < m->elen; e++) {
            const int *nd = m->el + e*nn;
            int has = 0;
            for (int k = 0; k < nn; k++) if (nd[k] == nmin) has = 1;
            if (!has) continue;
            for (int s = 0; s < m->nsd; s++) {
                int containsmin = 0, isbnd = 1;
                for (int k = 0; k < m->nsn; k++) {
                    int v = nd[dim == 3 ? TET_SN[s][k] : TRI_SN[s][k]];
                    if (v == nmin) containsmin = 1;
                    else if (!m->bnd[v]) { isbnd = 0; break; }
                }
                if (!containsmin || !isbnd) continue;
                double gc[3] = {0, 0, 0}, dc[3];
                for (int k = 0; k < m->nsn; k++)
                    for (int d = 0; d < dim; d++) gc[d] += m->x[(size_t)nd[dim == 3 ? TET_SN[s][k] : TRI_SN[s][k]]*dim + d] / m->nsn;
                /* outward direction cosine: minus the gradient direction of the
                 * shape function of the node opposite the side (tri3.cc:155-163, tet4.cc:230-241) */
                static const int opp3[4] = {3, 2, 1, 0}, opp2[3] = {2, 0, 1};
                int o = dim == 3 ? opp3[s] : opp2[s];
                dc[0] = -m->gb[e*nn + o]; dc[1] = -m->gc[e*nn + o]; dc[2] = dim == 3 ? -m->gd[e*nn + o] : 0.0;
                double len = 0.0;
                for (int d = 0; d < dim; d++) len += dc[d]*dc[d];
                len = sqrt(len);
                for (int d = 0; d < dim; d++) dc[d] /= len;
                double dist = pdist(p, gc, dim);
                if (dist < EPS) dist = EPS;
                double w = 1.0/dist;
                if (w > wmax) {
                    wmax = w;
                    double pd = 0.0;
                    for (int d = 0; d < dim; d++) pd += (p[d] - gc[d]) * dc[d];
                    for (int d = 0; d < dim; d++) pshift[d] = p[d] - dc[d]*pd;
                    found = 1;
                }
            }
        }
        return found;
    }
}

om_qm *om_qm_load(const om_mesh *m, const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return NULL;
    char line[4096], flag[128];
    om_qm *qm = (om_qm*)xcalloc(1, sizeof(om_qm));
    if (!fgets(line, sizeof line, f) || strncmp(line, "QM file", 7)) { fclose(f); return NULL; }
    if (!fgets(line, sizeof line, f)) { fclose(f); return NULL; }
    qm->dim = atoi(line+9);
    int fixedq = 0, fixedm = 0;
    while (fgets(line, sizeof line, f)) if (!strncmp(line, "SourceList", 10)) break;
    flag[0] = 0; int ni = sscanf(line+10, "%d%127s", &qm->nq, flag);
    fixedq = (ni > 1 && !strcasecmp(flag, "fixed")); (void)fixedq;
    qm->Q = (double*)xcalloc((size_t)qm->nq*qm->dim, sizeof(double));
    for (int i = 0; i < qm->nq; i++) {
        if (!fgets(line, sizeof line, f)) { fclose(f); return NULL; }
        char *p = line, *e;
        for (int d = 0; d < qm->dim; d++) { qm->Q[i*qm->dim+d] = strtod(p, &e); p = e; }
    }
    while (fgets(line, sizeof line, f)) if (!strncmp(line, "MeasurementList", 15)) break;
    flag[0] = 0; ni = sscanf(line+15, "%d%127s", &qm->nm, flag);
    fixedm = (ni > 1 && !strcasecmp(flag, "fixed"));
    qm->M = (double*)xcalloc((size_t)qm->nm*qm->dim, sizeof(double));
    for (int i = 0; i < qm->nm; i++) {
        if (!fgets(line, sizeof line, f)) { fclose(f); return NULL; }
        char *p = line, *e;
        for (int d = 0; d < qm->dim; d++) { qm->M[i*qm->dim+d] = strtod(p, &e); p = e; }
    }
    while (fgets(line, sizeof line, f)) if (!strncmp(line, "LinkList", 8)) break;
    qm->lptr = (int*)xcalloc(qm->nq+1, sizeof(int));
    int cap = qm->nq*qm->nm + 1;
    qm->lidx = (int*)xcalloc(cap, sizeof(int));
    for (int i = 0; i < qm->nq; i++) {
        int cnt; if (fscanf(f, "%d", &cnt) != 1) { fclose(f); return NULL; }
        int c; do { c = fgetc(f); } while (c != ':' && c != EOF);
        qm->lptr[i+1] = qm->lptr[i] + cnt;
        for (int j = 0; j < cnt; j++) if (fscanf(f, "%d", &qm->lidx[qm->lptr[i]+j]) != 1) { fclose(f); return NULL; }
    }
    qm->nqm = qm->lptr[qm->nq];
    fclose(f);
    if (!fixedm) {
        for (int i = 0; i < qm->nm; i++) {
            double ps[3];
            if (pull_to_boundary(m, qm->M + i*qm->dim, ps)) memcpy(qm->M + i*qm->dim, ps, sizeof(double)*qm->dim);
        }
    }
    return qm;
}

int om_qm_nq(const om_qm *q) { return q->nq; }
int om_qm_nm(const om_qm *q) { return q->nm; }
int om_qm_nqm(const om_qm *q) { return q->nqm; }
void om_qm_get(const om_qm *q, double *Q, double *M, int *lptr, int *lidx)
{
    memcpy(Q, q->Q, sizeof(double)*(size_t)q->nq*q->dim);
    memcpy(M, q->M, sizeof(double)*(size_t)q->nm*q->dim);
    memcpy(lptr, q->lptr, sizeof(int)*(size_t)(q->nq+1));
    memcpy(lidx, q->lidx, sizeof(int)*(size_t)q->nqm);
}
void om_qm_free(om_qm *q) { if (!q) return; free(q->Q); free(q->M); free(q->lptr); free(q->lidx); free(q); }

/* Pull one point (exposed for tests) */
int om_pull_to_boundary(const om_mesh *m, const double *p, double *out) { return pull_to_boundary(m, p, out); }

/* ------------------------------------------------------------------------
 * Krylov solvers.  CSR SpMV (crmatrix_imp.hpp:1173-1191); diagonal
 * preconditioner (precon_imp.hpp:56-80: 1/diag, 1 where diag == 0).
 * ------------------------------------------------------------------------ */
static void spmv_r(int n, const int *rp, const int *ci, const double *v, const double *x, double *y)
{
    for (int r = 0; r < n; r++) {
        double br = 0.0;
        for (int i = rp[r]; i < rp[r+1]; i++) br += v[i] * x[ci[i]];
        y[r] = br;
    }
}

static double dot_r(int n, const double *a, const double *b) { double s = 0.0; for (int i = 0; i < n; i++) s += a[i]*b[i]; return s; }

/* PCG (matrix_imp.hpp:93-154); val is real (one double per nonzero) */
int om_pcg(int n, const int *rp, const int *ci, const double *val, const double *b, double *x,
           double tol, int maxit, int diag, double *relerr)
{
    double *r = (double*)xcalloc(n, 8), *d = (double*)xcalloc(n, 8), *q = (double*)xcalloc(n, 8), *id = (double*)xcalloc(n, 8);
    for (int i = 0; i < n; i++) {
        double dg = 0.0;
        for (int k = rp[i]; k < rp[i+1]; k++) if (ci[k] == i) dg = val[k];
        id[i] = dg != 0.0 ? 1.0/dg : 1.0;
    }
    if (!maxit) maxit = n+1;
    spmv_r(n, rp, ci, val, x, r);
    for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
    for (int i = 0; i < n; i++) d[i] = diag ? r[i]*id[i] : r[i];
    double dnew = dot_r(n, r, d), dold, alpha, beta, err = 0.0;
    double bnorm = sqrt(dot_r(n, b, b));
    int it;
    for (it = 0; it < maxit; it++) {
        spmv_r(n, rp, ci, val, d, q);
        alpha = dnew / dot_r(n, d, q);
        if (alpha == 0.0 || isnan(alpha)) break;
        for (int i = 0; i < n; i++) x[i] += d[i]*alpha;
        for (int i = 0; i < n; i++) r[i] -= q[i]*alpha;
        dold = dnew;
        if (diag) {
            for (int i = 0; i < n; i++) q[i] = r[i]*id[i];
            dnew = dot_r(n, r, q);
            beta = dnew/dold;
            for (int i = 0; i < n; i++) d[i] = d[i]*beta + q[i];
        } else {
            dnew = dot_r(n, r, r);
            beta = dnew/dold;
            for (int i = 0; i < n; i++) d[i] = d[i]*beta + r[i];
        }
        if (beta == 0.0 || isnan(beta)) break;
        err = sqrt(dot_r(n, r, r));
        if (err <= tol*bnorm) break;
    }
    if (relerr) *relerr = err/bnorm;
    free(r); free(d); free(q); free(id);
    return it;
}

/* complex helpers on interleaved arrays */
static void spmv_c(int n, const int *rp, const int *ci, const double *v, const double *x, double *y)
{
    for (int r = 0; r < n; r++) {
        double br = 0.0, bi = 0.0;
        for (int i = rp[r]; i < rp[r+1]; i++) {
            double vr = v[2*i], vi = v[2*i+1], xr = x[2*ci[i]], xi = x[2*ci[i]+1];
            br += vr*xr - vi*xi; bi += vr*xi + vi*xr;
        }
        y[2*r] = br; y[2*r+1] = bi;
    }
}
static double dotR_c(int n, const double *a, const double *b) { double s = 0.0; for (int i = 0; i < n; i++) s += a[2*i]*b[2*i] + a[2*i+1]*b[2*i+1]; return s; }
static void cmul(const double *a, const double *b, double *o) { double r = a[0]*b[0]-a[1]*b[1], i = a[0]*b[1]+a[1]*b[0]; o[0] = r; o[1] = i; }

/* BiCGSTAB<complex<double>> (matrix_imp.hpp:350-413): real inner products */
int om_bicgstab_c(int n, const int *rp, const int *ci, const double *val, const double *b, double *x,
                  double tol, int maxit, int diag, double *relerr)
{
    size_t N = 2*(size_t)n*sizeof(double);
    double *r = (double*)xcalloc(2*n, 8), *rd = (double*)xcalloc(2*n, 8), *p = (double*)xcalloc(2*n, 8), *pd = (double*)xcalloc(2*n, 8);
    double *v = (double*)xcalloc(2*n, 8), *s = (double*)xcalloc(2*n, 8), *sd = (double*)xcalloc(2*n, 8), *t = (double*)xcalloc(2*n, 8);
    double *id = (double*)xcalloc(2*n, 8);
    for (int i = 0; i < n; i++) {
        double dr = 0.0, di = 0.0;
        for (int k = rp[i]; k < rp[i+1]; k++) if (ci[k] == i) { dr = val[2*k]; di = val[2*k+1]; }
        if (dr != 0.0 || di != 0.0) { double den = dr*dr + di*di; id[2*i] = dr/den; id[2*i+1] = -di/den; }
        else { id[2*i] = 1.0; id[2*i+1] = 0.0; }
    }
    if (!maxit) maxit = n+1;
    spmv_c(n, rp, ci, val, x, r);
    for (int i = 0; i < 2*n; i++) r[i] = b[i] - r[i];
    memcpy(rd, r, N);
    double bnorm = sqrt(dotR_c(n, b, b));
    double rho = 0.0, rhop, alpha = 0.0, beta, omega = 0.0, aden, onum, oden, err = 0.0;
    int it;
    for (it = 0; it < maxit; it++) {
        rhop = rho;
        rho = dotR_c(n, rd, r);
        if (rho == 0.0) break;
        if (!it) memcpy(p, r, N);
        else {
            beta = (rho*alpha)/(rhop*omega);
            for (int i = 0; i < 2*n; i++) p[i] = r[i] + (p[i]-v[i]*omega)*beta;
        }
        if (diag) for (int i = 0; i < n; i++) cmul(p+2*i, id+2*i, pd+2*i); else memcpy(pd, p, N);
        spmv_c(n, rp, ci, val, pd, v);
        aden = dotR_c(n, rd, v);
        alpha = rho/aden;
        for (int i = 0; i < 2*n; i++) s[i] = r[i] - v[i]*alpha;
        if (diag) for (int i = 0; i < n; i++) cmul(s+2*i, id+2*i, sd+2*i); else memcpy(sd, s, N);
        spmv_c(n, rp, ci, val, sd, t);
        onum = dotR_c(n, t, s); oden = dotR_c(n, t, t);
        omega = onum/oden;
        for (int i = 0; i < 2*n; i++) { x[i] += pd[i]*alpha + sd[i]*omega; r[i] = s[i] - t[i]*omega; }
        err = sqrt(dotR_c(n, r, r));
        if (err <= tol*bnorm) break;
        if (omega == 0.0) break;
    }
    if (relerr) *relerr = err/bnorm;
    free(r); free(rd); free(p); free(pd); free(v); free(s); free(sd); free(t); free(id);
    return it;
}

/* FP64 residual norm ||b - A x|| / ||b|| (complex interleaved) */
double om_relres_c(int n, const int *rp, const int *ci, const double *val, const double *b, const double *x)
{
    double *y = (double*)xcalloc(2*n, 8);
    spmv_c(n, rp, ci, val, x, y);
    double num = 0.0, den = 0.0;
    for (int i = 0; i < 2*n; i++) { double d = b[i]-y[i]; num += d*d; den += b[i]*b[i]; }
    free(y);
    return sqrt(num/den);
}

/* ------------------------------------------------------------------------
 * ProjectAll (fwdsolver.cc:1415-1447): proj[Qofs[q]+i] = dot(phi_q, mvec_m)
 * unconjugated (vector_imp.hpp:1045-1052), then complex log if logscale.
 * phi: nq x n complex, mvec: nm x n complex (dense rows).
 * ------------------------------------------------------------------------ */
void om_project(int n, int nq, int nm, const double *phi, const double *mvec, const int *lptr, const int *lidx,
                int logscale, double *proj)
{
    (void)nm;
    for (int q = 0; q < nq; q++)
        for (int k = lptr[q]; k < lptr[q+1]; k++) {
            int mi = lidx[k];
            double sr = 0.0, si = 0.0;
            const double *f = phi + 2*(size_t)q*n, *g = mvec + 2*(size_t)mi*n;
            for (int j = 0; j < n; j++) { sr += f[2*j]*g[2*j] - f[2*j+1]*g[2*j+1]; si += f[2*j]*g[2*j+1] + f[2*j+1]*g[2*j]; }
            if (logscale) { double c = log(hypot(sr, si)), a = atan2(si, sr); sr = c; si = a; }
            proj[2*k] = sr; proj[2*k+1] = si;
        }
}

/* ------------------------------------------------------------------------
 * Nodal Jacobian (jacobian.cc:441-497 with IntFG :806-835 and
 * IntGradFGradG :839-872).  dphi: nq x n, aphi: nm x n (complex), proj: nQM
 * complex linear projections; logscale divides by proj.  J: 2nQM x 2n
 * row-major.  Rows: q ascending, linked m ascending in link-list order.
 * ------------------------------------------------------------------------ */
void om_jacobian_nodal(const om_mesh *m, int nq, int nm, const double *dphi, const double *aphi,
                       const int *lptr, const int *lidx, const double *proj, int logscale, double *J)
{
    (void)nm;
    int n = m->n, nn = m->nn, nqm = lptr[nq];
    double *pm = (double*)xcalloc(2*(size_t)n, 8), *pk = (double*)xcalloc(2*(size_t)n, 8);
    size_t rowlen = 2*(size_t)n;
    for (int q = 0; q < nq; q++)
        for (int k = lptr[q]; k < lptr[q+1]; k++) {
            const double *f = dphi + 2*(size_t)q*n, *g = aphi + 2*(size_t)lidx[k]*n;
            memset(pm, 0, 16*(size_t)n); memset(pk, 0, 16*(size_t)n);
            for (int e = 0; e < m->elen; e++) {
                const int *nd = m->el + e*nn;
                for (int i = 0; i < nn; i++) {
                    int bs = nd[i];
                    for (int j = 0; j < nn; j++) {
                        int nj = nd[j];
                        double fg[2], t[2], sm[2], sk[2];
                        cmul(f+2*nj, g+2*nj, fg);
                        double w = el_intfff(m, e, i, j, j), wd = el_intfdd(m, e, i, j, j);
                        sm[0] = fg[0]*w; sm[1] = fg[1]*w; sk[0] = fg[0]*wd; sk[1] = fg[1]*wd;
                        for (int kk = 0; kk < j; kk++) {
                            int nk = nd[kk];
                            double a1[2], a2[2];
                            cmul(f+2*nj, g+2*nk, a1); cmul(f+2*nk, g+2*nj, a2);
                            t[0] = a1[0]+a2[0]; t[1] = a1[1]+a2[1];
                            w = el_intfff(m, e, i, j, kk); wd = el_intfdd(m, e, i, j, kk);
                            sm[0] += t[0]*w; sm[1] += t[1]*w; sk[0] += t[0]*wd; sk[1] += t[1]*wd;
                        }
                        pm[2*bs] += sm[0]; pm[2*bs+1] += sm[1];
                        pk[2*bs] += sk[0]; pk[2*bs+1] += sk[1];
                    }
                }
            }
            double pr = proj[2*k], pi = proj[2*k+1], den = pr*pr + pi*pi;
            double *R0 = J + (size_t)k*rowlen, *R1 = J + (size_t)(k+nqm)*rowlen;
            for (int i = 0; i < n; i++) {
                double a[2] = {pm[2*i], pm[2*i+1]}, b[2] = {pk[2*i], pk[2*i+1]};
                if (logscale) {   /* complex division by proj */
                    double ar = (a[0]*pr + a[1]*pi)/den, ai = (a[1]*pr - a[0]*pi)/den;
                    double br = (b[0]*pr + b[1]*pi)/den, bi = (b[1]*pr - b[0]*pi)/den;
                    a[0] = ar; a[1] = ai; b[0] = br; b[1] = bi;
                }
                R0[i] = -a[0]; R1[i] = -a[1]; R0[i+n] = -b[0]; R1[i+n] = -b[1];
            }
        }
    free(pm); free(pk);
}

/* ------------------------------------------------------------------------
 * Raster_Pixel with basis grid == intermediate grid (raster.cc:18-100,
 * raster_px.cc:10-54): grid points p = bbmin + i*dW, dW = (bbmax-bbmin)/(g-1).
 * ------------------------------------------------------------------------ */

/* Mesh::BoundingBox (pad 0): node coordinate extremes */
void om_bbox(const om_mesh *m, double *bbmin, double *bbmax)
{
    for (int d = 0; d < m->dim; d++) { bbmin[d] = 1e300; bbmax[d] = -1e300; }
    for (int i = 0; i < m->n; i++)
        for (int d = 0; d < m->dim; d++) {
            double v = m->x[(size_t)i*m->dim + d];
            if (v < bbmin[d]) bbmin[d] = v;
            if (v > bbmax[d]) bbmax[d] = v;
        }
}

/* Element_Unstructured::GContains (element.cc:553-563): element bounding box
 * test with EPS 1e-8, then Local + LContains (pad = true) */
static int el_gcontains(const om_mesh *m, int e, const double *p)
{
    const double EPS = 1e-8;
    int dim = m->dim, nn = m->nn;
    for (int d = 0; d < dim; d++) {
        double lo = 1e300, hi = -1e300;
        for (int i = 0; i < nn; i++) {
            double v = m->x[(size_t)m->el[e*nn+i]*dim + d];
            if (v < lo) lo = v;
            if (v > hi) hi = v;
        }
        if (p[d] < lo-EPS || p[d] > hi+EPS) return 0;
    }
    double loc[3];
    el_local(m, e, p, loc);
    return el_lcontains(dim, loc);
}

/* GenerateElementPixelRef (mesh.cc:2944-3015): for every grid point the first
 * element (ascending) whose EPS-inflated bounding box range holds it and
 * which contains it; -1 outside the mesh */
void om_raster_elref(const om_mesh *m, const int *gdim, const double *bbmin, const double *bbmax, int *elref)
{
    const double EPS = 1e-6;
    int dim = m->dim, nn = m->nn;
    int nx = gdim[0], ny = gdim[1], nz = dim > 2 ? gdim[2] : 1;
    double dW[3];
    for (int d = 0; d < dim; d++) dW[d] = (bbmax[d]-bbmin[d])/(double)(gdim[d]-1);
    for (long i = 0; i < (long)nx*ny*nz; i++) elref[i] = -1;
    for (int e = 0; e < m->elen; e++) {
        double lo[3], hi[3];
        int rmin[3] = {0, 0, 0}, rmax[3] = {1, 1, 1};
        for (int d = 0; d < dim; d++) { lo[d] = 1e10; hi[d] = -1e10; }
        for (int i = 0; i < nn; i++)
            for (int d = 0; d < dim; d++) {
                double v = m->x[(size_t)m->el[e*nn+i]*dim + d];
                if (v < lo[d]) lo[d] = v;
                if (v > hi[d]) hi[d] = v;
            }
        int overlap = 1;
        for (int d = 0; d < dim; d++) {
            lo[d] -= EPS; hi[d] += EPS;
            if (lo[d] > bbmax[d] || hi[d] < bbmin[d]) { overlap = 0; break; }
            int a = (int)ceil((lo[d]-bbmin[d])/dW[d]), b = (int)floor((hi[d]-bbmin[d])/dW[d]) + 1;
            rmin[d] = a > 0 ? a : 0;
            rmax[d] = b < gdim[d] ? b : gdim[d];
        }
        if (!overlap) continue;
        double p[3];
        for (int iz = rmin[2]; iz < rmax[2]; iz++) {
            if (dim > 2) p[2] = iz*dW[2] + bbmin[2];
            for (int iy = rmin[1]; iy < rmax[1]; iy++) {
                p[1] = iy*dW[1] + bbmin[1];
                for (int ix = rmin[0]; ix < rmax[0]; ix++) {
                    long idx = ix + (long)nx*(iy + (long)ny*iz);
                    if (elref[idx] >= 0) continue;
                    p[0] = ix*dW[0] + bbmin[0];
                    if (el_gcontains(m, e, p)) elref[idx] = e;
                }
            }
        }
    }
}

/* NodeMapMatrix (mesh.cc:2695-2853): per mesh node the bi/trilinear
 * weights of the grid points around it that lie inside the mesh (elref >= 0),
 * corners x fastest, renormalised over the present ones when some are
 * missing (equal weights if their sum is at most 1e-8); no entries for a
 * node outside the bounding box.  cols/w: n x 2^dim, unused slots -1 / 0. */
void om_raster_nodemap(const om_mesh *m, const int *gdim, const double *bbmin, const double *bbmax, const int *elref,
                       int *cols, double *w)
{
    const double EPS = 1e-8;
    int dim = m->dim, nc = 1 << dim;
    double dW[3];
    for (int d = 0; d < dim; d++) dW[d] = (bbmax[d]-bbmin[d])/(double)(gdim[d]-1);
    for (int i = 0; i < m->n; i++) {
        int *oc = cols + (size_t)i*nc;
        double *ow = w + (size_t)i*nc;
        for (int k = 0; k < nc; k++) { oc[k] = -1; ow[k] = 0.0; }
        int i0[3] = {0, 0, 0}, ok = 1;
        double d0[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < dim; d++) {
            double p = m->x[(size_t)i*dim + d];
            if (p >= bbmin[d] && p <= bbmax[d]) {
                i0[d] = (int)((p-bbmin[d])/dW[d]);
                if (i0[d] < 0) i0[d] = 0;
                else if (i0[d] >= gdim[d]-1) i0[d] = gdim[d]-2;
                d0[d] = (p - (i0[d]*dW[d] + bbmin[d]))/dW[d];
                if (d0[d] < 0.0) d0[d] = 0.0;
                else if (d0[d] > 1.0) d0[d] = 1.0;
            } else { ok = 0; break; }
        }
        if (!ok) continue;
        int cnt = 0;
        for (int k = 0; k < nc; k++) {
            int cx = k & 1, cy = (k >> 1) & 1, cz = (k >> 2) & 1;
            long pix = (i0[0]+cx) + (long)gdim[0]*((i0[1]+cy) + (dim > 2 ? (long)gdim[1]*(i0[2]+cz) : 0));
            if (elref[pix] < 0) continue;
            double v = (cx ? d0[0] : (1.0-d0[0])) * (cy ? d0[1] : (1.0-d0[1]));
            if (dim > 2) v = v * (cz ? d0[2] : (1.0-d0[2]));
            oc[cnt] = (int)pix; ow[cnt] = v; cnt++;
        }
        if (cnt > 0 && cnt < nc) {
            double weight = 0.0;
            for (int k = 0; k < cnt; k++) weight += ow[k];
            if (weight > EPS) for (int k = 0; k < cnt; k++) ow[k] /= weight;
            else for (int k = 0; k < cnt; k++) ow[k] = 1.0/(double)cnt;
        }
    }
}

/* GridMapMatrix (mesh.cc:2486-2554): row per grid point with the element's
 * global shape functions at the point (tet4.cc:302-313, tri3.cc:231-241);
 * cols/w: glen x nn, col -1 / weight 0 where the point is outside. */
void om_raster_bmap(const om_mesh *m, const int *gdim, const double *bbmin, const double *bbmax, const int *elref,
                    int *cols, double *w)
{
    int dim = m->dim, nn = m->nn;
    int nx = gdim[0], ny = gdim[1], nz = dim > 2 ? gdim[2] : 1;
    double dW[3];
    for (int d = 0; d < dim; d++) dW[d] = (bbmax[d]-bbmin[d])/(double)(gdim[d]-1);
    long i = 0;
    double p[3];
    for (int iz = 0; iz < nz; iz++) {
        if (dim > 2) p[2] = iz*dW[2] + bbmin[2];
        for (int iy = 0; iy < ny; iy++) {
            p[1] = iy*dW[1] + bbmin[1];
            for (int ix = 0; ix < nx; ix++, i++) {
                for (int j = 0; j < nn; j++) { cols[i*nn+j] = -1; w[i*nn+j] = 0.0; }
                int e = elref[i];
                if (e < 0) continue;
                p[0] = ix*dW[0] + bbmin[0];
                const double *a = m->ga + e*nn, *b = m->gb + e*nn, *c = m->gc + e*nn, *dd = m->gd + e*nn;
                double scale = dim == 3 ? 1.0/(6.0*m->size[e]) : 1.0/(2.0*m->size[e]);
                for (int j = 0; j < nn; j++) {
                    cols[i*nn+j] = m->el[e*nn+j];
                    w[i*nn+j] = dim == 3 ? scale * (a[j] + b[j]*p[0] + c[j]*p[1] + dd[j]*p[2])
                                         : scale * (a[j] + b[j]*p[0] + c[j]*p[1]);
                }
            }
        }
    }
}

/* ImageGradient (jacobian.cc:944-1060) of a complex grid image; spacing
 * gsize/n per axis; mask = elref.  grad: dim x glen complex */
void om_image_gradient(int dim, const int *gdim, const double *gsize, const double *im, const int *mask, double *grad)
{
    int nx = gdim[0], ny = gdim[1], nz = dim > 2 ? gdim[2] : 1;
    long n = (long)nx*ny*nz;
    long stride[3] = {1, nx, (long)nx*ny};
    int len[3] = {nx, ny, nz};
    for (int d = 0; d < dim; d++) {
        double h = gsize[d]/len[d], ih = 1.0/h, i2h = 0.5*ih;
        double *g = grad + 2*(size_t)d*n;
        long s = stride[d];
        for (long idx = 0; idx < n; idx++) {
            int pos = (int)((idx / s) % len[d]);
            int bcnt = mask[idx] >= 0;
            int blo = pos > 0 && mask[idx-s] >= 0;
            int bhi = pos < len[d]-1 && mask[idx+s] >= 0;
            double *o = g + 2*idx;
            if (blo && bhi) { o[0] = (im[2*(idx+s)]-im[2*(idx-s)]) * i2h; o[1] = (im[2*(idx+s)+1]-im[2*(idx-s)+1]) * i2h; }
            else if (!bcnt) { o[0] = 0.0; o[1] = 0.0; }
            else if (blo) { o[0] = (im[2*idx]-im[2*(idx-s)]) * ih; o[1] = (im[2*idx+1]-im[2*(idx-s)+1]) * ih; }
            else if (bhi) { o[0] = (im[2*(idx+s)]-im[2*idx]) * ih; o[1] = (im[2*(idx+s)+1]-im[2*idx+1]) * ih; }
            else { o[0] = 0.0; o[1] = 0.0; }
        }
    }
}

/* Raster::Map_MeshToGrid (raster.cc:188-199, Ax_cplx of the real B) */
static void map_mesh_to_grid(long glen, int nn, const int *cols, const double *w, const double *f, double *out)
{
    for (long i = 0; i < glen; i++) {
        double re = 0.0, im = 0.0;
        for (int j = 0; j < nn; j++) {
            if (cols[i*nn+j] < 0) continue;
            re += w[i*nn+j] * f[2*(size_t)cols[i*nn+j]];
            im += w[i*nn+j] * f[2*(size_t)cols[i*nn+j]+1];
        }
        out[2*i] = re; out[2*i+1] = im;
    }
}

/* ---------------------------------------------------------------------
 * Linear pixel basis coarser than the intermediate grid (Raster_Pixel with
 * bdim != gdim, raster_px.cc:10-35).  Basis voxel ix (of bx per axis, at
 * x = ix/(bx-1) in unit box coordinates) covers the grid points xx of
 * [floor(x0 (gx-1)), ceil(x1 (gx-1))], x0/x1 its neighbours (0 and 1 at the
 * ends), with d = |xx/(gx-1) - x| (bx-1) < 1 per axis; the tent weight is
 * prod (1 - d).  Loops, operation order and the elref mask follow
 * SubsampleLinPixel (mesh.cc:3047-3136), Grid2LinPixMatrix (:3138-3288,
 * rows normalised by their weight sum) and LinPix2GridMatrix (:3540-3718,
 * unnormalised weights, one row per grid point). */

typedef struct { double x, nb, dg; int g0, g1; } lp_axis;

static lp_axis lp_axis_of(int b, int nbv, int ng)
{
    lp_axis a;
    double db = 1.0/(nbv-1.0);
    a.nb = nbv-1.0;
    a.dg = 1.0/(ng-1.0);
    a.x = (double)b*db;
    double x0 = (b ? (b-1)*db : 0.0);
    double x1 = (b < nbv-1 ? (b+1)*db : 1.0);
    a.g0 = (int)floor(x0*(ng-1));
    a.g1 = (int)ceil(x1*(ng-1));
    return a;
}

/* grid points and tent weights of basis voxel b in the reference's loop
 * order; returns their count (gi, t may be NULL to count only) */
static int lp_row(int dim, const int *gdim, const int *bdim, const int *elref, long b, long *gi, double *t)
{
    int gx = gdim[0], gy = gdim[1];
    int bx = bdim[0], by = bdim[1];
    int ix = (int)(b % bx), iy = (int)((b / bx) % by), iz = dim > 2 ? (int)(b / ((long)bx*by)) : 0;
    lp_axis ax = lp_axis_of(ix, bx, gx), ay = lp_axis_of(iy, by, gy);
    int cnt = 0;
    if (dim == 2) {
        for (int yy = ay.g0; yy <= ay.g1; yy++) {
            double dy = fabs(yy*ay.dg - ay.x)*ay.nb;
            if (dy >= 1.0) continue;
            double idy = 1.0-dy;
            for (int xx = ax.g0; xx <= ax.g1; xx++) {
                long gidx = xx + (long)yy*gx;
                if (elref && elref[gidx] < 0) continue;
                double dx = fabs(xx*ax.dg - ax.x)*ax.nb;
                if (dx >= 1.0) continue;
                if (gi) { gi[cnt] = gidx; t[cnt] = (1.0-dx)*idy; }
                cnt++;
            }
        }
    } else {
        lp_axis az = lp_axis_of(iz, bdim[2], gdim[2]);
        for (int zz = az.g0; zz <= az.g1; zz++) {
            double dz = fabs(zz*az.dg - az.x)*az.nb;
            if (dz >= 1.0) continue;
            double idz = 1.0-dz;
            for (int yy = ay.g0; yy <= ay.g1; yy++) {
                double dy = fabs(yy*ay.dg - ay.x)*ay.nb;
                if (dy >= 1.0) continue;
                double idyz = (1.0-dy)*idz;
                long gidx = ax.g0 + ((long)yy + (long)zz*gy)*gx;
                for (int xx = ax.g0; xx <= ax.g1; xx++, gidx++) {
                    if (elref && elref[gidx] < 0) continue;
                    double dx = fabs(xx*ax.dg - ax.x)*ax.nb;
                    if (dx >= 1.0) continue;
                    if (gi) { gi[cnt] = gidx; t[cnt] = (1.0-dx)*idyz; }
                    cnt++;
                }
            }
        }
    }
    return cnt;
}

static long lp_len(int dim, const int *d) { long n = 1; for (int i = 0; i < dim; i++) n *= d[i]; return n; }

static int lp_maxrow(int dim, const int *gdim, const int *bdim)
{
    int m = 1;
    for (int d = 0; d < dim; d++) m *= 2*((gdim[d]-1)/(bdim[d]-1) + 2) + 1;
    return m;
}

/* SubsampleLinPixel (mesh.cc:3047-3136): bf[b] = sum t gf / sum t over the
 * grid points of basis voxel b (0 where there are none); elref may be NULL */
void om_subsample_linpixel(int dim, const int *gdim, const int *bdim, const int *elref, const double *gf, double *bf)
{
    long blen = lp_len(dim, bdim);
    int mx = lp_maxrow(dim, gdim, bdim);
    long *gi = (long*)xcalloc(mx, sizeof(long));
    double *t = (double*)xcalloc(mx, sizeof(double));
    for (long b = 0; b < blen; b++) {
        int c = lp_row(dim, gdim, bdim, elref, b, gi, t);
        double sum = 0.0, wgt = 0.0;
        for (int k = 0; k < c; k++) { wgt += t[k]; sum += t[k]*gf[gi[k]]; }
        bf[b] = wgt ? sum/wgt : 0.0;
    }
    free(gi); free(t);
}

/* Grid2LinPixMatrix (mesh.cc:3138-3288): blen x glen CSR, row b the
 * normalised tent weights of basis voxel b.  With ci == NULL only rp
 * (blen+1) is filled. */
void om_grid2linpix(int dim, const int *gdim, const int *bdim, const int *elref, int *rp, int *ci, double *val)
{
    long blen = lp_len(dim, bdim);
    int mx = lp_maxrow(dim, gdim, bdim);
    long *gi = (long*)xcalloc(mx, sizeof(long));
    double *t = (double*)xcalloc(mx, sizeof(double));
    rp[0] = 0;
    for (long b = 0; b < blen; b++) {
        int c = lp_row(dim, gdim, bdim, elref, b, gi, t);
        rp[b+1] = rp[b] + c;
        if (!ci) continue;
        double wgt = 0.0;
        for (int k = 0; k < c; k++) { ci[rp[b]+k] = (int)gi[k]; val[rp[b]+k] = t[k]; wgt += t[k]; }
        for (int k = 0; k < c; k++) val[rp[b]+k] = wgt ? val[rp[b]+k]/wgt : 0.0;
    }
    free(gi); free(t);
}

/* LinPix2GridMatrix (mesh.cc:3540-3718): glen x blen CSR, row g the
 * (unnormalised) tent weights of the basis voxels covering grid point g,
 * basis voxels ascending.  With ci == NULL only rp (glen+1) is filled. */
void om_linpix2grid(int dim, const int *gdim, const int *bdim, const int *elref, int *rp, int *ci, double *val)
{
    long blen = lp_len(dim, bdim), glen = lp_len(dim, gdim);
    int mx = lp_maxrow(dim, gdim, bdim);
    long *gi = (long*)xcalloc(mx, sizeof(long));
    double *t = (double*)xcalloc(mx, sizeof(double));
    int *len = (int*)xcalloc(glen, sizeof(int));
    for (long b = 0; b < blen; b++) {
        int c = lp_row(dim, gdim, bdim, elref, b, gi, t);
        for (int k = 0; k < c; k++) len[gi[k]]++;
    }
    rp[0] = 0;
    for (long g = 0; g < glen; g++) rp[g+1] = rp[g] + len[g];
    if (ci) {
        for (long g = 0; g < glen; g++) len[g] = 0;
        for (long b = 0; b < blen; b++) {
            int c = lp_row(dim, gdim, bdim, elref, b, gi, t);
            for (int k = 0; k < c; k++) {
                long g = gi[k];
                int idx = rp[g] + len[g]++;
                ci[idx] = (int)b;
                val[idx] = t[k];
            }
        }
    }
    free(gi); free(t); free(len);
}

/* Map_GridToSol of Raster_Pixel (raster_px.cc:102-114,148-151): with
 * bdim == gdim a gather of the grid points inside the mesh (G = NULL); else
 * Grid2LinPixMatrix then the basis voxels with support (SubsampleLinPixel
 * of the inside mask without elref, raster.cc:54-73).  sz: prod
 * gsize/(bdim-1), the voxel size of the Jacobians (jacobian.cc:368-372). */
typedef struct {
    int has_g;
    long slen, gslen;
    long *s2b;            /* slen solution voxels (grid points, or basis voxels with G) */
    long *g2s;            /* gslen grid points inside the mesh */
    int *rp, *ci;
    double *val, sz;
} om_gmap;

static void gmap_init(om_gmap *G, int dim, const int *gdim, const int *bdim, const double *bbmin,
                      const double *bbmax, const int *elref)
{
    memset(G, 0, sizeof *G);
    long glen = lp_len(dim, gdim);
    G->sz = 1.0;
    for (int d = 0; d < dim; d++) {
        G->sz *= (bbmax[d]-bbmin[d])/(bdim[d]-1.0);
        if (bdim[d] != gdim[d]) G->has_g = 1;
    }
    for (long i = 0; i < glen; i++) G->gslen += elref[i] >= 0;
    G->g2s = (long*)xcalloc((size_t)G->gslen, sizeof(long));
    for (long i = 0, s = 0; i < glen; i++) if (elref[i] >= 0) G->g2s[s++] = i;
    if (!G->has_g) {
        G->slen = G->gslen;
        G->s2b = (long*)xcalloc((size_t)G->slen, sizeof(long));
        memcpy(G->s2b, G->g2s, sizeof(long)*G->slen);
        return;
    }
    long blen = lp_len(dim, bdim);
    double *mask = (double*)xcalloc(glen, 8), *bs = (double*)xcalloc(blen, 8);
    for (long i = 0; i < glen; i++) if (elref[i] >= 0) mask[i] = 1.0;
    om_subsample_linpixel(dim, gdim, bdim, NULL, mask, bs);
    for (long b = 0; b < blen; b++) G->slen += bs[b] > 0.0;
    G->s2b = (long*)xcalloc((size_t)G->slen, sizeof(long));
    for (long b = 0, s = 0; b < blen; b++) if (bs[b] > 0.0) G->s2b[s++] = b;
    G->rp = (int*)xcalloc(blen+1, sizeof(int));
    om_grid2linpix(dim, gdim, bdim, elref, G->rp, NULL, NULL);
    G->ci = (int*)xcalloc(G->rp[blen], sizeof(int));
    G->val = (double*)xcalloc(G->rp[blen], sizeof(double));
    om_grid2linpix(dim, gdim, bdim, elref, G->rp, G->ci, G->val);
    free(mask); free(bs);
}

static void gmap_free(om_gmap *G)
{
    free(G->s2b); free(G->g2s); free(G->rp); free(G->ci); free(G->val);
}

/* solution voxel s of a complex grid image f (2 glen doubles) */
static void gmap_sol(const om_gmap *G, const double *f, long s, double *re, double *im)
{
    if (!G->has_g) { long g = G->s2b[s]; *re = f[2*g]; *im = f[2*g+1]; return; }
    long b = G->s2b[s];
    double br = 0.0, bi = 0.0;
    for (int k = G->rp[b]; k < G->rp[b+1]; k++) { br += f[2*(size_t)G->ci[k]]*G->val[k]; bi += f[2*(size_t)G->ci[k]+1]*G->val[k]; }
    *re = br; *im = bi;
}

long om_raster_slen(int dim, const int *gdim, const int *bdim, const double *bbmin, const double *bbmax,
                    const int *elref, long *s2b)
{
    om_gmap G;
    gmap_init(&G, dim, gdim, bdim, bbmin, bbmax, elref);
    long n = G.slen;
    if (s2b) memcpy(s2b, G.s2b, sizeof(long)*n);
    gmap_free(&G);
    return n;
}

/* GenerateJacobian_grid (jacobian.cc:293-373) for Raster_Pixel: PMDF_mua =
 * -d a, PMDF_kap = -sum_d grad_d d grad_d a (jacobian.cc:57-79) on the
 * grid, divided by the projection for log data, mapped into the solution
 * basis (Map_GridToSol), scaled by prod gsize/(bdim-1).  dphi: nq x n,
 * aphi: nm x n complex; J: (2 nqm) x (2 slen). */
void om_jacobian_grid_b(const om_mesh *m, const int *gdim, const int *bdim, const double *bbmin,
                        const double *bbmax, const int *elref, int nq, int nm, const double *dphi,
                        const double *aphi, const int *lptr, const int *lidx, const double *proj, int logscale,
                        double *J)
{
    int dim = m->dim, nn = m->nn;
    long glen = lp_len(dim, gdim);
    double gsize[3];
    for (int d = 0; d < dim; d++) gsize[d] = bbmax[d]-bbmin[d];
    om_gmap G;
    gmap_init(&G, dim, gdim, bdim, bbmin, bbmax, elref);
    int *cols = (int*)xcalloc((size_t)glen*nn, sizeof(int));
    double *w = (double*)xcalloc((size_t)glen*nn, sizeof(double));
    om_raster_bmap(m, gdim, bbmin, bbmax, elref, cols, w);
    long slen = G.slen;
    int nqm = lptr[nq];
    double *ca = (double*)xcalloc(2*(size_t)glen*nm, 8);
    for (int j = 0; j < nm; j++) map_mesh_to_grid(glen, nn, cols, w, aphi + 2*(size_t)j*m->n, ca + 2*(size_t)j*glen);
    double *cd = (double*)xcalloc(2*(size_t)glen, 8);
    double *gd = (double*)xcalloc(2*(size_t)glen*dim, 8), *ga = (double*)xcalloc(2*(size_t)glen*dim, 8);
    double *pmg = (double*)xcalloc(2*(size_t)glen, 8), *pkg = (double*)xcalloc(2*(size_t)glen, 8);
    size_t rowlen = 2*(size_t)slen;
    for (int q = 0; q < nq; q++) {
        map_mesh_to_grid(glen, nn, cols, w, dphi + 2*(size_t)q*m->n, cd);
        om_image_gradient(dim, gdim, gsize, cd, elref, gd);
        for (int k = lptr[q]; k < lptr[q+1]; k++) {
            const double *a = ca + 2*(size_t)lidx[k]*glen;
            om_image_gradient(dim, gdim, gsize, a, elref, ga);
            double pr = proj[2*k], pi = proj[2*k+1], den = pr*pr + pi*pi;
            /* PMDFs at the grid points inside the mesh (the only ones the
             * basis map reads) */
            for (long i = 0; i < G.gslen; i++) {
                long g = G.g2s[i];
                double pm[2], pk[2] = {0.0, 0.0}, t[2];
                cmul(cd + 2*g, a + 2*g, t);
                pm[0] = -t[0]; pm[1] = -t[1];
                for (int d = 0; d < dim; d++) {
                    cmul(gd + 2*((size_t)d*glen + g), ga + 2*((size_t)d*glen + g), t);
                    pk[0] -= t[0]; pk[1] -= t[1];
                }
                if (logscale) {
                    double ar = (pm[0]*pr + pm[1]*pi)/den, ai = (pm[1]*pr - pm[0]*pi)/den;
                    double br = (pk[0]*pr + pk[1]*pi)/den, bi = (pk[1]*pr - pk[0]*pi)/den;
                    pm[0] = ar; pm[1] = ai; pk[0] = br; pk[1] = bi;
                }
                pmg[2*g] = pm[0]; pmg[2*g+1] = pm[1]; pkg[2*g] = pk[0]; pkg[2*g+1] = pk[1];
            }
            double *R0 = J + (size_t)k*rowlen, *R1 = J + (size_t)(k+nqm)*rowlen;
            for (long s = 0; s < slen; s++) {
                double mr, mi, kr, ki;
                gmap_sol(&G, pmg, s, &mr, &mi);
                gmap_sol(&G, pkg, s, &kr, &ki);
                R0[s] = mr*G.sz; R1[s] = mi*G.sz; R0[s+slen] = kr*G.sz; R1[s+slen] = ki*G.sz;
            }
        }
    }
    free(cols); free(w); free(ca); free(cd); free(gd); free(ga); free(pmg); free(pkg);
    gmap_free(&G);
}

void om_jacobian_grid(const om_mesh *m, const int *gdim, const double *bbmin, const double *bbmax, const int *elref,
                      int nq, int nm, const double *dphi, const double *aphi, const int *lptr, const int *lidx,
                      const double *proj, int logscale, double *J)
{
    om_jacobian_grid_b(m, gdim, gdim, bbmin, bbmax, elref, nq, nm, dphi, aphi, lptr, lidx, proj, logscale, J);
}

/* GenerateJacobian_cw_grid (jacobian.cc:610-721) for Raster_Pixel,
 * followed by the log row scaling of GenerateJacobian_cw (:160-186): real
 * fields dphi (nq x n), aphi (nm x n); PMDF_mua = -d a, PMDF_kap = -sum_d
 * grad_d d grad_d a with the direct gradients masked by the element
 * reference and the adjoint gradients unmasked (ImageGradient with mask 0,
 * as the reference calls it there), mapped into the solution basis, times
 * prod gsize/(bdim-1), then each row times 1/proj for log data.
 * Jmua, Jkap: nqm x slen (either may be NULL). */
void om_jacobian_grid_cw_b(const om_mesh *m, const int *gdim, const int *bdim, const double *bbmin,
                           const double *bbmax, const int *elref, int nq, int nm, const double *dphi,
                           const double *aphi, const int *lptr, const int *lidx, const double *proj, int logscale,
                           double *Jmua, double *Jkap)
{
    int dim = m->dim, nn = m->nn;
    long glen = lp_len(dim, gdim);
    double gsize[3];
    for (int d = 0; d < dim; d++) gsize[d] = bbmax[d]-bbmin[d];
    om_gmap G;
    gmap_init(&G, dim, gdim, bdim, bbmin, bbmax, elref);
    int *cols = (int*)xcalloc((size_t)glen*nn, sizeof(int));
    double *w = (double*)xcalloc((size_t)glen*nn, sizeof(double));
    om_raster_bmap(m, gdim, bbmin, bbmax, elref, cols, w);
    long slen = G.slen;
    int *nomask = (int*)xcalloc((size_t)glen, sizeof(int));   /* all inside */
    /* the image helpers work on complex arrays: real parts carry the fields */
    double *f = (double*)xcalloc(2*(size_t)m->n, 8);
    double *ca = (double*)xcalloc(2*(size_t)glen*nm, 8), *ga = (double*)xcalloc(2*(size_t)glen*dim*nm, 8);
    for (int j = 0; j < nm; j++) {
        for (int i = 0; i < m->n; i++) { f[2*i] = aphi[(size_t)j*m->n + i]; f[2*i+1] = 0.0; }
        map_mesh_to_grid(glen, nn, cols, w, f, ca + 2*(size_t)j*glen);
        om_image_gradient(dim, gdim, gsize, ca + 2*(size_t)j*glen, nomask, ga + 2*(size_t)j*glen*dim);
    }
    double *cd = (double*)xcalloc(2*(size_t)glen, 8), *gd = (double*)xcalloc(2*(size_t)glen*dim, 8);
    double *pmg = (double*)xcalloc(2*(size_t)glen, 8), *pkg = (double*)xcalloc(2*(size_t)glen, 8);
    for (int q = 0; q < nq; q++) {
        for (int i = 0; i < m->n; i++) { f[2*i] = dphi[(size_t)q*m->n + i]; f[2*i+1] = 0.0; }
        map_mesh_to_grid(glen, nn, cols, w, f, cd);
        om_image_gradient(dim, gdim, gsize, cd, elref, gd);
        for (int k = lptr[q]; k < lptr[q+1]; k++) {
            const int j = lidx[k];
            const double *a = ca + 2*(size_t)j*glen, *gaj = ga + 2*(size_t)j*glen*dim;
            const double ip = logscale ? 1.0 / proj[k] : 1.0;
            for (long i = 0; i < G.gslen; i++) {
                long g = G.g2s[i];
                double pk = 0.0;
                for (int d = 0; d < dim; d++) pk -= gd[2*((size_t)d*glen + g)] * gaj[2*((size_t)d*glen + g)];
                pmg[2*g] = -cd[2*g] * a[2*g];
                pkg[2*g] = pk;
            }
            for (long s = 0; s < slen; s++) {
                double pm, pk, im;
                gmap_sol(&G, pmg, s, &pm, &im);
                gmap_sol(&G, pkg, s, &pk, &im);
                pm *= G.sz; pk *= G.sz;
                if (logscale) { pm *= ip; pk *= ip; }
                if (Jmua) Jmua[(size_t)k*slen + s] = pm;
                if (Jkap) Jkap[(size_t)k*slen + s] = pk;
            }
        }
    }
    free(cols); free(w); free(nomask); free(f); free(ca); free(ga); free(cd); free(gd); free(pmg); free(pkg);
    gmap_free(&G);
}

void om_jacobian_grid_cw(const om_mesh *m, const int *gdim, const double *bbmin, const double *bbmax,
                         const int *elref, int nq, int nm, const double *dphi, const double *aphi, const int *lptr,
                         const int *lidx, const double *proj, int logscale, double *Jmua, double *Jkap)
{
    om_jacobian_grid_cw_b(m, gdim, gdim, bbmin, bbmax, elref, nq, nm, dphi, aphi, lptr, lidx, proj, logscale,
                          Jmua, Jkap);
}

/* AddDataGradient's raster part (toastmodule.cc:1767-1843): for each
 * source q in order, the direct field dphi_q and its adjoint field aphi_q
 * (nq x n complex each) mapped to the grid, gradients with the element
 * reference as mask; grad (2 slen, cmua then ckappa) -= Re(Map_GridToSol(d a))
 * and -= Re(Map_GridToSol(sum_j grad_j d grad_j a)). */
void om_data_gradient_b(const om_mesh *m, const int *gdim, const int *bdim, const double *bbmin,
                        const double *bbmax, const int *elref, int nq, const double *dphi, const double *aphi,
                        double *grad)
{
    int dim = m->dim, nn = m->nn;
    long glen = lp_len(dim, gdim);
    double gsize[3];
    for (int d = 0; d < dim; d++) gsize[d] = bbmax[d]-bbmin[d];
    om_gmap G;
    gmap_init(&G, dim, gdim, bdim, bbmin, bbmax, elref);
    int *cols = (int*)xcalloc((size_t)glen*nn, sizeof(int));
    double *w = (double*)xcalloc((size_t)glen*nn, sizeof(double));
    om_raster_bmap(m, gdim, bbmin, bbmax, elref, cols, w);
    long slen = G.slen;
    double *cd = (double*)xcalloc(2*(size_t)glen, 8), *ca = (double*)xcalloc(2*(size_t)glen, 8);
    double *gd = (double*)xcalloc(2*(size_t)glen*dim, 8), *ga = (double*)xcalloc(2*(size_t)glen*dim, 8);
    double *pm = (double*)xcalloc(2*(size_t)glen, 8), *pk = (double*)xcalloc(2*(size_t)glen, 8);
    for (int q = 0; q < nq; q++) {
        map_mesh_to_grid(glen, nn, cols, w, dphi + 2*(size_t)q*m->n, cd);
        om_image_gradient(dim, gdim, gsize, cd, elref, gd);
        map_mesh_to_grid(glen, nn, cols, w, aphi + 2*(size_t)q*m->n, ca);
        om_image_gradient(dim, gdim, gsize, ca, elref, ga);
        for (long i = 0; i < G.gslen; i++) {
            long g = G.g2s[i];
            double t[2], gk[2] = {0.0, 0.0};
            cmul(cd + 2*g, ca + 2*g, pm + 2*g);
            for (int j = 0; j < dim; j++) {
                cmul(gd + 2*((size_t)j*glen + g), ga + 2*((size_t)j*glen + g), t);
                gk[0] += t[0]; gk[1] += t[1];
            }
            pk[2*g] = gk[0]; pk[2*g+1] = gk[1];
        }
        for (long s = 0; s < slen; s++) {
            double re, im;
            gmap_sol(&G, pm, s, &re, &im);
            grad[s] -= re;
            gmap_sol(&G, pk, s, &re, &im);
            grad[slen + s] -= re;
        }
    }
    free(cols); free(w); free(cd); free(ca); free(gd); free(ga); free(pm); free(pk);
    gmap_free(&G);
}

void om_data_gradient(const om_mesh *m, const int *gdim, const double *bbmin, const double *bbmax, const int *elref,
                      int nq, const double *dphi, const double *aphi, double *grad)
{
    om_data_gradient_b(m, gdim, gdim, bbmin, bbmax, elref, nq, dphi, aphi, grad);
}

/* QVec_Point (source.cc:157-196): the element holding p (first in
 * ascending order that contains it, GContains with its padding), else the
 * element of the nearest node pair; shape functions at p, boundary nodes
 * only for Neumann sources.  out: n values. */
void om_qvec_point(const om_mesh *m, const double *p, int neumann, double *out)
{
    int el = -1, dim = m->dim, nn = m->nn;
    for (int e = 0; e < m->elen && el < 0; e++)
        if (el_gcontains(m, e, p)) el = e;
    if (el < 0) {
        double dmin = 1e10, dmin2 = 1e10;
        const double eps = 1e-8;
        for (int e = 0; e < m->elen; e++)
            for (int j = 0; j < nn; j++) {
                double d = pdist(m->x + (size_t)m->el[e*nn+j]*dim, p, dim);
                if (d < dmin+eps)
                    for (int k = 0; k < nn; k++) {
                        if (k == j) continue;
                        double d2 = pdist(m->x + (size_t)m->el[e*nn+k]*dim, p, dim);
                        if (d2 < dmin2) { dmin2 = d2; dmin = d; el = e; }
                    }
            }
    }
    for (int i = 0; i < m->n; i++) out[i] = 0.0;
    if (el < 0) return;
    double loc[3], fun[4];
    el_local(m, el, p, loc);
    if (dim == 3) { fun[0] = 1.0-loc[0]-loc[1]-loc[2]; fun[1] = loc[0]; fun[2] = loc[1]; fun[3] = loc[2]; }
    else { fun[0] = 1.0-loc[0]-loc[1]; fun[1] = loc[0]; fun[2] = loc[1]; }
    for (int i = 0; i < nn; i++) {
        int nd = m->el[el*nn+i];
        if (m->bnd[nd] || !neumann) out[nd] = fun[i];
    }
}
