// Device-resident finite-element mesh (Tri3 / Tet4) and its QM description.
#pragma once
#include "common.hpp"

namespace tmi {

// Per-element geometry record (doubles), one record per element:
//   Tet4: a[0..3] b[4..7] c[8..11] d[12..15] size[16] side_area[17..20]
//   Tri3: a[0..2] b[3..5] c[6..8] size[9] bndff[10..15] (lower triangle of
//         the boundary mass matrix: (0,0) (1,0) (1,1) (2,0) (2,1) (2,2))
constexpr int GEO3 = 24;
constexpr int GEO2 = 16;
// Per-element flag word: bits 0..3 boundary sides, bit 4 boundary element.
constexpr int BNDEL_BIT = 16;

struct QM {
    int nq = 0, nm = 0, nqm = 0;
    std::vector<double> Q, M;          // nq x dim, nm x dim
    std::vector<int> lptr, lidx;       // link list (CSR)
};

struct Mesh {
    int dim = 0, n = 0, elen = 0, nn = 0, nbnd = 0;
    std::vector<double> h_nodes;       // host copies of the inputs
    std::vector<int> h_elist;
    DBuf<double> nodes;                // n x dim
    DBuf<int> elist;                   // elen x nn
    DBuf<double> geo;                  // elen x GEO{2,3}
    DBuf<int> eflags;                  // elen
    DBuf<int> bnd;                     // n (0/1)
    DBuf<int> n2e_ptr, n2e;            // node -> elements, ascending
    // CSR pattern of the system matrix (built lazily)
    bool have_csr = false;
    int nnz = 0;
    int dmax = 0;                      // longest CSR row (0 = not yet computed)
    DBuf<int> rowptr, colidx, rowidx, diagpos;
    // host-side caches
    std::vector<int> h_bnd;
    QM qm;
    // device copy of the QM link list for the Jacobian (row pointer, source
    // and detector per link; rebuilt when the QM changes) and the Jacobian's
    // node-major intermediate, kept between calls
    bool links_valid = false;
    DBuf<int> d_lptr, d_lq, d_lm, d_lk;   // link list; d_lk[q * nm + m] = link of (q, m) or -1
    DBuf<double> jac_ws;
    // node orders of the solvers' private copies of the system (order.cpp),
    // one per kind (ORDER_NATURAL, ORDER_MORTON, ORDER_CLUSTER): row r of the relabelled
    // system is node perm[r]; its CSR pattern with slot[k] = nonzero of the
    // reference-order pattern held by entry k
    struct NodeOrder {
        bool valid = false;
        DBuf<int> perm, rowptr, colidx, slot, diagpos;
        // row blocks of the LDS-staged SpMM (complex FP32, 64 columns)
        int nb = 0;
        DBuf<int> brow, bu, ucols;
        DBuf<unsigned short> lcol;
        // blocks of RB_ROWS consecutive rows with the union of their columns
        // (row-block SpMM): per block the ascending distinct columns
        // (rb_col[rb_ofs[b] .. rb_ofs[b+1])), per distinct column the mask
        // of block rows holding it, and the CSR slot of each coefficient in
        // the kernel's order (distinct column ascending, then row)
        int rb_nb = 0;
        DBuf<int> rb_ofs, rb_col, rb_slot;
        DBuf<rb_mask_t> rb_mask;
        // per distinct column: row mask | (offset of its first coefficient
        // in the block << 16), for the pipelined row-block SpMM (valid when
        // every block has fewer than 65536 coefficients)
        DBuf<unsigned> rb_word;
        bool rb_word_ok = false;
    } order[3];

    int geo_stride() const { return dim == 3 ? GEO3 : GEO2; }
    void setup(hipStream_t st);        // Mesh::Setup on the GPU
    void build_csr(hipStream_t st);    // Mesh::SparseRowStructure on the GPU
    const NodeOrder &build_order(int kind, hipStream_t st);
};

enum { ORDER_NATURAL = 0, ORDER_MORTON = 1, ORDER_CLUSTER = 2 };

// node orders of the solver's copy of the system (order.cpp, host code):
// perm[r] = mesh node at row r
std::vector<int> morton_order(int dim, int n, const double *x);
std::vector<int> cluster_order(int n, const std::vector<int> &rp, const std::vector<int> &ci,
                               const std::vector<int> &seed, int group);

// host parser helpers
void read_mesh_file(const char *path, int &dim, std::vector<double> &nodes, std::vector<int> &elist,
                    std::vector<int> &region);
void read_qm_file(const char *path, int dim, QM &qm, bool &fixed_m);
// Mesh::PullToBoundary restated on the host (mesh.cc:1052-1111)
bool pull_to_boundary(const Mesh &m, const std::vector<double> &geo, const double *p, double *out);
// QVec_Point (source.cc:157-196) on the host: out has n entries
void qvec_point(const Mesh &m, const double *p, bool neumann, double *out);

// GPU kernels launched from mesh_kernels.hip
void launch_exclusive_scan(const int *in, int *out, int n, hipStream_t st);  // out has n+1 entries
void launch_qvec_gaussian(const Mesh &m, int npts, const double *d_centres, double width, int neumann,
                          const double *d_scale, double *d_out, hipStream_t st);
// elbasis = 0: nodal coefficients (n each); 1: per-element coefficients (elen each)
void launch_assemble(const Mesh &m, const double *d_cmua, const double *d_ckap, const double *d_c2a, double omega,
                     int complex_valued, double *d_val64, void *d_val32, hipStream_t st, int elbasis = 0);

// projections (mesh_kernels.hip): proj[k] = sum_j phi[q][j] * mvec[m][j] for
// the link k = lk[m * nq + q] of every linked pair (lk = -1: not linked).
// The plan fixes the segment length S of the mvec rows, the sources per
// wave and the partial-sum buffer (partial_doubles) the launch needs;
// maxrow = the longest mvec row.
struct ProjectPlan {
    int qw = 1, S = 1, nseg = 1;
    size_t partial_doubles = 0;
};
ProjectPlan project_plan(int nq, int nm, int maxrow, int cplx);
void launch_project(const ProjectPlan &pl, int nq, int nm, const int *d_lk, const int *mrp, const int *mci,
                    const double *mval, const double *phi, long sq, long sj, int cplx, double *partial, double *proj,
                    hipStream_t st);

}  // namespace tmi
