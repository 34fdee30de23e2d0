/*
 * toastpp_amd.h -- C-ABI of the MI355X diffusion forward/adjoint solve,
 * projection and Jacobian library (libtoastpp_amd.so).
 *
 * Plain C types only: pointers, sizes, int status codes.  Complex arrays are
 * interleaved (re, im) doubles.  "Host" arrays are ordinary CPU memory; the
 * *_device entry points take pointers to GPU memory (hipMalloc or a torch
 * tensor's data_ptr) on the library's device.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * toastpp/toastpp).  Unlike the reference (xERROR -> exit(1)), no call ever
 * terminates the process: errors come back as status codes and
 * tmi_last_error() gives the message of the calling thread's last failure.
 *
 * Thread safety: distinct handles may be used from distinct threads; a
 * single handle must not be used concurrently.
 */
#ifndef TOASTPP_AMD_H
#define TOASTPP_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TMI_OK                 0
#define TMI_ERR_INVALID_ARG    1
#define TMI_ERR_NOT_CONVERGED  2   /* solve finished, some column above tol */
#define TMI_ERR_OOM            3
#define TMI_ERR_DEVICE         4
#define TMI_ERR_IO             5
#define TMI_ERR_STATE          6   /* call out of order (e.g. CalcFields before Reset) */
#define TMI_ERR_UNSUPPORTED    7
#define TMI_ERR_TIMEOUT        8   /* tmi_server_wait: request not finished yet */

/* element type ids (libfe/element.h:30-49) */
#define TMI_ELID_TET4  3
#define TMI_ELID_TRI3  15

/* linear solvers (TFwdSolver LSOLVER / ITMETHOD, fwdsolver.h, gsmatrix.h:32-38) */
#define TMI_LSOLVER_DIRECT    0   /* FP64-accurate solve: BiCGSTAB/PCG to tol 1e-14 on the GPU */
#define TMI_LSOLVER_CG        1
#define TMI_LSOLVER_BICGSTAB  2
#define TMI_LSOLVER_GMRES     3   /* GMRES(20), FP64 in every mode (gmres_imp.hpp:41-230) */
/* COCG: conjugate orthogonal CG for the complex symmetric system K = A + i w B
   (PCG with the unconjugated bilinear form; one SpMM per iteration against
   BiCGSTAB's two).  No reference counterpart: an additional iterative method
   for complex systems, selected like the reference's ITMETHOD names. */
#define TMI_LSOLVER_COCG      4

/* preconditioners (PreconType, precon.h) */
#define TMI_PRECON_NONE  0
#define TMI_PRECON_DIAG  1

/* precision / batching mode of CalcFields (no reference counterpart) */
#define TMI_MODE_FP64_SINGLE  0   /* FP64 Krylov, one right-hand side at a time */
#define TMI_MODE_MIXED_BLOCK  1   /* complex-FP32 block Krylov inside FP64 iterative refinement */
#define TMI_MODE_FP64_BLOCK   2   /* FP64 block Krylov (all columns in one block) */

/* data scaling (DataScale, fwdsolver.h) */
#define TMI_DATA_LIN  0
#define TMI_DATA_LOG  1

typedef struct tmi_mesh_s   *tmi_mesh;
typedef struct tmi_solver_s *tmi_solver;
typedef struct tmi_raster_s *tmi_raster;
typedef struct tmi_server_s *tmi_server;

/* per right-hand-side convergence record (IterativeSolverResult, matrix.h:34-37,
 * extended with refinement sweeps and a converged flag) */
typedef struct {
    int    iters;      /* Krylov iterations spent on this column (all sweeps);
                          GMRES: restart cycles (gmres() returns them) */
    int    sweeps;     /* FP64 refinement sweeps (1 for FP64 modes); GMRES: Arnoldi steps */
    double relres;     /* final FP64 ||q - K phi|| / ||q||; GMRES: of the left-
                          preconditioned system, ||M^-1 (q - K phi)|| / ||M^-1 q|| */
    int    converged;  /* relres <= tol */
} tmi_colstat;

/* ---------------------------------------------------------------- library */
const char *tmi_last_error(void);
/* Device name and HBM capacity of the library's GPU (device 0). */
int tmi_device_info(char *name, int name_len, double *hbm_gib);
/* Wait for all queued GPU work of the library. */
int tmi_synchronize(void);
/* Device buffers for the *_device entry points (no reference counterpart;
 * plumbing for callers without their own allocator). */
int tmi_device_alloc(unsigned long long bytes, void **ptr);
int tmi_device_free(void *ptr);
int tmi_memcpy_h2d(void *dst, const void *src, unsigned long long bytes);
int tmi_memcpy_d2h(void *dst, const void *src, unsigned long long bytes);
int tmi_memcpy_d2d(void *dst, const void *src, unsigned long long bytes);

/* ------------------------------------------------------------------- mesh */
/* Host-only parser of a "MeshData 5.0" file (Mesh operator>>, mesh.cc:1787-1825,
 * ndlist.cc:146-159, node.cc:130-162, ellist.cc:154-190).  Call once with
 * nodes == NULL to get the sizes, then again with buffers of those sizes.
 * Only Tri3 and Tet4 elements are accepted (TMI_ERR_UNSUPPORTED otherwise). */
int tmi_read_mesh_file(const char *path, int *dim, int *nlen, int *elen, int *nnode,
                       double *nodes, int *elist, int *node_region);

/* Create a device mesh and run Mesh::Setup on the GPU (mesh.cc:96-158):
 * element geometry, MarkBoundary (mesh.cc:1614-1703), node-to-element
 * adjacency.  nodes: nlen x dim, elist: elen x (dim+1), 0-based.
 * Replaces toast.Mesh.Make (src/python/toast/mesh.py:71-89). */
int tmi_mesh_make(int dim, int nlen, const double *nodes, int elen, const int *elist, tmi_mesh *out);
/* tmi_read_mesh_file + tmi_mesh_make (toast.Mesh.Read, mesh.py:46-59) */
int tmi_mesh_read(const char *path, tmi_mesh *out);
int tmi_mesh_free(tmi_mesh m);
int tmi_mesh_dims(tmi_mesh m, int *dim, int *nlen, int *elen, int *nbnd);
int tmi_mesh_get_nodes(tmi_mesh m, double *nodes);
int tmi_mesh_get_elements(tmi_mesh m, int *elist);
/* node boundary flags after MarkBoundary (1 = boundary) */
int tmi_mesh_get_bnd(tmi_mesh m, int *bnd);
/* signed element sizes (Element_Unstructured::size, element.cc:516-530) */
int tmi_mesh_element_size(tmi_mesh m, double *size);
/* Mesh::SparseRowStructure (mesh.cc:422-549), built on the GPU.  With
 * rowptr == NULL only *nnz is returned. */
int tmi_mesh_sparse_structure(tmi_mesh m, int *rowptr, int *colidx, int *nnz);

/* QM description: QMMesh::LoadQM (qmmesh.cc:319-466) including the pulling of
 * non-fixed measurement positions onto the boundary (qmmesh.cc:105-145,
 * mesh.cc:1052-1111). */
int tmi_mesh_read_qm(tmi_mesh m, const char *path);
/* Same from arrays: Q nq x dim, M nm x dim, link list in CSR form
 * (lptr nq+1, lidx nqm).  pull_m != 0 pulls M onto the boundary. */
int tmi_mesh_set_qm(tmi_mesh m, int nq, const double *Q, int nm, const double *M,
                    const int *lptr, const int *lidx, int pull_m);
int tmi_mesh_qm_dims(tmi_mesh m, int *nq, int *nm, int *nqm);
int tmi_mesh_get_qm(tmi_mesh m, double *Q, double *M, int *lptr, int *lidx);

/* Gaussian source/measurement profiles: QVec_Gaussian (source.cc:4-41) for
 * npts centres, evaluated on the GPU.  neumann != 0: boundary source
 * (SRCMODE_NEUMANN); else isotropic volume source.  out: npts x nlen
 * (dense, real).  scale (nlen values or NULL) multiplies each profile
 * nodally, as Mvec's c2a factor (fwdfem.cc:151-176, toastmodule.cc:1189-1220). */
int tmi_qvec_gaussian(tmi_mesh m, int npts, const double *centres, double width, int neumann,
                      const double *scale, double *out);

/* Point source profiles: QVec_Point (source.cc:157-196), evaluated on the
 * host: shape functions of the element holding each centre (nearest node
 * pair if outside the mesh), on boundary nodes only if neumann != 0.
 * out: npts x nlen (dense, real).  The reference's measurement vectors have
 * no point profile (CalcMvec, toastmodule.cc:1189-1220, leaves them zero). */
int tmi_qvec_point(tmi_mesh m, int npts, const double *centres, int neumann, double *out);

/* Optical coefficients -> FEM coefficients: cmua = c mua / n,
 * ckappa = c / (3 n (mua + mus')), c2a = c / (2 n A_Keijzer(n)), c = 0.3 mm/ps
 * (toastmodule.cc:1308-1318, param.cc:24-32).  Host arrays of length nlen. */
int tmi_optical_to_coeffs(int nlen, const double *mua, const double *mus, const double *ref,
                          double *cmua, double *ckappa, double *c2a);

/* --------------------------------------------------------- forward solver */
/* TFwdSolver<T> (fwdsolver.h:77-508, fwdsolver.cc:17-1299).
 * complex_valued = 0: TFwdSolver<double>, real CW system (omega must be 0);
 * complex_valued = 1: TFwdSolver<std::complex<double>>.
 * The mesh is borrowed and must outlive the solver. */
int tmi_solver_create(tmi_mesh m, int complex_valued, int lsolver, double tol, tmi_solver *out);
int tmi_solver_free(tmi_solver s);
/* SetLinSolver("DIRECT"|"CG"|"BICGSTAB", tol) (fwdsolver.cc:1279-1299).
 * CG on a complex solver is rejected (the reference's complex CG is a stub
 * that returns the initial guess, gsmatrix_imp.hpp:289-296).  Names: DIRECT,
 * CG, BICGSTAB, GMRES (restart 20, as IterativeSolve, gsmatrix_imp.hpp:477). */
int tmi_solver_set_linsolver(tmi_solver s, const char *name, double tol);
int tmi_solver_set_precon(tmi_solver s, int precon);            /* SetPrecon */
int tmi_solver_set_maxit(tmi_solver s, int maxit);              /* SetLinSolverMaxit; 0 = n+1 */
int tmi_solver_set_mode(tmi_solver s, int mode);                /* TMI_MODE_* */
/* Column tile of the block modes (RHS columns per Krylov block; 0 = auto) */
int tmi_solver_set_tile(tmi_solver s, int tile);
/* Node order of the solver's private copy of the system (no reference
 * counterpart): TMI_ORDER_MORTON relabels the nodes along a Z-order curve of
 * their coordinates so that the sparse kernels' neighbour gathers stay in
 * cache; TMI_ORDER_CLUSTER (default) starts from the Morton order and groups
 * the rows, 8 at a time, so that each group's rows share most of their
 * columns (heavy-edge matching on the shared-column counts), which the
 * row-block SpMM turns into fewer gathers; TMI_ORDER_NATURAL keeps the mesh
 * order.  Inputs, outputs and tmi_solver_get_sysmat always use the mesh
 * order.  Changing it requires tmi_solver_allocate again. */
#define TMI_ORDER_NATURAL 0
#define TMI_ORDER_MORTON  1
#define TMI_ORDER_CLUSTER 2
int tmi_solver_set_order(tmi_solver s, int order);
/* The node order of a kind for a mesh given by its node coordinates (n x dim,
 * row-major) and the CSR pattern of its system (rowptr n+1, colidx, mesh
 * order): perm[r] = mesh node at row r of the relabelled system.  Host code
 * only (no device work); the order the solver builds in tmi_solver_allocate. */
int tmi_node_order(int kind, int dim, int n, const double *nodes, const int *rowptr, const int *colidx, int *perm);
/* Rows per run of the 64-column block SpMM (default 4): runs of this many
 * consecutive rows are dealt round-robin to the waves of each XCD; 0 gives
 * every wave one contiguous share instead.  Tuning parameter of the row
 * schedule, no effect on results. */
int tmi_solver_set_rows_per_wave(tmi_solver s, int rows);
/* 64-column complex-FP32 SpMM variant: 1 gathers neighbour rows from
 * global memory per nonzero; 0 stages each block of rows' neighbour
 * rows in LDS first (through registers); 2 stages them with the LDS DMA
 * into two alternating buffers, the next block's copy overlapping the
 * current block's compute; 3 walks blocks of 8 consecutive rows over the
 * union of their columns and gathers each shared neighbour row once; 4 is
 * 3 with double-buffered gathers; 5 (default) computes each block of 8 rows
 * as a dense product on the matrix cores (16x16x4 FP32 MFMA) over the same
 * shared neighbour rows, and also runs the 32- and 16-column tiles.
 * Tuning parameter, no effect on results beyond rounding. */
int tmi_solver_set_spmm_variant(tmi_solver s, int variant);
/* Autotuner (no reference counterpart): times the 64-column SpMM for the
 * row-schedule / variant candidates (among the row-block variants 3 and 5,
 * whole 64-column solves decide), then whole CalcFields of the caller's nq
 * right-hand sides (device, node-major [nlen][ld]) for tile widths 8..64,
 * keeps the fastest settings on the solver and reports them (any output
 * pointer may be NULL).  d_phi is scratch and ends holding a solution. */
int tmi_solver_autotune(tmi_solver s, int nq, const void *d_rhs, void *d_phi, int ld, int *tile,
                        int *rows_per_wave, int *spmm_variant, double *solve_ms);
/* Column compaction of the block solve (no reference counterpart): when at
 * most half of a tile's columns are still iterating, their state moves to
 * a tile of the smallest power-of-two width >= max(active, min_width) and
 * the finished columns are written out.  min_width 0 turns it off; default
 * 2.  Results are unchanged up to the rounding of narrower kernels. */
int tmi_solver_set_compaction(tmi_solver s, int min_width);
/* FP64 fallback of the mixed mode (no reference counterpart): columns the
 * FP32 iteration leaves above the tolerance (stagnation, breakdown or its
 * iteration cap) are solved again by the FP64 block iteration, starting
 * from where they stopped; their records add the FP64 iterations and one
 * sweep.  enable: 1 (default) or 0 (such columns come back unconverged);
 * mixed_maxit: iteration cap of the FP32 iteration (0, the default: the
 * solver's maxit). */
int tmi_solver_set_fallback(tmi_solver s, int enable, int mixed_maxit);
/* Device bytes a CalcFields of nq columns may use for its workspace,
 * including the FP64 fallback's (no reference counterpart; the serving
 * loop's admission unit). */
int tmi_solver_work_bytes(tmi_solver s, int nq, double *bytes);
/* Allocate (fwdsolver.cc:193-261): CSR pattern of the system matrix on the GPU */
int tmi_solver_allocate(tmi_solver s);
/* AssembleSystemMatrix (fwdsolver.cc:338-357, :405-424): K = PFF(cmua) +
 * PDD(ckappa) + BNDPFF(c2a) [+ i omega FF] on the GPU.  Host arrays (nlen). */
int tmi_solver_assemble(tmi_solver s, const double *cmua, const double *ckappa, const double *c2a, double omega);
/* Reset (fwdsolver.cc:494-505): assemble + preconditioner setup */
int tmi_solver_reset(tmi_solver s, const double *cmua, const double *ckappa, const double *c2a, double omega);
/* AssembleSystemMatrix / Reset with the basis argument (fwdsolver.cc:338-357,
 * :405-424, `bool elbasis`).  elbasis = 0: nodal coefficients (nlen each),
 * as tmi_solver_assemble / tmi_solver_reset.  elbasis = 1: one coefficient
 * per element (elen each; ASSEMBLE_PFF_EL / PDD_EL / BNDPFF_EL, AddToElMatrix
 * mesh.cc:1985-1993, :2112-2120): K = sum_el cmua_el IntFF + ckappa_el IntDD
 * + c2a_el BndIntFF [+ i omega FF]. */
int tmi_solver_assemble_basis(tmi_solver s, const double *cmua, const double *ckappa, const double *c2a,
                              double omega, int elbasis);
int tmi_solver_reset_basis(tmi_solver s, const double *cmua, const double *ckappa, const double *c2a,
                           double omega, int elbasis);
int tmi_solver_nnz(tmi_solver s, int *nnz);
/* Copy the FP64 system matrix back (public member F, toastmodule.cc:1069-1077).
 * val: nnz complex (interleaved) for complex solvers, nnz real otherwise. */
int tmi_solver_get_sysmat(tmi_solver s, int *rowptr, int *colidx, double *val);

/* CalcFields (fwdsolver.cc:653-810).  qvec: nq x nlen CSR (row q = source q),
 * values complex interleaved (complex solver) or real.  phi: nq x nlen
 * (complex interleaved or real), input = initial guess (the reference's
 * callers pass zeros), output = fields.  stats: nq records (may be NULL).
 * Returns TMI_ERR_NOT_CONVERGED if a column ended above tol. */
int tmi_solver_calc_fields(tmi_solver s, int nq, const int *qrowptr, const int *qcolidx,
                           const double *qval, double *phi, tmi_colstat *stats);
/* CalcFields on GPU-resident data.  d_rhs, d_phi: nlen x nq node-major
 * (element [j, q] at j*nq + q), complex128 or float64.  d_phi is overwritten
 * (zero initial guess).  Nothing is copied to the host except the stats. */
int tmi_solver_calc_fields_device(tmi_solver s, int nq, const void *d_rhs, void *d_phi, tmi_colstat *stats);
/* Several frequencies of one mesh in one block solve (no reference
 * counterpart: the reference assembles and solves once per frequency,
 * AssembleSystemMatrix + CalcFields, fwdsolver.cc:405-424, :653-810).
 * Column q of d_rhs is solved at omega[q] (host array, nq values, the unit
 * of tmi_solver_reset's omega) with the optical coefficients of the last
 * reset / assemble: (A + i omega[q] B) phi_q = rhs_q, A and B the real and
 * mass parts of the system.  Complex solvers with COCG, block modes (mixed,
 * fp64_block); layouts as tmi_solver_calc_fields_device. */
int tmi_solver_calc_fields_freqs_device(tmi_solver s, int nq, const void *d_rhs, void *d_phi, const double *omega,
                                        tmi_colstat *stats);
/* ProjectAll (fwdsolver.cc:1433-1447, :1415-1429): proj[Qofs[q]+i] =
 * sum_j phi_q[j] mvec[QMref[q][i]][j] (unconjugated), log if
 * dscale == TMI_DATA_LOG.  Uses the mesh's QM link list.  mvec: nm x nlen CSR. */
int tmi_solver_project_all(tmi_solver s, int nm, const int *mrowptr, const int *mcolidx,
                           const double *mval, const double *phi, int dscale, double *proj);
/* Same with GPU-resident fields d_phi: nlen rows of ld entries (first nq
 * used), node-major as tmi_solver_calc_fields_device writes them. */
int tmi_solver_project_all_device(tmi_solver s, int nm, const int *mrowptr, const int *mcolidx,
                                  const double *mval, const void *d_phi, int ld, int dscale, double *proj);
/* No reference counterpart: tmi_solver_project_all_device without any device
 * allocation, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as tmi_solver_project_all_device (bit-identical
 * projections: the same kernel).  The link table, the mvec CSR arrays and the
 * projections are staged in a fixed device array of
 * TMI_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a call
 * needing more (about 4 nq nm + 4 (nm + nqm) + (4 + 8 cw) mvec_nnz + 8 cw nqm
 * bytes, cw = 2 for complex) returns TMI_ERR_INVALID_ARG before touching the
 * device.  Calls on one GPU share that array and are serialised by a lock. */
#define TMI_NOALLOC_SCRATCH_BYTES (64u << 20)
int tmi_solver_project_all_device_noalloc(tmi_solver s, int nm, const int *mrowptr, const int *mcolidx,
                                          const double *mval, const void *d_phi, int ld, int dscale, double *proj);
/* Last CalcFields: GPU time (ms, HIP events) and mean Krylov iterations per column */
int tmi_solver_last_timing(tmi_solver s, double *solve_ms, double *iters_per_col);
/* Last CalcFields' host-side work (no reference counterpart): kernel launch
 * calls, HIP graph launches and host synchronisations (any may be NULL). */
int tmi_solver_last_counts(tmi_solver s, long long *kernel_launches, long long *graph_launches,
                           long long *host_syncs);
/* HIP graphs for the block Krylov iteration (no reference counterpart): with
 * enable = 1 (default) each check period of the BiCGSTAB / PCG iteration --
 * 8 iterations and the reliable update when it falls due, 64 to 104 kernel
 * launches -- is captured once per workspace and operator and replayed as
 * one graph launch; 0 launches the kernels one by one.  Same kernels, same
 * arguments, same order: results are identical either way. */
int tmi_solver_set_graphs(tmi_solver s, int enable);
/* Measurement hook (no reference counterpart): time `reps` launches of one
 * kernel of the last CalcFields tile with HIP events on the library stream.
 * which = 0: the inner block SpMM, 1: the FP64 residual SpMM.  Returns the
 * average launch time (ms) and the algorithmic HBM bytes of one launch. */
int tmi_solver_time_kernel(tmi_solver s, int which, int reps, double *avg_ms, double *bytes);

/* --------------------------------------------------------------- Jacobian */
/* GenerateJacobian, nodal basis (jacobian.cc:109-126 with RASTER_NDBASIS ->
 * GenerateJacobian_mesh :441-497, IntFG :806-835, IntGradFGradG :839-872).
 * dphi: nlen x nq complex, aphi: nlen x nm complex, both node-major as
 * toast.Mesh.Fields returns them (toastmodule.cc:1454-1485); proj: nqm complex
 * linear projections in link-list order.  J: (2 nqm) x (2 nlen) row-major
 * doubles [dlnA/dmua | dlnA/dkappa ; dphase/dmua | dphase/dkappa] for
 * dscale == TMI_DATA_LOG, [Re ; Im] of the linear PMDFs otherwise.
 * Columns are derivatives with respect to the nodal cmua and ckappa. */
int tmi_jacobian_nodal(tmi_mesh m, int nq, int nm, const double *dphi, const double *aphi,
                       const double *proj, int dscale, double *J);
/* Same on GPU-resident arrays: d_dphi nlen rows of ldd complex (first nq
 * used), d_aphi nlen rows of lda complex (first nm used), so that direct and
 * adjoint fields of one block solve can be passed in place; d_J must hold
 * 4 nqm nlen doubles. */
int tmi_jacobian_nodal_device(tmi_mesh m, int nq, int nm, const void *d_dphi, int ldd, const void *d_aphi,
                              int lda, const void *d_proj, int dscale, void *d_J);
/* GenerateJacobian_cw, nodal basis (jacobian.cc:132-186 with raster ==
 * RASTER_NDBASIS, GenerateJacobian_cw_mesh :727-790): CW (real) fields,
 * dphi nlen x nq and aphi nlen x nm node-major; proj: nqm real linear
 * projections (may be NULL for TMI_DATA_LIN).  Jmua = -IntFG(dphi_q, aphi_m)
 * and Jkap = -IntGradFGradG(dphi_q, aphi_m), one row per linked measurement
 * (nqm x nlen row-major), each row times 1/proj for TMI_DATA_LOG; either
 * output may be NULL. */
int tmi_jacobian_cw(tmi_mesh m, int nq, int nm, const double *dphi, const double *aphi, const double *proj,
                    int dscale, double *Jmua, double *Jkap);

/* GenerateJacobian, element basis (jacobian.cc:95-96,117-119 with
 * RASTER_ELBASIS -> GenerateJacobian_mesh :441-497, IntFG_el :879-906,
 * IntGradFGradG_el :911-939): fields and proj as tmi_jacobian_nodal; J:
 * (2 nqm) x (2 elen) row-major, per element -Re/-Im of
 * sum_ij (f_i g_j) IntFF(i,j) and sum_ij (f_i g_j) IntDD(i,j), each over
 * proj for TMI_DATA_LOG.  Columns are derivatives with respect to the
 * per-element cmua and ckappa of tmi_solver_assemble_basis(elbasis = 1). */
int tmi_jacobian_element(tmi_mesh m, int nq, int nm, const double *dphi, const double *aphi,
                         const double *proj, int dscale, double *J);
/* Same on GPU-resident arrays (as tmi_jacobian_nodal_device); d_J must hold
 * 4 nqm elen doubles. */
int tmi_jacobian_element_device(tmi_mesh m, int nq, int nm, const void *d_dphi, int ldd, const void *d_aphi,
                                int lda, const void *d_proj, int dscale, void *d_J);
/* GenerateJacobian_cw, element basis (jacobian.cc:141-142 ->
 * GenerateJacobian_cw_mesh :727-790 with elbasis = true): real fields as for
 * tmi_jacobian_cw; Jmua = -IntFG_el, Jkap = -IntGradFGradG_el (nqm x elen),
 * each row times 1/proj for TMI_DATA_LOG; either output may be NULL. */
int tmi_jacobian_cw_element(tmi_mesh m, int nq, int nm, const double *dphi, const double *aphi,
                            const double *proj, int dscale, double *Jmua, double *Jkap);

/* basis argument of the mvec overloads (the raster argument of
 * GenerateJacobian: RASTER_NDBASIS / RASTER_ELBASIS, jacobian.h:33-35, or a
 * Raster) */
#define TMI_BASIS_NODAL    0
#define TMI_BASIS_ELEMENT  1
#define TMI_BASIS_RASTER   2   /* the raster handle's solution basis */
/* GenerateJacobian with the measurement vectors instead of the projections
 * (jacobian.cc:86-101 -> GenerateJacobian_mesh :380-435 /
 * GenerateJacobian_grid :202-286): the linear projections are computed
 * from mvec (nm x nlen CSR, complex interleaved values; ProjectSingle with
 * DATA_LIN) and the Jacobian follows as with given projections.  J:
 * (2 nqm) x (2 len), len = nlen, elen or the raster's slen.  raster: used
 * with TMI_BASIS_RASTER only.  For TMI_BASIS_ELEMENT the result is laid out
 * as the reference's overload lays it out: its kappa PMDF is a view at
 * offset nlen instead of elen (jacobian.cc:397), so mua columns
 * nlen .. elen-1 hold the kappa PMDF of elements 0 .. elen-nlen-1 (the
 * given-projections overload, tmi_jacobian_element, has no such overlap). */
int tmi_jacobian_mvec(tmi_mesh m, tmi_raster raster, int basis, int nq, int nm, const int *mrowptr,
                      const int *mcolidx, const double *mval, const double *dphi, const double *aphi, int dscale,
                      double *J);
/* GenerateJacobian_cw with mvec (jacobian.cc:160-177): real fields and real
 * mvec values; the rows are scaled by 1/ProjectAll(mvec, dphi) for
 * TMI_DATA_LOG.  Jmua / Jkap: nqm x len, either may be NULL. */
int tmi_jacobian_cw_mvec(tmi_mesh m, tmi_raster raster, int basis, int nq, int nm, const int *mrowptr,
                         const int *mcolidx, const double *mval, const double *dphi, const double *aphi,
                         int dscale, double *Jmua, double *Jkap);

/* ----------------------------------------------------------------- raster */
/* Raster_Pixel with basis grid == intermediate grid (raster.cc:18-100,
 * raster_px.cc:10-54; toast.Raster(mesh, grd), raster.py): the element of
 * every grid point (GenerateElementPixelRef, mesh.cc:2944-3015), the
 * mesh -> grid map B (GridMapMatrix, mesh.cc:2486-2554) and the solution
 * mask (grid points inside the mesh).  gdim: dim grid sizes (each >= 2);
 * bb: NULL for the mesh bounding box (Mesh::BoundingBox, pad 0), else
 * [bbmin(dim), bbmax(dim)].  Built on the GPU. */
int tmi_raster_create(tmi_mesh m, const int *gdim, const double *bb, tmi_raster *out);
/* Raster_Pixel(bdim, gdim) (raster_px.cc:10-54; toast.Raster(mesh, grd,
 * intgrd) with grd = bdim, intgrd = gdim): fields live on the intermediate
 * grid gdim, the solution on the basis grid bdim.  With bdim != gdim the
 * grid -> basis map G (Grid2LinPixMatrix, mesh.cc:3138-3288) is built, the
 * solution voxels are the basis voxels with support (SubsampleLinPixel of
 * the inside mask, raster.cc:54-73), and the Jacobians and the data
 * gradient map their grid PMDFs through G (Map_GridToSol).  Built on the
 * GPU. */
int tmi_raster_create_basis(tmi_mesh m, const int *bdim, const int *gdim, const double *bb, tmi_raster *out);
int tmi_raster_free(tmi_raster r);
/* glen = prod(gdim), slen = solution voxels (grid points inside the mesh
 * when bdim == gdim) */
int tmi_raster_dims(tmi_raster r, int *glen, int *slen);
/* blen = prod(bdim); has_g = 1 if bdim != gdim (either may be NULL) */
int tmi_raster_basis_len(tmi_raster r, int *blen, int *has_g);
/* G = Grid2LinPixMatrix (blen x glen CSR, columns grid indices) and
 * GI = LinPix2GridMatrix (mesh.cc:3540-3718, glen x blen CSR); rowptr
 * always, colidx / val when not NULL, nnz the entry count.
 * TMI_ERR_UNSUPPORTED when bdim == gdim (the reference has no G then). */
int tmi_raster_get_gmap(tmi_raster r, int *rowptr, int *colidx, double *val, int *nnz);
int tmi_raster_get_gimap(tmi_raster r, int *rowptr, int *colidx, double *val, int *nnz);
/* Map_GridToBasis / Map_BasisToGrid (raster_px.cc:102-130, Ax_cplx of G /
 * GI) for ncols complex columns: gvec glen x ncols, bvec blen x ncols */
int tmi_raster_map_grid_to_basis(tmi_raster r, int ncols, const double *gvec, double *bvec);
int tmi_raster_map_basis_to_grid(tmi_raster r, int ncols, const double *bvec, double *gvec);
/* bounding box: bbmin(dim), bbmax(dim) */
int tmi_raster_bbox(tmi_raster r, double *bbmin, double *bbmax);
/* element per grid point (-1 outside), glen; sol2basis (Raster::sol2basis,
 * raster.cc:64-77: basis index of each solution voxel), slen; either may
 * be NULL */
int tmi_raster_get_elref(tmi_raster r, int *elref, int *sol2basis);
/* rows of B: glen x (dim+1) node indices (-1 outside) and weights */
int tmi_raster_get_bmap(tmi_raster r, int *cols, double *w);
/* NodeMapMatrix (mesh.cc:2695-2853), the grid -> mesh map BI: per mesh
 * node 2^dim grid points (-1: unused slot) and their bi/trilinear weights,
 * renormalised over the points inside the mesh; built on first use. */
int tmi_raster_get_nodemap(tmi_raster r, int *cols, double *w);
/* Map_GridToMesh (raster.cc:212-224, Ax_cplx of BI) for ncols complex
 * columns: gvec glen x ncols, mvec nlen x ncols (node-major).  With the
 * basis grid equal to the intermediate grid this is also Map_BasisToMesh,
 * and Map_SolToMesh after scattering the solution voxels into the grid;
 * else Map_BasisToMesh is tmi_raster_map_basis_to_grid followed by this. */
int tmi_raster_map_grid_to_mesh(tmi_raster r, int ncols, const double *gvec, double *mvec);
/* Map_MeshToGrid (raster.cc:188-199, Ax_cplx of B) for ncols complex
 * columns: mvec nlen x ncols (node-major), gvec glen x ncols */
int tmi_raster_map_mesh_to_grid(tmi_raster r, int ncols, const double *mvec, double *gvec);
/* GenerateJacobian_grid (jacobian.cc:293-373): dphi nlen x nq, aphi
 * nlen x nm complex (node-major, as toast.Mesh.Fields returns them), proj
 * nqm complex linear projections; J (2 nqm) x (2 slen) row-major
 * [dlnA/dmua | dlnA/dkappa ; dphase/dmua | dphase/dkappa] on the solution
 * voxels, scaled by the voxel size prod gsize/(gdim-1). */
int tmi_jacobian_raster(tmi_raster r, int nq, int nm, const double *dphi, const double *aphi,
                        const double *proj, int dscale, double *J);
/* Same on GPU-resident arrays (field rows of ldd / lda complex entries);
 * d_J must hold 4 nqm slen doubles. */
int tmi_jacobian_raster_device(tmi_raster r, int nq, int nm, const void *d_dphi, int ldd, const void *d_aphi,
                               int lda, const void *d_proj, int dscale, void *d_J);
/* AddDataGradient's raster part (toastmodule.cc:1745-1843, used by
 * toast.Gradient :1884-1941): dphi the direct and aphi the adjoint complex
 * fields of the nq sources (nlen x nq node-major, column q of aphi the
 * adjoint field of source q); grad (2 slen: cmua then ckappa, solution
 * voxels) is decremented by Re(d_q a_q) and Re(grad d_q . grad a_q) for
 * q = 0, 1, ... in order (gradients with the inside-mesh mask, no voxel
 * scaling), as the reference accumulates it. */
int tmi_raster_data_gradient(tmi_raster r, int nq, const double *dphi, const double *aphi, double *grad);
/* Same on GPU-resident fields (rows of ldd / lda complex) and gradient. */
int tmi_raster_data_gradient_device(tmi_raster r, int nq, const void *d_dphi, int ldd, const void *d_aphi, int lda,
                                    void *d_grad);
/* GenerateJacobian_cw on the raster basis (jacobian.cc:132-186,
 * GenerateJacobian_cw_grid :610-721): real fields as for tmi_jacobian_cw;
 * Jmua / Jkap nqm x slen row-major, PMDFs -d a and -grad d . grad a on the
 * solution voxels (the adjoint gradients without the inside-mesh mask, as the
 * reference takes them), times the voxel size, each row times 1/proj for
 * TMI_DATA_LOG; either output may be NULL. */
int tmi_jacobian_raster_cw(tmi_raster r, int nq, int nm, const double *dphi, const double *aphi, const double *proj,
                           int dscale, double *Jmua, double *Jkap);

/* ---------------------------------------------------------------- serving */
/* Serving loop (no reference counterpart; the reference's CalcFields runs
 * one caller's sources to completion, fwdsolver.cc:653-810).  A server owns
 * `slots` worker threads, each with its own HIP stream and workspace.
 * Requests -- a prepared solver (Allocate + Reset done, matrix resident)
 * and nq right-hand sides in device memory, node-major [nlen][ld] --
 * are admitted in arrival order as soon as a slot is free and the
 * workspaces of the running requests fit hbm_budget_bytes; admitted
 * requests run concurrently.  One solver may serve several requests at
 * once.  Results do not depend on what else runs (fixed-order reductions).
 * The device buffers must stay valid until tmi_server_wait returns. */
int tmi_server_create(int slots, double hbm_budget_bytes, tmi_server *out);
int tmi_server_free(tmi_server s);          /* finishes queued requests first */
int tmi_server_submit(tmi_server s, tmi_solver solver, int nq, const void *d_rhs, void *d_phi, int ld,
                      long long *id);
/* Wait for a request (timeout_s < 0: no limit; TMI_ERR_TIMEOUT if not done).
 * stats: nq records; times: [queued, started, finished] seconds on the
 * server clock; status: the request's own status (TMI_OK,
 * TMI_ERR_NOT_CONVERGED or an error). */
int tmi_server_wait(tmi_server s, long long id, double timeout_s, tmi_colstat *stats, double *times, int *status);
int tmi_server_counts(tmi_server s, int *running, int *queued, double *hbm_bytes);
double tmi_server_clock(tmi_server s);      /* seconds since creation */

/* ------------------------------------------------------------- data files */
/* Nodal image (NIM) files: "NIM" header, "ImageSize = n", "EndHeader", then
 * blocks "Image k" of n values.  Host only, no device needed. */
#define TMI_NIM_FWDFEM    0   /* values "%.10g " (WriteNIM, fwdfem.cc:316-323) */
#define TMI_NIM_TOASTMOD  1   /* values "%.12e " (toast_write_nim, toastmodule.cc:942-972) */
/* toast.ReadNim / toast_read_nim (toastmodule.cc:905-940, mesh.py:206-209,
 * 338-339) and ReadNim (fwdfem.cc:327-347): image idx (0-based; idx < 0 or
 * past the end: the last image).  size / nimg (may be NULL) return the
 * image length and the number of images; img (may be NULL) receives the
 * image. */
int tmi_read_nim(const char *path, int idx, int *size, int *nimg, double *img);
/* toast.WriteNim (toastmodule.cc:942-972) for nimg == 1, OpenNIM + WriteNIM
 * (fwdfem.cc:306-323) for a series: imgs is nimg x size, image k written
 * as "Image k". */
int tmi_write_nim(const char *path, const char *meshname, int size, int nimg, const double *imgs, int style);
/* Vector text "[v0 v1 ...]" of the .fem data files: TVector operator>>
 * (vector_imp.hpp:1370-1420).  Call with v == NULL for the length; else *n
 * is the buffer length on input and the vector length on output. */
int tmi_read_vector(const char *path, int *n, double *v);
/* WriteData (fwdfem.cc:457-462): "[" values "%.10e" separated by blanks "]" */
int tmi_write_vector(const char *path, int n, const double *v);

#ifdef __cplusplus
}
#endif
#endif /* TOASTPP_AMD_H */
