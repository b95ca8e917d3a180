/*
 * snoo_hip.h -- C ABI of libsnoo_hip.so, the MI355X (gfx950) kernels behind the
 * nonlinear reduced-vector-potential solve.
 *
 * Conventions
 *   - Every pointer argument named *_dev is a device pointer (HBM); sizes are
 *     element counts.  Host pointers are named *_host.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).  All
 *     calls are asynchronous with respect to the host unless stated.
 *   - Every function returns 0 on success or an SN_ERR_* code; the message of
 *     the last failure on the calling thread is returned by sn_last_error().
 *   - Ownership: the caller owns every buffer it passes.  Handles
 *     (sn_pattern, sn_cg) own their internal device memory and are released
 *     with their destroy call.
 *   - Indices are int32, 0-based (the reference C kernels take 0-based int32
 *     ids: fem_c/interface_c.pyx:248-256 after curl_curl_assembler.py:916-917).
 */
#ifndef SNOO_HIP_H
#define SNOO_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SN_OK 0
#define SN_ERR_ARG 1            /* bad size, null pointer or unsupported option   */
#define SN_ERR_HIP 2            /* a HIP runtime call failed                      */
#define SN_ERR_NOT_CONVERGED 3  /* iteration cap reached                          */
#define SN_ERR_BREAKDOWN 4      /* CG breakdown: p.Ap <= 0 or non-finite scalar   */

/* reluctivity model codes */
#define SN_NU_CONST 0  /* ConstantReluctance (snoopy/materials.py:486-536)        */
#define SN_NU_PCHIP 1  /* Reluctance, PCHIP in |B| (snoopy/materials.py:271-437) */
#define SN_NU_BSQ 2    /* cubic B-spline in |B|^2, piecewise-polynomial form     */

const char *sn_last_error(void);
int sn_version(void);

/* ---------------------------------------------------------------- sources */

/* Biot-Savart of straight segments with core radius; B (+)= I*1e-7*sum.
 * Replaces compute_B_line_segs (fem_c/ctools.c:1761-1923) as bound by
 * compute_B_line_segs_cpp (fem_c/interface_c.pyx:928-991).
 * src_dev [num_src*6], tar_dev [num_tar*3], B_dev [num_tar*3].
 * accumulate=0 overwrites B, 1 adds (sum over coils without a host copy). */
int sn_biot_savart_f64(const double *src_dev, int num_src, const double *tar_dev,
                       int num_tar, double current, double radius, double *B_dev,
                       int accumulate, void *stream);

/* Brick-field source: piecewise-constant B in 3 bricks.
 * Replaces BrickField.compute_B (snoopy/brick_field.py:400-433).
 * faces_host [3*6*6] = (centre, unit normal) per face, fields_host [9]. */
int sn_brick_field_f64(const double *faces_host, const double *fields_host,
                       const double *tar_dev, int num_tar, double *B_dev,
                       int accumulate, void *stream);

/* Physical quadrature points sum_a phi_a(xi_q) X_a, out_dev [num_el*Q*3].
 * Replaces the element loop of get_quadrature_points
 * (snoopy/curl_curl_assembler.py:1521-1593, evaluate_finite_element at
 * fem_c/ctools.c:480-504).  phi_dev [Q*4]. */
int sn_quad_points_f64(int num_el, const double *nodes_dev, const int *cells_dev,
                       const double *phi_dev, int Q, double *out_dev, void *stream);

/* Points and curl A at the quadrature points, points_dev/B_dev [num_el*Q*3].
 * Replaces compute_B_Hcurl (fem_c/ctools.c:1500-1563) as bound by
 * fem_c/interface_c.pyx:555-696.  curls_dev [64*Q*6*3], phi_dev [Q*4].
 * The element is the affine tetrahedron: the q = 0 slice of the curl table is
 * used for every q (lowest-order curls are constant on the element). */
int sn_B_Hcurl_f64(int num_el, const double *nodes_dev, const int *cells_dev,
                   const int *glob_ids_dev, const int *orient_dev,
                   const double *curls_dev, const double *phi_dev, int Q,
                   const double *x_dev, double *points_dev, double *B_dev,
                   void *stream);

/* Element averaging of element vectors at the mesh nodes: out [n_nodes*3] =
 * mean of Be [*, 3] over the elements node_el[node_ptr[i] .. node_ptr[i+1])
 * (0 for an empty list).  Replaces the node loop of compute_nodal_fields
 * (snoopy/curl_curl_assembler.py:2089-2197) for the chosen domains. */
int sn_nodal_average_f64(int n_nodes, const int *node_ptr_dev, const int *node_el_dev,
                         const double *Be_dev, double *out_dev, void *stream);

/* P1 interpolation of nodal vector fields at points: point p lies in element
 * el_dev[p] (cells_dev [*, 4], -1: outside) and uses field field_dev[p] of
 * nodal_dev [n_fields][n_nodes][3] (-1: outside); out [n_pts*3], 0 outside.
 * Replaces the point loop of evaluate_interpolation
 * (snoopy/evaluators.py:297-395) once points are located. */
int sn_p1_interp_f64(int n_pts, const double *pts_dev, const int *el_dev, const int *field_dev,
                     int n_nodes, const double *nodes_dev, const int *cells_dev,
                     const double *nodal_dev, double *out_dev, void *stream);

/* ------------------------------------------- the reference's own tables
 * Element kernels that take the arrays the reference's FFI receives,
 * unchanged apart from living on the device: 0-based cells [num_el*4] and
 * DOF ids [num_el*6], a per-element orientation number orient [num_el] in
 * [0, n_orient), the curl table of every orientation curls
 * [n_orient*Q*6*3] (c_{j,q} of orientation o at ((o*Q + q)*6 + j)*3), the P1
 * gradients d_phi [Q*4*3] and weights w [Q].  The element map and the curls
 * are evaluated at every quadrature point as the tables give them.  These
 * calls synchronise the stream: an orientation outside [0, n_orient) makes
 * them return SN_ERR_ARG (its element is skipped, never read out of
 * bounds). */

/* Local K, dK (21 upper-triangle pairs each, loc[pr*num_el + e]) and rhs
 * (loc6[j*num_el + e]; 0 without a source field) of one domain.  Replaces
 * compute_K_dK_Hcurl_red (fem_c/ctools.c:1195-1388, Bs_dev != NULL,
 * [num_el*Q*3]) and compute_K_dK_Hcurl (fem_c/ctools.c:960-1193, Bs_dev ==
 * NULL) as called by compute_stiffness_and_jacobi_matrix_curl_curl
 * (fem_c/interface_c.pyx:162-374, from snoopy/curl_curl_assembler.py:915-927);
 * sn_pattern_create on glob_ids_dev and sn_csr_gather_f64 /
 * sn_dof_gather_f64 then give the CSR matrices and rhs that call returns. */
int sn_element_table_f64(int num_el, const double *nodes_dev, const int *cells_dev,
                         const int *glob_ids_dev, const int *orient_dev, int n_orient,
                         const double *curls_dev, const double *d_phi_dev, const double *w_dev,
                         int Q, int type_spec, int n_int, const double *coeffs_dev,
                         const double *knots_dev, const double *x_dev, const double *Bs_dev,
                         double *locK_dev, double *locdK_dev, double *loc6_dev, void *stream);

/* Quadrature points and curl A, points_dev/B_dev [num_el*Q*3]; phi_dev [Q*4].
 * Replaces compute_B_Hcurl (fem_c/ctools.c:1500-1563) as bound by
 * fem_c/interface_c.pyx:555-696 with the reference's tables. */
int sn_B_Hcurl_table_f64(int num_el, const double *nodes_dev, const int *cells_dev,
                         const int *glob_ids_dev, const int *orient_dev, int n_orient,
                         const double *curls_dev, const double *phi_dev,
                         const double *d_phi_dev, int Q, const double *x_dev,
                         double *points_dev, double *B_dev, void *stream);

/* ---------------------------------------------------- pattern and scatter */

typedef struct sn_pattern sn_pattern;

/* Build, on the device, the CSR pattern of the union of the per-element 6x6
 * blocks (rows and columns sorted, as scipy's canonical CSR) plus the ordered
 * contribution lists that make the scatter a deterministic gather:
 *   nonzero p <- sum of local pair values, in ascending element order;
 *   DOF i     <- sum of local DOF values, in ascending element order.
 * glob_ids_dev [num_el*6] covers every element of every domain. */
int sn_pattern_create(sn_pattern **out, int num_dofs, int num_el,
                      const int *glob_ids_dev, void *stream);
void sn_pattern_destroy(sn_pattern *pat);
/* nnz and device views of indptr [num_dofs+1], indices [nnz], diag [num_dofs]
 * (position of (i,i)); the views stay owned by the pattern. */
int sn_pattern_info(const sn_pattern *pat, long long *nnz, const int **indptr_dev,
                    const int **indices_dev, const int **diag_dev);
/* Copy indptr [num_dofs+1], indices [nnz] and diag [num_dofs] into caller
 * buffers (device); any output may be NULL. */
int sn_pattern_copy(const sn_pattern *pat, int *indptr_dev, int *indices_dev, int *diag_dev,
                    void *stream);

/* Local element data layout (structure of arrays, element fastest):
 *   pair values   loc[pr*num_el_total + e], pr = index of (j<=k) in row-major
 *                 upper-triangle order, 21 pairs;
 *   DOF values    loc6[j*num_el_total + e], 6 entries. */

/* Element kernel of the drop-in assembly: local K, dK (21 pairs each) and
 * rhs (6) for elements [el_begin, el_begin+el_count) of one domain.
 * Replaces compute_K_dK_Hcurl_red (fem_c/ctools.c:1195-1388, Bs_dev != NULL)
 * and compute_K_dK_Hcurl (fem_c/ctools.c:960-1193, Bs_dev == NULL) as called
 * from compute_stiffness_and_jacobi_matrix_curl_curl
 * (fem_c/interface_c.pyx:162-374).  Bs_dev is indexed from the domain's
 * first element: [el_count*Q*3].  curls_dev [64*Q*6*3], d_phi_dev [Q*4*3],
 * w_dev [Q]; coeffs_dev [4*n_int] (or [1] for SN_NU_CONST), knots_dev
 * [n_int+1].  The element is the affine tetrahedron: the q = 0 slices of the
 * curl and gradient tables are used for every q (both are constant on the
 * lowest-order element); B_s, nu and nu' vary with q.
 * Diluted steel: dilution_host = {c_num, m_num, c_den, m_den} (NULL: none)
 * replaces compute_K_dK_Hcurl_red_dil (fem_c/ctools.c:1565-1759) as bound by
 * fem_c/interface_c.pyx:698-926: with z the physical z of quadrature point q
 * (phi_dev [Q*4], the P1 basis at the points), s = (c_num + z m_num) /
 * (c_den + z m_den), nu_q = s nu(s|B_q|), nu'_q = s^2 nu'(s|B_q|).  Like the
 * reference, only source (Bs_dev != NULL) domains are diluted. */
int sn_element_local_f64(int num_el_total, int el_begin, int el_count,
                         const double *nodes_dev, const int *cells_dev,
                         const int *glob_ids_dev, const int *orient_dev,
                         const double *curls_dev, const double *d_phi_dev,
                         const double *w_dev, int Q, int type_spec, int n_int,
                         const double *coeffs_dev, const double *knots_dev,
                         const double *x_dev, const double *Bs_dev,
                         double *locK_dev, double *locdK_dev, double *loc6_dev,
                         const double *phi_dev, const double *dilution_host, void *stream);

/* CSR values of K, dK and J = dK + K from the local pair buffers (the COO->CSR
 * sums of fem_c/interface_c.pyx:367-368 and the K += ..., dK + K of
 * snoopy/curl_curl_assembler.py:928-932).  Any output may be NULL. */
int sn_csr_gather_f64(const sn_pattern *pat, const double *locK_dev,
                      const double *locdK_dev, double *K_dev, double *dK_dev,
                      double *J_dev, void *stream);

/* out[i] = sum of the local DOF values of DOF i (the rhs scatter of
 * fem_c/ctools.c:1332-1335 made deterministic). */
int sn_dof_gather_f64(const sn_pattern *pat, const double *loc6_dev,
                      double *out_dev, void *stream);

/* Fused Newton element kernel: local Jacobian J_e = K_e + dK_e (21 pairs,
 * float64 or float32 output) and one local vector in float64 (6):
 *   mode 0: b_e = dK_e x_e + rhs_e  (b = Jx - Kx + rhs, red_mvp_solver.py:186)
 *   mode 1: R_e = K_e x_e - rhs_e   (Newton residual R(x) = K(x)x - rhs(x))
 *   mode 2: R_e as mode 1; instead of the 21 pairs, locJ_dev receives the
 *           Jacobian in reference-curl form, 6 values per element at
 *           u*num_el_total + e (u over 00, 01, 02, 11, 12, 22 of the
 *           symmetric 3x3 Mr), such that J_e[j][k] = s_j s_k c_j . Mr c_k
 *           with c_j the integer reference curl of local edge j (rows
 *           (0,-1,1), (1,0,-1), (-1,1,0), (0,0,1), (0,-1,0), (1,0,0) for the
 *           edges (0,1), (0,2), (0,3), (1,2), (1,3), (2,3)) and s_j the edge's
 *           orientation sign (-1 where bit j of orient is set).  The operand
 *           of sn_cg_set_element_operator.
 * Element arithmetic is float64 in both precisions.  Bs_dev == NULL marks a
 * source-free (air) domain.  nu_el_dev (may be NULL) receives the element's
 * mean reluctivity sum_q nu_q w_q / sum_q w_q (float32, indexed like the
 * elements), the coefficient of the auxiliary-space preconditioner.
 * phi_dev / dilution_host: diluted steel, as for sn_element_local_f64. */
int sn_element_newton(int out_f32, int mode, int num_el_total, int el_begin,
                      int el_count, const double *nodes_dev, const int *cells_dev,
                      const int *glob_ids_dev, const int *orient_dev,
                      const double *curls_dev, const double *d_phi_dev,
                      const double *w_dev, int Q, int type_spec, int n_int,
                      const double *coeffs_dev, const double *knots_dev,
                      const double *x_dev, const double *Bs_dev,
                      void *locJ_dev, double *loc6_dev, float *nu_el_dev,
                      const double *phi_dev, const double *dilution_host, void *stream);

/* CSR values from one local pair buffer, float64 or float32. */
int sn_csr_gather_one(const sn_pattern *pat, int f32, const void *loc_dev,
                      void *vals_dev, void *stream);

/* ------------------------------------------------------- sparse algebra */

/* y = A x over rows with mask != 0 and columns with mask != 0 (mask may be
 * NULL); float64. */
int sn_spmv_f64(const sn_pattern *pat, const double *vals_dev,
                const unsigned char *mask_dev, const double *x_dev, double *y_dev,
                void *stream);

/* Reductions over entries with mask != 0 (mask may be NULL), float64 result
 * written to the host. */
int sn_dot_f64(int n, const double *a_dev, const double *b_dev,
               const unsigned char *mask_dev, double *out_host, void *stream);
int sn_absmax_f64(int n, const double *a_dev, const unsigned char *mask_dev,
                  double *out_host, void *stream);
/* sn_dot_f64 with the result left on the device (*out_dev, the same bits):
 * nothing waits, so a caller can queue several and read them with one copy.
 * Calls of one host thread on one stream share a partials buffer in stream
 * order. */
int sn_dot_f64_dev(int n, const double *a_dev, const double *b_dev,
                   const unsigned char *mask_dev, double *out_dev, void *stream);
/* out_host[v] = sum over rows [row0[v], row0[v+1]) with mask != 0 of a*b. */
int sn_seg_dot_f64(int nseg, const int *row0_dev, const double *a_dev, const double *b_dev,
                   const unsigned char *mask_dev, double *out_host, void *stream);

/* Allocation-free variant of sn_seg_dot_f64.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The per-segment sums go to a file-scope device array of 4096 doubles, one
 * per device, instead of a buffer allocated per call; the call allocates and
 * frees nothing and synchronises the stream to copy the sums out.  At most
 * 4096 segments (more return SN_ERR_ARG).  Calls that overlap on one device
 * share that array and race: serialise them (one stream, or events). */
int sn_seg_dot_f64_noalloc(int nseg, const int *row0_dev, const double *a_dev,
                           const double *b_dev, const unsigned char *mask_dev,
                           double *out_host, void *stream);

/* ------------------------------------------------------------------- PCG */

typedef struct sn_cg sn_cg;

/* Workspace for Jacobi-PCG on one pattern.  precision 64: float64 matrix and
 * vectors; precision 32: float32 matrix and vectors with float64 dot-product
 * accumulation; precision 3264: float32 matrix values (the bulk of the bytes)
 * with float64 vectors, behaving as precision 64 otherwise.  lanes_per_row in {0 (auto), 4, 8, 16}; chunk = iterations
 * per captured hipGraph (0 = default). */
int sn_cg_create(sn_cg **out, const sn_pattern *pat, int precision,
                 int lanes_per_row, int chunk, void *stream);
void sn_cg_destroy(sn_cg *cg);

/* Workspace for a batch of independent systems stored as one block-diagonal
 * matrix: segment v owns rows [seg_row0_host[v], seg_row0_host[v+1]) and no
 * entry couples two segments.  Each segment runs its own CG (own alpha, beta,
 * tolerance and stop flag) inside the same kernels and the same captured
 * graph; blocks are shared out over the segments in proportion to their
 * rows. */
int sn_cg_create_batched(sn_cg **out, const sn_pattern *pat, int precision, int lanes_per_row,
                         int chunk, int nseg, const int *seg_row0_host, void *stream);
/* sn_cg_solve for every segment at once: rtol_host [nseg]; per-segment
 * iteration counts, final residual norms and status (SN_OK,
 * SN_ERR_NOT_CONVERGED, SN_ERR_BREAKDOWN) to the host arrays (may be NULL).
 * Returns the worst status. */
int sn_cg_solve_batched(sn_cg *cg, const void *vals_dev, const unsigned char *mask_dev,
                        const double *b_dev, double *x_dev, const double *rtol_host, double atol,
                        int maxiter, int *iters_host, double *rnorm_host, int *status_host,
                        void *stream);

/* Jacobi-PCG on the system restricted to mask != 0, scipy 1.15.3 cg
 * semantics as driven by snoopy/red_mvp_solver.py:195-205:
 *   M = diag(A_c)^-1; stop when ||r||_2 < max(atol, rtol*||b_c||_2) is seen at
 *   the top of an iteration; at most maxiter iterations (maxiter < 0:
 *   10 * number of free rows).
 * precision 64: vals_dev is double*, x_dev holds x0 on entry and the solution
 *   on exit (entries with mask == 0 are left untouched).
 * precision 32: vals_dev is float*, the float64 b is rounded to float32, CG
 *   starts from 0 and the float32 result is ADDED to the float64 x_dev
 *   (one refinement step: x += A^-1 b).
 * iters_host, rnorm_host (final recursive-residual norm, float64) may be NULL.
 * Returns SN_OK, SN_ERR_NOT_CONVERGED or SN_ERR_BREAKDOWN. */
int sn_cg_solve(sn_cg *cg, const void *vals_dev, const unsigned char *mask_dev,
                const double *b_dev, double *x_dev, double rtol, double atol,
                int maxiter, int *iters_host, double *rnorm_host, void *stream);

/* Row schedule of the CG SpMV kernel (a tuning knob, results identical up
 * to the summation order of p.q): bit 0 = each workgroup takes one
 * contiguous span of rows (else workgroups stride over the rows), bit 1 =
 * workgroups renumbered so each of the 8 XCDs takes one contiguous eighth of
 * them (also the element operator's two kernels).  Default 0; enabling the
 * element operator sets 2. */
int sn_cg_set_schedule(sn_cg *cg, int schedule);

/* Column storage of the CG SpMV (no counterpart in the reference, whose
 * scipy CG multiplies a CSR matrix with int32 indices).  At creation the
 * workspace stores every column as a 16-bit offset j - i when all of them
 * fit (|j - i| <= 32767, true for the edge numbering of structured meshes up
 * to a few million elements), halving the SpMV's index stream; results are
 * bit-identical either way.  enable = 0 goes back to the int32 indices;
 * enable = 1 fails with SN_ERR_ARG when the offsets do not fit. */
int sn_cg_set_col_offsets(sn_cg *cg, int enable);

/* Bytes per column index the SpMV reads: 2 (offsets) or 4. */
int sn_cg_index_bytes(const sn_cg *cg);

/* Average duration (ms) of the CG SpMV kernel (kernel A: p = z + beta p,
 * q = A p, p.q) over `reps` back-to-back launches on `stream`, timed with HIP
 * events; run after a solve on the same workspace (it overwrites the CG
 * scratch vectors).  With the element operator: its two kernels, the product
 * with the CG update of x and r. */
int sn_cg_time_spmv(sn_cg *cg, const void *vals_dev, const unsigned char *mask_dev, int reps,
                    double *avg_ms_host, void *stream);

/* ------------------------------------------ auxiliary-space preconditioner
 * No counterpart in the reference, whose CG uses the Jacobi diagonal only
 * (snoopy/red_mvp_solver.py:198-199).  Hiptmair-Xu preconditioner of the
 * singular curl-curl Jacobian:  M r = D^-1 r + weight * Pi B Pi^T r, with
 * (Pi y)_e = (x_hi - x_lo)/2 . (y_lo + y_hi) the Whitney interpolation of a
 * P1 vector field and B one V(1,1)-cycle (damped Jacobi, weight omega) of
 * unsmoothed-aggregation AMG on the P1 Laplacian with the element
 * reluctivity, applied to the three components at once (float32). */

typedef struct sn_ams sn_ams;

/* Edge side of the preconditioner (host arrays, copied).
 * edge_nodes_host [2*n_edges]: level-0 ids of the (lo, hi) nodes of each
 *   edge; ids >= n_free_nodes mark Dirichlet nodes (their value is 0).
 * edge_mesh_nodes_host [2*n_edges]: mesh node ids of (lo, hi) (the edge
 *   tangents are formed from the node coordinates at every sn_ams_update).
 * inc_ptr_host [n_free_nodes+1], inc_edge_host: free edges at each free node.
 * n_el: elements of the Laplacian (the element order of sn_ams_update). */
int sn_ams_create(sn_ams **out, int n_edges, int n_free_nodes, int n_el,
                  const int *edge_nodes_host, const int *edge_mesh_nodes_host,
                  const int *inc_ptr_host, const int *inc_edge_host, float omega,
                  float weight, void *stream);
/* Append a level (host arrays, copied), finest first.  CSR pattern
 * (indptr [n+1], indices [nnz], with a diagonal in every row); gather plan:
 * vals[k] = sum of src[gather_list[gather_ptr[k] .. gather_ptr[k+1])] in that
 * order, where src is the element pair values (entry slot*n_el + e, slot =
 * index of (a <= b) in the row-major upper triangle of the 4x4 element
 * matrix, in level-0 ids) for level 0 and the previous level's values
 * otherwise.  agg_ptr [n+1] (NULL for level 0): row k of this level is the
 * aggregate of rows [agg_ptr[k], agg_ptr[k+1]) of the previous level. */
int sn_ams_add_level(sn_ams *ams, int n, int nnz, const int *indptr_host,
                     const int *indices_host, const int *gather_ptr_host,
                     const int *gather_list_host, const int *agg_ptr_host, void *stream);
/* Numeric refresh for new element reluctivities (and node coordinates):
 * edge tangents, element Laplacians, every level's values and diagonal,
 * dense inverse of the coarsest level (at most 1024 rows).  cells_dev
 * [4*n_el] and the edge mesh nodes use the ids of nodes_dev. */
int sn_ams_update(sn_ams *ams, const double *nodes_dev, const int *cells_dev,
                  const float *nu_el_dev, void *stream);
/* z = M r (float64 vectors; entries with mask == 0 set to 0). */
int sn_ams_apply(sn_ams *ams, const unsigned char *mask_dev, const double *r_dev,
                 const double *dinv_dev, double *z_dev, void *stream);
int sn_ams_num_levels(const sn_ams *ams);
/* Lanes per row (4, 8 or 16; default 8) of the restriction, the up steps and
 * the tail of the apply: a launch-shape knob for the autotuner (results
 * identical up to the summation order).  A CG workspace that holds the
 * handle recaptures its graphs on its next solve only when its own pointers
 * change, so set the lanes before the first solve, or call sn_cg_set_grids
 * (which drops the graphs) afterwards. */
int sn_ams_set_lanes(sn_ams *ams, int lanes);
/* Down steps of the single-system apply through each coarse level's
 * combined restriction P^T (I - w A D^-1), a gather per distinct fine row
 * (combined != 0), or through the fine residual of the pre-smoothed iterate
 * (0, the default: one system's step is bound by its dependent loads, not
 * their number): the same operator up to float rounding.  The batched CG has
 * its own switch (sn_vcg_set_schedule bit 2) and uses them by default.  Like
 * the lanes, set it before the solves whose graphs capture the apply, or
 * call sn_cg_set_ams again (which drops them). */
int sn_ams_set_restriction(sn_ams *ams, int combined);
/* Host set-up helper: greedy aggregation of the strength graph of a CSR
 * matrix (|a_ij| >= theta sqrt|a_ii a_jj|); agg_host [n] receives aggregate
 * ids 0..n_agg-1 (deterministic, index order). */
int sn_amg_aggregate(int n, const int *indptr_host, const int *indices_host,
                     const double *vals_host, double theta, int *agg_host, int *n_agg);
/* rows and nonzeros of one level; copy of its current values (float32 [nnz]) */
int sn_ams_level_info(const sn_ams *ams, int level, int *n, int *nnz);
int sn_ams_level_values(const sn_ams *ams, int level, float *vals_dev, void *stream);
void sn_ams_destroy(sn_ams *ams);
/* Use `ams` (NULL: the Jacobi diagonal) as the preconditioner of every later
 * sn_cg_solve / sn_cg_solve_batched on this workspace; the handle must stay
 * alive and be updated (sn_ams_update) before the solves.  For a batched
 * workspace the preconditioner is built on the joined (block-diagonal) mesh,
 * so its cycle never couples two segments. */
int sn_cg_set_ams(sn_cg *cg, sn_ams *ams);

/* Element operator for the float32 CG.  No counterpart in the reference,
 * whose CG multiplies the assembled CSR Jacobian (snoopy/red_mvp_solver.py:
 * 195-205); this is the same matrix, J = sum_e P_e^T S_e C^T Mr_e C S_e P_e,
 * applied as a sum of element products from the reference-curl form of
 * sn_element_newton mode 2.  Rounding Mr to float32 keeps J's gradient null
 * space exact, which rounding the assembled CSR values does not: the float32
 * CG on the rounded CSR matrix can stall far above the tolerance of a tight
 * Newton step.  After enable = 1 every sn_cg_solve on this workspace takes
 * vals_dev = Mr (float32 [6*num_el], layout of mode 2) and starts from 0;
 * enable = 0 returns to CSR values.  glob_ids_dev [6*num_el] and orient_dev
 * [num_el] as given to the element kernels (num_el of the pattern).  Float32
 * workspaces (precision 32) with one segment only. */
int sn_cg_set_element_operator(sn_cg *cg, const int *glob_ids_dev, const int *orient_dev,
                               int enable, void *stream);
/* 1 when the workspace multiplies with the element operator, else 0. */
int sn_cg_element_operator(const sn_cg *cg);
/* Launch shapes of the CG kernels (tuning knobs, no counterpart in the
 * reference; results identical up to the summation order of the
 * reductions).  grid_el: workgroups of the element-operator pass (0: one
 * per 256 elements; capped at the SpMV grid, rounded down to a multiple of
 * 8); vec_per_cu: workgroups of 1024 threads per CU for the DOF pass, the
 * initialisation and the preconditioner's last kernel (default 1).
 * Synchronises the stream; the next solve recaptures its graphs.
 * sn_cg_grids reports the current shapes. */
int sn_cg_set_grids(sn_cg *cg, int grid_el, int vec_per_cu, void *stream);
int sn_cg_grids(const sn_cg *cg, int *grid_el, int *grid_vec);

/* ------------------------------------------------- variant-minor batched CG
 * No counterpart in the reference, which solves one magnet per call
 * (snoopy/red_mvp_solver.py:121-224, its scipy cg at :202-205).  nvar systems
 * on one mesh topology (the pattern's DOF numbering and element lists, one
 * Dirichlet mask) with their own element operators, right-hand sides and
 * stop tests, solved by the same kernels and captured graphs: every vector
 * is stored variant-minor (entry i*nvar + v), the element matrices as
 * Mr[(k*num_el + e)*nvar + v] (k < 6, the layout of sn_element_newton mode 2
 * per variant, interleaved), so each index is read once for all systems.
 * Per system the arithmetic is the float32 element-operator CG of
 * sn_cg_set_element_operator (scipy's cg rule, float64 inner products), with
 * per-system reductions in a fixed order.
 * nvar in {1, 2, 4, 8}; glob_ids_dev [6*num_el], orient_dev [num_el] as
 * given to the element kernels. */
typedef struct sn_vcg sn_vcg;
int sn_vcg_create(sn_vcg **out, const sn_pattern *pat, const int *glob_ids_dev,
                  const int *orient_dev, int nvar, int chunk, void *stream);
/* The auxiliary-space preconditioner shared by all systems (NULL: each
 * system's Jacobi diagonal): its hierarchy, values and coordinates are those
 * the caller last set with sn_ams_update (one system's, or any symmetric
 * positive definite choice); the workspace keeps per-system vectors of every
 * level.  Call again after the hierarchy changes. */
int sn_vcg_set_ams(sn_vcg *vcg, sn_ams *ams, void *stream);
/* Schedule flags (tuning knobs): bit 0 = blocks grouped per XCD (else dealt
 * round-robin; results identical up to the summation order of the
 * reductions); bit 1 = element products stored by element slot and gathered
 * through the DOF lists (else stored in DOF-list order and read as one run per
 * DOF; bit-identical results); bit 2 = the preconditioner's down steps form
 * the fine residual from the pre-smoothed iterate (else they apply each
 * coarse level's combined restriction P^T (I - w A D^-1); the same operator,
 * different float rounding).  Default 3. */
int sn_vcg_set_schedule(sn_vcg *vcg, int flags);
/* Solve all systems: Mr_dev float32 [6*num_el*nvar], b_dev float64 [n*nvar];
 * each CG starts from 0 and its float32 result is ADDED to x_dev (float64
 * [n*nvar], free rows only).  rtol_host [nvar]; stop when ||r_v|| <
 * max(atol, rtol_v ||b_v||) at the top of an iteration; at most maxiter
 * iterations (< 0: 10 n).  A system with b_v = 0 is done at once.  Per-system
 * iteration counts, final recursive-residual norms and status (SN_OK,
 * SN_ERR_NOT_CONVERGED, SN_ERR_BREAKDOWN) to the host arrays (may be NULL);
 * returns the first non-OK status. */
int sn_vcg_solve(sn_vcg *vcg, const float *Mr_dev, const unsigned char *mask_dev,
                 const double *b_dev, double *x_dev, const double *rtol_host, double atol,
                 int maxiter, int *iters_host, double *rnorm_host, int *status_host,
                 void *stream);
/* The operands of sn_vcg_solve from per-system arrays (host arrays of nvar
 * device pointers): Mr_dev[(j*n_el + e)*nvar + v] = Mr_src[v][j*n_el + e]
 * (each [6*n_el] float32, every entry non-null) and b_dev[i*nvar + v] =
 * -R_src[v][i] (each [n] float64; a null entry gives b = 0), in one pass. */
int sn_vcg_pack(sn_vcg *vcg, const float *const *Mr_src, const double *const *R_src,
                float *Mr_dev, double *b_dev, void *stream);
/* Average duration (ms) of the CG's element pass and DOF pass (the product
 * with the element operators and the CG update of every system) over `reps`
 * back-to-back pairs on `stream`, timed with HIP events, every system
 * running; overwrites the workspace's vectors (run after a solve). */
int sn_vcg_time_pair(sn_vcg *vcg, const float *Mr_dev, const unsigned char *mask_dev, int reps,
                     double *avg_ms_host, void *stream);
/* Launch shapes (tuning knobs; results identical up to the summation order of
 * the reductions): grid_el workgroups of the element pass (0: one per 256
 * element-variant lanes, at most 8 per CU), vec_per_cu workgroups of 1024
 * threads per CU for the DOF pass, the start and the prolongation (default
 * 1).  Synchronises the stream; the next solve recaptures its graphs. */
int sn_vcg_set_grids(sn_vcg *vcg, int grid_el, int vec_per_cu, void *stream);
int sn_vcg_info(const sn_vcg *vcg, int *nvar, int *grid_el, int *grid_vec);
void sn_vcg_destroy(sn_vcg *vcg);
/* y = J x with the element operator in float64 (Mr_dev float64, layout of
 * sn_element_newton mode 2; rows and columns with mask == 0 removed, mask may
 * be NULL); work_dev [6*num_el] receives the element products.  No
 * counterpart in the reference (a check of the element operator). */
int sn_element_apply_f64(const sn_pattern *pat, const int *glob_ids_dev, const int *orient_dev,
                         const double *Mr_dev, const unsigned char *mask_dev,
                         const double *x_dev, double *work_dev, double *y_dev, void *stream);

/* ------------------------------------------------- learned warm start
 * No counterpart in the reference (it starts Newton from A = 0 or a given
 * x_0, snoopy/red_mvp_solver.py:146-147,174); these kernels train and run the
 * network that predicts x_0. */

/* C[M,N] = op(A)[M,K] op(B)[K,N] on the BF16 matrix cores with FP32
 * accumulation.  ta/tb = 1 read A/B transposed (A stored [K][lda], B stored
 * [N][ldb]).  Epilogue: + bias[N] (may be NULL); epi 1: Zout = pre-activation
 * (may be NULL), C = GELU; epi 2: C *= GELU'(Zin).  Outputs Cf (float32)
 * and/or Cb (bfloat16), leading dimension ldc. */
int sn_gemm_bf16(int ta, int tb, int epi, int M, int N, int K, const void *A_dev, int lda,
                 const void *B_dev, int ldb, const float *bias_dev, const float *Zin_dev,
                 float *Zout_dev, float *Cf_dev, void *Cb_dev, int ldc, void *stream);
/* dY = 2 (Y - T) scale (float32 and bfloat16 copies); mean squared error to
 * loss_host (may be NULL, then no host synchronisation). */
int sn_mse_grad(int n, const float *Y_dev, const float *T_dev, float scale, float *dYf_dev,
                void *dYb_dev, double *loss_host, void *stream);
/* db[n] = sum_m dZ[m][n] */
int sn_colsum_f32(int M, int N, const float *dZ_dev, float *db_dev, void *stream);
/* one Adam step (bias-corrected, step >= 1, decoupled weight decay wd) on
 * float32 parameters; pb_dev (bfloat16 copy, may be NULL) is refreshed. */
int sn_adam_f32(int n, float *p_dev, const float *g_dev, float *m1_dev, float *m2_dev,
                void *pb_dev, float lr, float b1, float b2, float eps, int step, float wd,
                void *stream);
int sn_f32_to_bf16(int n, const float *x_dev, void *y_dev, void *stream);
/* Reduced (POD-Galerkin) Newton step on the first r POD modes U [r][n]
 * (mode k at U[k*n .. k*n+n)), with the Jacobian in the reference-curl form
 * of sn_element_newton mode 2 (Mr float32 [6*num_el], element order of
 * glob_ids/orient): H = U^T J U = sum_e W_e^T Mr_e W_e with W_e the 3 x r
 * face circulations of the modes on element e.
 * sn_pod_element_modes: W [num_el][3][r] (float32), once per model and mesh.
 * sn_pod_reduced_jacobian: H (float64, r*r, row-major) to the host; r in
 *   {32, 64, 128}; synchronises the stream.
 * sn_pod_reduced_jacobian_range: the same sum over the elements [e0, e1) of
 *   the num_el elements only (H is additive over element ranges: a domain
 *   whose Mr does not depend on the state, constant reluctivity, can be
 *   summed once per geometry and added to the other domains' sums).
 * sn_pod_project: g[k] = sum_i U[k*n + i] v[i] (float64) to the host;
 *   synchronises the stream. */
int sn_pod_element_modes(int num_el, int n, int r, const float *U_dev, const int *glob_ids_dev,
                         const int *orient_dev, float *W_dev, void *stream);
int sn_pod_reduced_jacobian(int num_el, int r, const float *W_dev, const float *Mr_dev,
                            double *H_host, void *stream);
int sn_pod_reduced_jacobian_range(int num_el, int e0, int e1, int r, const float *W_dev,
                                  const float *Mr_dev, double *H_host, void *stream);
int sn_pod_project(int n, int r, const float *U_dev, const double *v_dev, double *g_host,
                   void *stream);
/* out[i] = mean[i] + sum_j U[j*n + i] c[j]  (mean may be NULL) */
int sn_pod_expand(int n, int r, const float *U_dev, const float *c_dev, const double *mean_dev,
                  double *out_dev, void *stream);
/* The same for nvec (1, 2, 4 or 8) vectors in one pass over the modes (the
 * starts of a batch of variants): c_dev [nvec][r] (r <= 256), out_dev
 * [nvec][n]; every output bit-identical to sn_pod_expand of its vector. */
int sn_pod_expand_many(int n, int r, int nvec, const float *U_dev, const float *c_dev,
                       const double *mean_dev, double *out_dev, void *stream);
/* sn_pod_project of nvec (1, 2, 4 or 8) vectors v_dev [nvec][n] in one pass
 * over the modes: g_host [nvec][r], each bit-identical to sn_pod_project of
 * its vector; synchronises the stream. */
int sn_pod_project_many(int n, int r, int nvec, const float *U_dev, const double *v_dev,
                        double *g_host, void *stream);
/* Device-output forms, queued on the stream without waiting:
 * sn_pod_reduced_jacobian_range_dev writes H (float64, r*r) to H_dev and
 * sn_pod_project_many_dev writes g [nvec][r] to g_dev, the same values as
 * the host forms. */
int sn_pod_reduced_jacobian_range_dev(int num_el, int e0, int e1, int r, const float *W_dev,
                                      const float *Mr_dev, double *H_dev, void *stream);
int sn_pod_project_many_dev(int n, int r, int nvec, const float *U_dev, const double *v_dev,
                            double *g_dev, void *stream);
/* Reduced Newton corrections of nsys systems on the device: d_dev[s] =
 * -((H_s + H_s^T)/2)^-1 g_dev[s] (float32) by Cholesky factorization, H_dev
 * [nsys][r][r] and g_dev [nsys][r] float64, r <= 128; d = 0 for a system
 * whose matrix is not numerically positive definite (a pivot not above
 * 1e-13 of the largest diagonal entry) or whose solution is not finite. */
int sn_pod_reduced_solve_many(int nsys, int r, const double *H_dev, const double *g_dev,
                              float *d_dev, void *stream);
/* x_f[i] = sum_k w[6i+k] x_c[idx[6i+k]] (idx < 0 skipped): edge DOFs of one
 * mesh interpolated from another (Whitney interpolation weights). */
int sn_edge_transfer(int nf, const int *idx_dev, const double *w_dev, const double *xc_dev,
                     double *xf_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif
