/*
 * tensormesh_amd C ABI: P1 finite-element assembly, condensation and Krylov
 * solves on one MI355X (gfx950).
 *
 * Conventions
 *  - Every function returns an int status: TM_OK (0) or a negative code;
 *    tm_last_error() gives a message for the calling thread.  Nothing throws
 *    across the ABI.
 *  - "host" functions take host pointers and run on the CPU (setup work).
 *    "device" functions take device pointers, enqueue work on the given HIP
 *    stream (a hipStream_t passed as void*, NULL = default stream) and return
 *    without synchronising unless stated otherwise.
 *  - The caller owns every buffer; the library never frees caller memory.
 *    Host-side plans are opaque handles with *_build / *_destroy.
 *  - Index types: connectivity int64 (the reference keeps cells as int64,
 *    tensormesh/mesh/mesh.py:126), CSR row_ptr/col_idx int32, element->slot
 *    map int64 (the reference's ReduceProjector indices, projector.py:24-93).
 */
#ifndef TENSORMESH_AMD_H
#define TENSORMESH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TM_OK 0
#define TM_ERR_ARG (-1)
#define TM_ERR_HIP (-2)
#define TM_ERR_LIMIT (-3)
#define TM_ERR_ALLOC (-4)
#define TM_ERR_NOTCONV (-5)

/* Forms recognised by the assembly kernels. */
#define TM_FORM_LAPLACE 0       /* scale * grad u . grad v                      */
#define TM_FORM_KAPPA_LAPLACE 1 /* kappa(x) grad u . grad v, kappa nodal        */
#define TM_FORM_MASS 2          /* scale * u v                                  */
/* OR-ed into the form of tm_asm_p1_matrix / tm_asm_p1_load, or the flags of
 * tm_asm_p1_general / tm_asm_p1_moments / tm_asm_p1_load_moments / tm_facet_load:
 * deterministic assembly.  Every element-matrix entry of a row is rounded to
 * a multiple of a power of two 2^k per row (2^(k+44) above the row's largest
 * |entry|, so each entry moves by at most 2^-44 of that maximum) before it is
 * accumulated; the FP64 sums are then exact, so the values do not depend on
 * the order of the on-chip atomics and repeat bit for bit.  Costs a second
 * pass over the block's elements. */
#define TM_ASM_DETERMINISTIC 0x100
#define TM_LOAD_CONST 0         /* c * v                                        */
#define TM_LOAD_NODAL 1         /* f(x) v, f nodal (interpolated at quadrature)  */

/* dtype codes */
#define TM_F32 0
#define TM_F64 1
#define TM_BF16 2

const char* tm_last_error(void);
int tm_version(void);

/* ---------------------------------------------------------------- host --- */

/* Canonical CSR pattern of all (conn[e,i], conn[e,j]) pairs: rows ascending,
 * columns ascending and unique.  Replaces the SciPy dedup of
 * tensormesh/assemble/element_assembler.py:838-857.  Two calls: the first
 * fills row_ptr[N+1] (nnz = row_ptr[N]), the second col_idx[nnz]. */
int tm_pattern_rowptr(const int64_t* conn, int64_t n_elem, int nb, int64_t n_points,
                      int32_t* row_ptr);
int tm_pattern_colidx(const int64_t* conn, int64_t n_elem, int nb, int64_t n_points,
                      const int32_t* row_ptr, int32_t* col_idx);

/* The same pattern and P_E built on the device from device connectivity
 * (conn [E, nb] int64): bit-identical to tm_pattern_rowptr / tm_pattern_colidx
 * / tm_element_slots.  work holds tm_pattern_dev_work_size(E, nb, N) bytes and
 * is shared by the three calls.  tm_pattern_dev_rows writes row_ptr[N+1] and
 * returns after one host synchronisation (row_ptr[N] is nnz); the caller then
 * sizes col_idx[nnz] for tm_pattern_dev_cols.  Rows are limited to 1024
 * element-vertex incidences. */
int64_t tm_pattern_dev_work_size(int64_t n_elem, int nb, int64_t n_points);
int tm_pattern_dev_rows(const int64_t* conn, int64_t n_elem, int nb, int64_t n_points, void* work,
                        int32_t* row_ptr, void* stream);
int tm_pattern_dev_cols(const int64_t* conn, int64_t n_elem, int nb, int64_t n_points, const void* work,
                        const int32_t* row_ptr, int32_t* col_idx, void* stream);
int tm_element_slots_dev(const int64_t* conn, int64_t n_elem, int nb, const int32_t* row_ptr,
                         const int32_t* col_idx, void* work, int64_t n_points, int64_t* slots, void* stream);

/* Element->edge map P_E[e*nb*nb + i*nb + j] = slot of (conn[e,i], conn[e,j]).
 * Replaces element_assembler.py:858-866 (eids_csr[elem_u, elem_v]). */
int tm_element_slots(const int64_t* conn, int64_t n_elem, int nb, const int32_t* row_ptr,
                     const int32_t* col_idx, int64_t* slots);

/* Block plan for the owner-computes assembly kernels (no reference
 * counterpart: it is the precomputed form of the projector P_E,
 * element_assembler.py:863-884).  Nodes are split by recursive coordinate
 * bisection into blocks of at most max_rows owned rows such that a block's
 * LDS footprint fits lds_budget bytes; entry_bytes is the accumulator size
 * per CSR entry (8 scalar, 72 for 3x3 blocks), node_bytes the per-halo-node
 * cache size, own_bytes the per-owned-row scratch. */
typedef struct tm_plan tm_plan;
int tm_plan_build(const double* points, int dim, const int64_t* conn, int64_t n_elem, int nb,
                  int64_t n_points, const int32_t* row_ptr, const int32_t* col_idx,
                  int max_rows, int lds_budget, int entry_bytes, int node_bytes, int own_bytes,
                  int n_threads, tm_plan** out);
/* info[0] n_blocks, [1] n_halo_total, [2] n_records, [3] n_runs,
 * [4] max LDS bytes of one block, [5] max halo, [6] max records, [7] max acc */
int tm_plan_info(const tm_plan* plan, int64_t* info);
/* blk[n_blocks*8] = {halo_off, n_halo, n_own, rec_off, n_rec, run_off, n_run, n_acc};
 * halo[n_halo_total] global node id of each LDS slot (slots [0, n_own) hold
 * the block's owned nodes, the others its halo; slots are ordered to spread
 * LDS bank accesses);
 * meta[n_halo_total] = acc_offset | diag_rank << 16 | row_length << 24 for
 * owned slots, 0 for the others;
 * recs[n_records*4] packed element records (an element's vertices may be
 * listed in any order); rec_elem[n_records] element ids;
 * runs[n_runs*4] = {value_start, acc_start, length, first_row}. */
int tm_plan_export(const tm_plan* plan, int32_t* blk, int32_t* halo, uint32_t* meta,
                   uint32_t* recs, int32_t* rec_elem, int32_t* runs);
void tm_plan_destroy(tm_plan* plan);

/* Edge records of an exported scalar plan (no reference counterpart; a
 * compact form of the same projector P_E, element_assembler.py:863-884, for
 * symmetric element matrices).  Each record lists its owned vertices first by
 * ascending slot, as LDS node words w(s) = s + (s & ~15), and carries one
 * in-row rank per vertex pair (i, j), i < j, the rank of j in the row of i:
 * tetrahedra 8 bytes (10-bit words, 4-bit ranks) when every such rank is
 * below 16 and every halo at most 512 nodes, else 12 (11-bit words, 6-bit
 * ranks); triangles 8 bytes (11-bit words, 6-bit ranks).  An edge with both
 * ends owned is accumulated in the row of the smaller slot; mir[] = src | dst
 * << 16 (block accumulator indices) fills its transposed entry.
 * eblk[n_blocks*4] = {mirror offset, mirror count | e2 << 16, e0 | e1 << 16,
 * row offset}: e0, e1, e2 end the block's record groups with nb, nb-1 and
 * nb-2 owned vertices; rows[2 * (row offset + i)] = {first value index,
 * accumulator byte offset | row length << 16} of the block's i-th owned row
 * in accumulator order (from the plan's runs).  info[0] words per record,
 * [1] mirror entries, [2] max per block, [3] max rank, [4] LDS bytes per
 * block of the edge kernel, [5] node-word bits, [6] owned rows. */
typedef struct tm_plan_edge tm_plan_edge;
int tm_plan_edge_build(int nb, int64_t n_blocks, const int32_t* blk, const uint32_t* meta,
                       const uint32_t* recs, const int32_t* runs, int n_threads, tm_plan_edge** out);
int tm_plan_edge_info(const tm_plan_edge* plan, int64_t* info);
/* recs[n_records * info[0]], eblk[n_blocks * 4], mir[info[1]], rows[info[6] * 2] */
int tm_plan_edge_export(const tm_plan_edge* plan, uint32_t* recs, int32_t* eblk, uint32_t* mir, int32_t* rows);
void tm_plan_edge_destroy(tm_plan_edge* plan);

/* -------------------------------------------------------------- device --- */

/* Device view of an exported plan. */
typedef struct {
  const int32_t* blk;
  const int32_t* halo;
  const uint32_t* meta;
  const uint32_t* recs;
  const int32_t* rec_elem;
  const int32_t* runs;
  int32_t n_blocks;
  int32_t lds_bytes;
  int32_t threads; /* workgroup size of the matrix kernel: 256, 512 or 1024 (0 = 256) */
  int32_t entry_bytes; /* tm_plan_build entry_bytes the plan was built with, -1 for a plan
                        * built with 0 (no CSR accumulators), 0 = not stated: the 3x3-block
                        * tangent and stiffness calls refuse a stated value below 72, whose
                        * blocks have no room for their accumulators */
  /* optional edge records of the same plan (tm_plan_edge_export; NULL: none):
   * tm_asm_p1_matrix uses them for its symmetric forms unless deterministic */
  const uint32_t* erecs;
  const int32_t* eblk;
  const uint32_t* emir;
  const int32_t* erows;
  int32_t erec_words; /* 2 or 3 */
  int32_t elds_bytes; /* tm_plan_edge_info info[4] */
} tm_plan_dev;

/* Reference-element moments of the quadrature rule: wsum = sum_q w_q,
 * W[c] = sum_q w_q phi_c(q), M[a][b] = sum_q w_q phi_a(q) phi_b(q).
 * Plan records may list an element's vertices in any order, so the moments
 * must be invariant under vertex relabelling (all W[c] equal, all M[a][a]
 * equal, all M[a][b], a != b, equal - true of every symmetric rule); the
 * assembly calls return TM_ERR_ARG otherwise. */
typedef struct {
  double wsum;
  double W[4];
  double M[16];
} tm_quad;

/* Boundary-facet load vector (FacetAssembler.__call__,
 * assemble/facet_assembler.py:123-277): for n_facets selected facets, each an
 * element cells[f][nb] (int64 node ids) and the local index opposite[f] of the
 * cell vertex off the facet, adds
 *   b[node_a*H + h] += sum_q w_q |F| (F0[f][q][h] phi_a(q) + F1[f][q][h][:] . grad phi_a)
 * where |F| is the facet Jacobian (edge length in 2D, twice the face area in
 * 3D), phi_a(q) the cell basis at the facet point with facet barycentric
 * coordinates q_bary[q][0..dim-1] (the facet vertices in ascending local
 * order; 0 for the opposite vertex), q_w the facet rule's weights (summing to
 * 1 on the unit interval, 1/2 on the reference triangle), F0 [n_facets][n_q][H]
 * and F1 [n_facets][n_q][H][dim] f64 (F1 may be NULL), H <= 3, n_q <= 16.
 * b [N*H] f64 is accumulated into (FP64 atomics), not overwritten.  flags:
 * TM_ASM_DETERMINISTIC (each entry's contributions rounded to a power-of-two
 * grid below its largest one, so the sums do not depend on the order of the
 * atomics); that mode needs scratch, N*H doubles zeroed by the caller
 * (NULL otherwise). */
int tm_facet_load(int nb, int dim, int64_t n_facets, const double* coords, const int64_t* cells,
                  const int32_t* opposite, int n_q, const double* q_bary, const double* q_w, int H,
                  const double* F0, const double* F1, double* b, int flags, double* scratch, void* stream);

/* Stiffness / mass values in the canonical CSR order (ElementAssembler.__call__,
 * element_assembler.py:243-466, for the forms above).  coords [N, dim] f64;
 * coef [N] nodal (f32 or f64, NULL when unused); values [nnz] f32 or f64.
 * Every CSR entry is written exactly once. */
int tm_asm_p1_matrix(int form, int nb, int dim, const tm_plan_dev* plan, const double* coords,
                     const void* coef, int coef_dtype, double scale, const tm_quad* quad,
                     void* values, int value_dtype, void* stream);

/* General constant-coefficient P1 bilinear form (a user forward(u, gradu, v,
 * gradv, scalars) recognised as U^T D V, U = (u, gradu), V = (v, gradv); the
 * reference evaluates such a forward under vmap, element_assembler.py:243-466):
 *   K_ab = int (phi_a, grad phi_a)^T D (phi_b, grad phi_b),  a = row (u), b = col (v),
 * D[(dim+1)^2] f64 host array, row-major, index 0 = value, 1.. = gradient.
 * No symmetry is assumed.  values [nnz] f32 or f64, every entry written once. */
int tm_asm_p1_general(int nb, int dim, const tm_plan_dev* plan, const double* coords, const double* D,
                      const tm_quad* quad, void* values, int value_dtype, int flags, void* stream);

/* Spatially varying bilinear forms (ElementAssembler.__call__ with a forward
 * reading x, point data, their gradients or element data,
 * element_assembler.py:332-346, 402-466): the integrand U_a^T D(q) V_b,
 * U = (phi, grad phi) of the row basis function, is given by per-element
 * moments in the element's vertex order, moments[E][(nb+dim)^2] f64 =
 * {MM[nb][nb] = sum_q w_q phi_a phi_b D00, P[nb][dim] = sum_q w_q phi_a D_0g,
 *  Qm[nb][dim] = sum_q w_q D_g0 phi_b, S[dim][dim] = sum_q w_q D_gg}
 * (w_q reference weights; the kernel supplies |det J| and grad phi).  values
 * [nnz] f32 or f64 in canonical CSR order, each entry written once. */
int tm_asm_p1_moments(int nb, int dim, const tm_plan_dev* plan, const double* coords,
                      const double* moments, void* values, int value_dtype, int flags, void* stream);
/* Spatially varying linear forms F0 v + F1 . grad v (NodeAssembler.__call__,
 * node_assembler.py:242-409, e.g. func_node_assembler, builtin.py:929-969):
 * moments[E][nb+dim] f64 = {L0[nb] = sum_q w_q phi_a F0, L1[dim] = sum_q w_q F1}
 * in the element's vertex order; b[N] f64 is overwritten. */
int tm_asm_p1_load_moments(int nb, int dim, const tm_plan_dev* plan, const double* coords,
                           const double* moments, double* b, int flags, void* stream);

/* Load vector b[N] (f64) (NodeAssembler.__call__, node_assembler.py:242-409):
 * TM_LOAD_CONST: c * int phi_a;  TM_LOAD_NODAL: int (sum_c f_c phi_c) phi_a. */
int tm_asm_p1_load(int form, int nb, int dim, const tm_plan_dev* plan, const double* coords,
                   const void* f, int f_dtype, double c, const tm_quad* quad, double* b,
                   void* stream);

/* Adjoint of the kappa-Laplace assembly: dkappa[N] += d(sum_s g_s K_s)/dkappa
 * for g[nnz] (f32 or f64).  dkappa must be zeroed by the caller.  Reference
 * gradient: autograd through element_assembler.py:402-466. */
int tm_asm_p1_kappa_adjoint(int nb, int dim, const tm_plan_dev* plan, const double* coords,
                            const void* g, int g_dtype, const tm_quad* quad, double* dkappa,
                            void* stream);

/* y = A x for CSR A (values f32 or f64), x and y f64.  nnz = row_ptr[n], given
 * by the caller (it picks the lanes per row; the call never synchronises the
 * stream).  Reference: SparseMatrix @ (matrix.py:172-176), SparseMVTorch
 * (mm/torch_mm.py:32-48). */
int tm_spmv(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx, const void* values,
            int value_dtype, const double* x, double* y, void* stream);

/* dst[i] = src[idx[i]] (4- or 8-byte elements): Condenser value gathers
 * (condense.py:294-300). */
int tm_gather(int64_t n, const int32_t* idx, const void* src, void* dst, int elem_bytes,
              void* stream);

/* Device condensation layout (Condenser._compute_layout, condense.py:182-220)
 * of a canonical CSR matrix [n x n] and a Dirichlet mask (uint8, nonzero =
 * Dirichlet).  tm_condense_count fills local[n] (an inner DOF's rank among
 * the inner DOFs, an outer DOF's rank among the outer DOFs, both in increasing
 * global order) and, when sizes is given, the device int64 sizes[4] = {n_inner,
 * n_outer, nnz_ii, nnz_io}; work holds tm_condense_work_size(n) bytes and must
 * be kept for tm_condense_fill, which writes, in the original edge order:
 * rp_ii[n_inner+1], ci_ii[nnz_ii] (local columns) and e_ii[nnz_ii] (edge ids)
 * of the inner-inner block, the same for the inner-outer block, inner_dof /
 * outer_dof (global id of each local DOF), optionally coo_ii[2][nnz_ii] and
 * coo_io[2][nnz_io] (int64 local row and column) and the uint8 edge masks
 * is_inner_edge[nnz] / is_ou2in_edge[nnz]. */
int64_t tm_condense_work_size(int64_t n);
int tm_condense_count(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const uint8_t* dirichlet,
                      void* work, int32_t* local, int64_t* sizes, void* stream);
int tm_condense_fill(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const uint8_t* dirichlet,
                     const void* work, const int32_t* local, int32_t* rp_ii, int32_t* ci_ii, int32_t* e_ii,
                     int32_t* rp_io, int32_t* ci_io, int32_t* e_io, int32_t* inner_dof, int32_t* outer_dof,
                     int64_t* coo_ii, int64_t* coo_io, uint8_t* is_inner_edge, uint8_t* is_ou2in_edge,
                     void* stream);
/* Condensed right-hand side f_i - K_io u_o (Condenser.__call__ / condense_rhs,
 * condense.py:294-309, 388-427): out[l] = rhs[inner_dof[l]] - (K_io u_outer)[l]
 * with K_io given by rp_io/ci_io and values v_io (f32 or f64); u_outer may be
 * NULL (zero Dirichlet values).  rhs, u_outer and out are f64. */
int tm_condense_rhs(int64_t n_inner, const int32_t* inner_dof, const double* rhs, const int32_t* rp_io,
                    const int32_t* ci_io, const void* v_io, int v_dtype, const double* u_outer, double* out,
                    void* stream);
/* Full solution from the inner one (Condenser.recover, condense.py:429-458):
 * out[i] = u_outer[local[i]] on Dirichlet DOFs (0 when u_outer is NULL, which
 * is Condenser.prolong), u_inner[local[i]] elsewhere; f64. */
int tm_condense_recover(int64_t n, const uint8_t* dirichlet, const int32_t* local, const double* u_inner,
                        const double* u_outer, double* out, void* stream);

/* Jacobi-PCG for SPD CSR A: stops when ||b - A x||_2 <= max(rtol ||b||_2, atol)
 * (recursive residual) or after maxiter iterations.  x holds the initial guess
 * on entry.  work must hold tm_pcg_work_size(n) bytes of device memory.
 * Synchronises the stream every check_every iterations to test convergence;
 * *iters and *relres receive the final count and ||r||/||b||.  Returns TM_OK
 * or TM_ERR_NOTCONV (x still holds the last iterate). */
int64_t tm_pcg_work_size(int64_t n);
int tm_pcg_jacobi(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const void* values,
                  int value_dtype, const double* b, double* x, void* work, double rtol,
                  double atol, int64_t maxiter, int check_every, int64_t* iters, double* relres,
                  void* stream);

/* Jacobi-PCG that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (serving loops and harnesses
 * that forbid hipMalloc).  Same arguments, kernels and results as
 * tm_pcg_jacobi without `work`: the workspace is a static device array of the
 * library, one per GPU, sized for n <= TM_PCG_NOALLOC_MAX_ROWS (2^20 rows,
 * 32 MiB; the 96^3 Poisson system has 857,375).  Larger n returns
 * TM_ERR_LIMIT before any device work.  The stream must belong to the current
 * device.  Calls on one GPU share that array: calls that overlap in time on
 * one GPU (two host threads, or two streams) race and must not be made. */
#define TM_PCG_NOALLOC_MAX_ROWS 1048576
int64_t tm_pcg_noalloc_max_rows(void);
int tm_pcg_jacobi_noalloc(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const void* values,
                          int value_dtype, const double* b, double* x, double rtol, double atol,
                          int64_t maxiter, int check_every, int64_t* iters, double* relres, void* stream);

/* Jacobi-preconditioned BiCGSTAB (right preconditioning) for square, possibly
 * non-symmetric CSR systems: the reference's lse_solver bicgstab_py
 * (sparse/solve/torch_solve.py:130-248) with device-side scalars, FP64 vectors
 * and reductions.  Stops when ||b - A x|| <= max(rtol ||b||, atol); x holds
 * the initial guess on entry.  nnz = row_ptr[n] (picks the lanes per row).
 * work: tm_bicgstab_work_size(n) bytes.  Returns TM_ERR_NOTCONV on breakdown
 * or when maxiter is reached (x is the last iterate, *relres its residual). */
int64_t tm_bicgstab_work_size(int64_t n);
int tm_bicgstab_jacobi(int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx, const void* values,
                       int value_dtype, const double* b, double* x, void* work, double rtol, double atol,
                       int64_t maxiter, int check_every, int64_t* iters, double* relres, void* stream);

/* y = A x for a block-CSR matrix with 3x3 blocks: values[nnz*9], block k at
 * values[9k..9k+8] row-major, on the scalar pattern (row_ptr, col_idx) of the
 * nodes; x and y have 3*n_block_rows entries (dof = node*3 + component).  This is
 * the storage order of SparseMatrix.from_block_coo (sparse/matrix.py:325-375).
 * nnz = row_ptr[n_block_rows] (blocks), given by the caller as for tm_spmv. */
int tm_spmv_bsr3(int64_t n_block_rows, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                 const void* values, int value_dtype, const double* x, double* y, void* stream);

/* Compact column format for the scalar Jacobi-PCG (no counterpart in the
 * reference, which streams int64 indices): col_delta[k] = col_idx[k] - row as
 * int16 for every entry of rows 0..n-1; *fits (device int) is set to 1, then
 * cleared if some entry lies more than 32767 from the diagonal (col_delta is
 * then unusable).  Stream-ordered; the caller reads *fits. */
int tm_csr_delta16(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, int16_t* col_delta, int* fits,
                   void* stream);
/* tm_pcg_jacobi with the columns given as tm_csr_delta16 offsets: the same
 * iteration (bit-identical iterates), 2 bytes less per matrix entry and SpMV. */
int tm_pcg_jacobi_d16(int64_t n, const int32_t* row_ptr, const int16_t* col_delta, const void* values,
                      int value_dtype, const double* b, double* x, void* work, double rtol, double atol,
                      int64_t maxiter, int check_every, int64_t* iters, double* relres, void* stream);
/* Jacobi-PCG on a 3x3 block-CSR matrix (point Jacobi on the block diagonals);
 * semantics as tm_pcg_jacobi with n = 3*n_block_rows unknowns; work must hold
 * tm_pcg_work_size(3*n_block_rows) bytes. */
int tm_pcg_jacobi_bsr3(int64_t n_block_rows, const int32_t* row_ptr, const int32_t* col_idx,
                       const void* values, int value_dtype, const double* b, double* x, void* work,
                       double rtol, double atol, int64_t maxiter, int check_every, int64_t* iters,
                       double* relres, void* stream);

/* Compressible neo-Hookean P1 tetrahedra (NeoHookeanModel, assemble/builtin.py:220-334):
 * for displacements u[N*3] (f64) and Young's modulus per element E_elem[E]
 * (f32/f64; NULL = the constant E0), Poisson ratio nu, writes any of
 *   kvals[nnz*9]  tangent 3x3 blocks on the scalar pattern (f32 or f64),
 *   R[N*3]        residual (gradient of the energy w.r.t. u, f64),
 *   *energy       += total strain energy (f64, caller zeroes it).
 * NULL outputs are skipped.  plan must be built with entry_bytes 72,
 * node_bytes 56, own_bytes 32; without kvals the kernel keeps no CSR
 * accumulators and a plan with entry_bytes 0, node_bytes 80, own_bytes 32
 * (that of tm_nh_apply) serves as well.  The reference computes the residual by
 * autograd of ElementAssembler.energy (element_assembler.py:468-680) and has
 * no tangent kernel. */
int tm_asm_nh(const tm_plan_dev* plan, const double* coords, const double* u, const void* E_elem, int E_dtype,
              double E0, double nu, void* kvals, int kv_dtype, double* R, double* energy, void* stream);

/* Isotropic linear-elastic stiffness of P1 tetrahedra
 * (LinearElasticityElementAssembler, assemble/builtin.py:103-179, with
 * voigt_shape_grad / voigt_stiffness, functional/elasticity.py:132-285):
 * block (a, b) = vol B_a^T C B_b, i.e.
 *   K_ab[i][k] = vol (mu (g_a . g_b) d_ik + lam g_a,i g_b,k + mu g_a,k g_b,i),
 * mu = E/(2(1+nu)), lam = E nu/((1+nu)(1-2nu)), E per element (E_elem f32/f64)
 * or the constant E0 (E_elem NULL).  Writes kvals[nnz*9] (f32 or f64) as 3x3
 * row-major blocks on the scalar pattern (from_block_coo order, row = a,
 * the u basis function; col = b).  plan: built like that of tm_asm_nh
 * (entry_bytes 72, node_bytes 56, own_bytes 32), of any block size. */
int tm_asm_linear_elastic(const tm_plan_dev* plan, const double* coords, const void* E_elem, int E_dtype, double E0,
                          double nu, void* kvals, int kv_dtype, void* stream);

/* Matrix-free tangent action Y = K(u) lam (f64 [N*3], every entry written) and,
 * when dE != NULL, dE[e] += lam . dR/dE_e (caller zeroes dE): the adjoint of
 * the residual assembly w.r.t. nodal displacements and per-element E.  The
 * kernel keeps no CSR accumulators: it stages 72 bytes per halo node and 28
 * per owned row, so plan may be built with entry_bytes 0, node_bytes 80,
 * own_bytes 32 (or be a plan of tm_asm_nh). */
int tm_nh_apply(const tm_plan_dev* plan, const double* coords, const double* u, const void* E_elem, int E_dtype,
                double E0, double nu, const double* lam, double* Y, double* dE, void* stream);

/* Batched Jacobi-PCG over independent SPD systems stacked into one
 * block-diagonal CSR (no reference counterpart: the reference solves varying
 * meshes one at a time in a Python loop, docs batched_workflows.rst:247-270).
 * Rows are cut into tiles [tile_row0[t], tile_row0[t+1]) that lie inside one
 * system tile_mesh[t]; system k owns tiles [mesh_tile0[k], mesh_tile0[k+1]).
 * Each system stops on its own relative residual rtol; x starts at zero.
 * One asynchronous persistent launch (capturable in a hipGraph); lanes = lanes
 * per row (2, 4, 8, 16), blocks = workgroups (<= 0: as many as are resident).
 * Per system: mesh_status (1 converged, 2 breakdown, 3 max iterations),
 * mesh_iters, mesh_relres. */
typedef struct {
  int64_t n_rows;
  int32_t n_tiles;
  int32_t n_mesh;
  const int32_t* row_ptr;
  const int32_t* col_idx;
  const void* values;
  int32_t value_dtype;
  int32_t reserved;
  const double* b;
  double* x;
  const int32_t* tile_row0;
  const int32_t* tile_mesh;
  const int32_t* mesh_tile0;
  double rtol;
  int64_t maxiter;
  void* work;
  int32_t* mesh_status;
  int32_t* mesh_iters;
  double* mesh_relres;
} tm_batched_args;
int64_t tm_batched_pcg_work_size(int64_t n_rows, int32_t n_tiles);
int tm_batched_pcg(const tm_batched_args* args, int lanes, int blocks, void* stream);
/* Copies the grid-barrier timeout flag (0 = every barrier completed) of the
 * last launch on this work buffer into a device int32. */
int tm_batched_timeout_flag(const tm_batched_args* args, int32_t* flag, void* stream);

/* ------------------------------------------- learned preconditioner --- */
/* No reference counterpart: the reference has no learned preconditioner and
 * no flexible CG (its Krylov methods are cg_py / bicgstab_py with Jacobi,
 * sparse/solve/torch_solve.py:70-248).
 *
 * Features are node-major [n][64] BF16 (uint16_t bit patterns).  A level is a
 * CSR graph with normalised values A^ = D^-1/2 A D^-1/2 (symmetric).  Layer
 * weights: w_t = W^T as [64][128] BF16 for the forward, w = W as [128][64]
 * BF16 for the backward, bias [64] FP32; channels < n_linear have identity
 * activation, the rest ReLU.  stop (nullable) points at a device int: kernels
 * return at once while it is non-zero (the flexible-CG status). */
int tm_gnn_mp_forward(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                      const uint16_t* h, const uint16_t* w_t, const float* bias, int n_linear, uint16_t* out,
                      uint16_t* save_msg, uint16_t* save_pre, const int* stop, void* stream);
/* FP32 floats of the weight-gradient partial buffer one backward call needs. */
int64_t tm_gnn_part_floats(void);
/* g_in = dL/dh from g_out = dL/dh'; grad_w[128*64 + 64] += (dW | db).
 * g_msg is [n][64] FP32 scratch. */
int tm_gnn_mp_backward(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                       const float* g_out, const uint16_t* h, const uint16_t* msg, const uint16_t* pre,
                       const uint16_t* w, int n_linear, float* g_in, float* g_msg, float* part, float* grad_w,
                       void* stream);
/* out[I] = sum over the children i of coarse node I of w_i h[i] (dtype TM_BF16 / TM_F32). */
int tm_gnn_pool(int64_t n_coarse, const int32_t* child_ptr, const int32_t* child, const float* w, const void* h,
                void* out, int dtype, const int* stop, void* stream);
/* out[i] = skip[i] + w_i hc[agg_i] (skip nullable; out may alias skip). */
int tm_gnn_unpool(int64_t n_fine, const int32_t* agg, const float* w, const void* hc, const void* skip, void* out,
                  int dtype, const int* stop, void* stream);
/* Per-workgroup partials of sum (d_i r_i)^2 (for the input scale), and the
 * scale s = sqrt(sum / n) from them. */
int tm_gnn_sumsq(int64_t n, const double* r, const float* d, double* ss_part, int* nparts, void* stream);
int tm_gnn_scale(int64_t n, const double* ss_part, int nparts, double* scale, void* stream);
/* Linear encoder h0 = [d r / s, lk] We + be, s = rms(d r) from the partials;
 * writes s.  We is [2][64], be [64]. */
int tm_gnn_encode(int64_t n, const double* r, const float* d, const float* lk, const double* ss_part, int nparts,
                  const float* we, const float* be, uint16_t* h0, double* scale, const int* stop, void* stream);
/* grad[2*64 + 64] += (dWe | dbe) for dL/dh0 = g_h0. */
int tm_gnn_encode_backward(int64_t n, const double* r, const float* d, const float* lk, const double* scale,
                           const float* g_h0, float* part, float* grad, void* stream);
/* y = h wd + bd (nullable output), z = s d y. */
int tm_gnn_decode(int64_t n, const uint16_t* h, const float* wd, const float* bd, const float* d,
                  const double* scale, float* y, double* z, const int* stop, void* stream);
/* Inference fusions of the finest level (the unfused kernels compute the same
 * values up to BF16 rounding of intermediate level-0 features):
 *  - tm_gnn_encode_pool: encoder + first pooling, writing the coarse
 *    features, s and xs[i] = (d_i r_i / s, lk_i) (float2 per node; lk may be
 *    NULL for a network without the log-coefficient input; with nparts = 0, s
 *    is read from *scale instead of reduced from ss_part);
 *  - tm_gnn_final: unpool (skip = encoder output) + message passing on level 0
 *    + decode, using xs, the row sums of A^ and B = A^ P^ (b_* arrays, see
 *    tm_gnn_level_prolong); writes z. */
int tm_gnn_encode_pool(int64_t n, int64_t n_coarse, const int32_t* child_ptr, const int32_t* child, const float* w,
                       const double* r, const float* d, const float* lk, const double* ss_part, int nparts,
                       const float* we, const float* be, uint16_t* out, float* xs, double* scale, const int* stop,
                       void* stream);
int tm_gnn_final(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values, const float* xs,
                 const float* row_sums, const int32_t* agg, const float* w, const uint16_t* hc,
                 const int32_t* b_row_ptr, const int32_t* b_col, const float* b_values, const float* we,
                 const float* be, const uint16_t* w_t, const float* bias, int n_linear, const float* wd,
                 const float* bd, const float* d, const double* scale, double* z, const int* stop,
                 void* stream);
/* Finest level without a layer (inference): z = s d ((e + w hc[agg]) wd + bd)
 * with e the encoder output from xs; coarse_dot [n_coarse] is scratch for
 * hc . wd. */
int tm_gnn_unpool_decode(int64_t n, int64_t n_coarse, const float* xs, const int32_t* agg, const float* w,
                         const uint16_t* hc, const float* we, const float* be, const float* wd, const float* bd,
                         const float* d, const double* scale, float* coarse_dot, double* z, const int* stop,
                         void* stream);
/* Symmetric network with a layer pair on level 0 (inference).  The level-0
 * features of the rank-one encoder output stay in the span of x = d r / s and
 * A^x, so the finest level reduces to scalar CSR products.  Both calls take
 * A's own FP32 values (A^ = D A D with D = diag d is applied as d (A (d v)),
 * so no normalised copy is streamed):
 *  - tm_gnn_sym0_down writes xs = (x, A^x) from u = d^2 r [n] (FP32; with
 *    u_ready = 0 it is formed from r first, otherwise it is the u the
 *    flexible-CG update wrote into the view's dr buffer); tm_gnn_encode_pool
 *    with r = NULL, nparts = 0, we = [a; b], be = 0 then pools
 *    x (x) a + (A^x) (x) b;
 *  - tm_gnn_sym0_up takes the level-1 up-path features hc, the vectors a, b and
 *    c3 = (a.a, a.b, b.b), and writes z; ua, ub [n_coarse] and t [n] are scratch. */
int tm_gnn_sym0_down(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                     const double* r, const float* d, float* u, int u_ready, const double* scale, float* xs,
                     const int* stop, void* stream);
int tm_gnn_sym0_up(int64_t n, int64_t n_coarse, const int32_t* row_ptr, const int32_t* col_idx,
                   const float* values, const float* xs, const int32_t* agg, const float* w,
                   const uint16_t* hc, const float* va, const float* vb, const float* c3, const float* d,
                   const double* scale, float* ua, float* ub, float* t, double* z, const int* stop, void* stream);
/* Finest-level prolongation B = A^ P^: b_values[s] = sum of A^_k w_col(k) over
 * the entries k in perm[seg[s] .. seg[s+1]).  Row sums of a CSR. */
int tm_gnn_level_prolong(int64_t nnz_b, const int32_t* seg, const int32_t* perm, const int32_t* col_idx,
                         const float* values_normalized, const float* w, float* b_values, void* stream);
int tm_gnn_row_sums(int64_t n, const int32_t* row_ptr, const float* values, float* out, void* stream);
/* g_h[i] = g_y[i] wd; grad[64 + 1] += (dwd | dbd). */
int tm_gnn_decode_backward(int64_t n, const uint16_t* h, const float* wd, const float* g_y, float* g_h,
                           float* part, float* grad, void* stream);
/* Energy loss of the normalised system: g_y = (A^ y - d r / s) / n and
 * partials [y.A^y | y.r~ | e~.r~] (e~ = e / (d s), e nullable). */
int tm_gnn_energy_loss(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                       const float* y, const double* r, const float* d, const double* scale, const double* e,
                       float* g_y, double* part, int* nparts, void* stream);
/* Scale-invariant loss gradient, in place on the g_y of tm_gnn_energy_loss:
 * L = 1 - (y.r~)^2 / ((y.A^y)(e~.r~)) (the error left after the optimal step
 * along y), from that call's partial sums. */
int tm_gnn_cosine_grad(int64_t n, const double* r, const float* d, const double* scale, const double* part,
                       int nparts, float* g_y, void* stream);
/* Level values of one matrix: d = diag^-1/2, values_normalized = D A D
 * (values_normalized may be NULL: then only d is formed). */
int tm_gnn_level_setup(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const float* values, float* d,
                       float* values_normalized, void* stream);
/* Galerkin coarse values: coarse[s] = sum of fine[perm[k]], k in [seg[s], seg[s+1]). */
int tm_gnn_level_coarsen(int64_t nnz_coarse, const int32_t* seg, const int32_t* perm, const float* fine_values,
                         float* coarse_values, void* stream);
/* w_i = d_coarse[agg_i] / d_fine[i]. */
int tm_gnn_pool_weights(int64_t n_fine, const int32_t* agg, const float* d_fine, const float* d_coarse, float* w,
                        void* stream);
/* Adam step on the FP32 master vector; entries whose mask is 0 stay fixed
 * (mask nullable). */
int tm_gnn_adam(int64_t n, float* param, const float* grad, float* m, float* v, const float* mask, float lr,
                float beta1, float beta2, float eps, int64_t step, void* stream);

/* Flexible CG truncated to the last m directions (m = 1, 2, 4, 8) around an
 * external preconditioner:
 * tm_fcg_start, then per iteration the preconditioner writes z = M(r) into the
 * view and tm_fcg_iterate finishes the iteration (one SpMV with its dot
 * products, one update, one single-workgroup convergence check).  Everything
 * is stream-ordered with device-side scalars, so iterations can be captured
 * in a hipGraph.  ss_part holds the nparts partials of sum (d r)^2, and the
 * start / check kernels store s = rms(d r) at *scale, so an encoder may read
 * it directly (tm_gnn_encode_pool with nparts = 0). */
typedef struct {
  double* r;
  double* z;
  double* ss_part;
  double* scale;
  double* relres;
  int32_t* status; /* 0 running, 1 converged, 2 breakdown, 3 max iterations */
  int32_t* iter;
  int32_t nparts;
  int32_t reserved;
  float* dr; /* d^2 r in FP32 (the u of tm_gnn_sym0_down), refreshed with r by the start and update kernels */
} tm_fcg_view;
int64_t tm_fcg_work_size(int64_t n, int m);
int tm_fcg_layout(int64_t n, int m, void* work, tm_fcg_view* view);
int tm_fcg_start(int64_t n, int m, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                 const double* b, const double* x, const float* d, void* work, double rtol, double atol,
                 int64_t maxiter, void* stream);
int tm_fcg_iterate(int64_t n, int m, const int32_t* row_ptr, const int32_t* col_idx, const float* values,
                   double* x, const float* d, void* work, void* stream);

#ifdef __cplusplus
}
#endif
#endif
