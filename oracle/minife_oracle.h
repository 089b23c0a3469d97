/*
 * minife_oracle.h -- CPU restatement of miniFE's ref/ assemble-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle for the HIP library
 * (libminife_hip.so) and the "port" CPU baseline.  Only tests/, the smoke()
 * entry point and bench.py's CPU-baseline leg may load it, and only as a
 * checker.  The product never links or calls it.
 *
 * Scope: one rank (local box == global box), so the id->row map of
 * simple_mesh_description (ref/utils/box_utils.hpp:233-315) is the identity.
 * Types follow ref/src/Makefile.gnu.serial: Scalar double, ordinals int32.
 * Every function follows the reference's operation order so results are
 * bit-identical to the compiled reference built without FMA contraction;
 * each cites the file:line it restates (paths relative to the reference root).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   structure  -> ref_dump fixtures (bit-exact), closed-form nnz, utest counts
 *   element    -> ref_dump elem fixtures (bit-exact), symmetry, zero row sums,
 *                 source = h^3/8 per node
 *   assembly   -> ref_dump A_prebc/b_prebc fixtures (bit-exact)
 *   dirichlet  -> ref_dump A_bc0/A/b fixtures (bit-exact)
 *   matvec     -> utest_cases.hpp:927-1033 4x4 example, fixtures
 *   dot/waxpby -> textbook definitions checked on fixtures (CG history)
 *   cg_solve   -> ref_dump per-iteration (pAp, rtrans, alpha) log, bit-exact
 *   soln/verify-> centre value 1/6 of the Laplace problem, ref verify rc
 */
#ifndef MINIFE_ORACLE_H
#define MINIFE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Box as ref/utils/Box.hpp: {x0,x1,y0,y1,z0,z1}, element ranges [lo,hi). */

/* Number of rows and non-zeros generate_matrix_structure produces. */
int orc_structure_sizes(const int gbox[6], int64_t* nrows, int64_t* nnz);

/* ref/src/generate_matrix_structure.hpp:47-149 + MatrixInitOp.hpp:127-156.
 * rows[nrows], row_offsets[nrows+1], cols[nnz], vals[nnz] (zeroed). */
int orc_generate_matrix_structure(const int gbox[6], int* rows, int* row_offsets,
                                  int* cols, double* vals);

/* ref/fem/gauss_pts.hpp:37-62 (2-point rule). */
void orc_gauss_pts(double pts[2], double wts[2]);
/* ref/fem/Hex8.hpp:43-65 */
void orc_shape_fns(const double x[3], double values_at_nodes[8]);
/* ref/fem/Hex8.hpp:67-113 */
void orc_gradients(const double x[3], double values_per_fn[24]);
/* ref/fem/Hex8_ElemData.hpp:37-61 -> grad_vals[192] */
void orc_compute_gradient_values(double grad_vals[192]);
/* ref/fem/Hex8.hpp:259-361; k scales the diffusion coefficient (ref: k=1.0). */
void orc_diffusion_matrix_symm(const double coords[24], const double grad_vals[192],
                               double k, double elem_mat[36]);
/* ref/fem/Hex8.hpp:363-410; Q is the source (ref: Q=1.0). */
void orc_source_vector(const double coords[24], const double grad_vals[192], double Q,
                       double elem_vec[8]);
/* ref/src/Hex8_box_utils.hpp:101-159 */
void orc_get_elem_nodes_and_coords(const int gbox[6], int elem_id, int node_ids[8],
                                   double coords[24]);

/* ref/src/perform_element_loop.hpp:40-104 over elem_box, scattering with
 * sum_into_global_linear_system (SparseMatrix_functions.hpp:183-255,
 * Vector_functions.hpp:84-103).  kappa (nullable) scales element e's matrix:
 * K_e(kappa) = kappa[e] * K_e(1) computed as ref with k = kappa[e]. */
int orc_perform_element_loop(const int gbox[6], const int elem_box[6], int64_t nrows,
                             const int* row_offsets, const int* cols, double* vals,
                             double* b, const double* kappa);
/* ref/src/assemble_FE_data.hpp:38-73 (one rank: no ghost layer). */
int orc_assemble_FE_data(const int gbox[6], int64_t nrows, const int* row_offsets,
                         const int* cols, double* vals, double* b, const double* kappa);

/* simple_mesh_description.hpp:77-224: which==0 -> bc_rows_0 (faces x=0,y=0,
 * z=0,y=1,z=1), which==1 -> bc_rows_1 (face x=1).  Sorted, unique.  Returns
 * the count; out may be NULL to query. */
int64_t orc_bc_rows(const int gbox[6], int which, int* out);
/* The same with the u = 1 face selectable: hot_face 1 = x=1 (reference),
 * 0 = x=0 (the x=1 face then joins bc_rows_0).  Pinned by mirror symmetry. */
int64_t orc_bc_rows_face(const int gbox[6], int which, int hot_face, int* out);
/* assemble_FE_data with source Q (reference Q = 1; Q = 0: no source). */
int orc_assemble_FE_data_q(const int gbox[6], int64_t nrows, const int* row_offsets,
                           const int* cols, double* vals, double* b, const double* kappa,
                           double Q);

/* ref/src/SparseMatrix_functions.hpp:431-480 (+ zero_row_and_put_1 :411-428). */
void orc_impose_dirichlet(double value, int64_t nrows, const int* row_offsets,
                          const int* cols, double* vals, double* b, const int* bc_rows,
                          int64_t nbc);

/* ref/src/SparseMatrix_functions.hpp:545-578 (y initial sum = 0*y[row]). */
void orc_matvec(int64_t nrows, const int* row_offsets, const int* cols,
                const double* vals, const double* x, double* y);
/* The mixed-precision solve's inner SpMV (no reference counterpart): values and
 * x rounded to float, float products summed in stored column order from +0.0,
 * y widened to double.  Pinned by the FP64 matvec it rounds (error bound). */
void orc_matvec_f32(int64_t nrows, const int* row_offsets, const int* cols,
                    const double* vals, const double* x, double* y);
/* ref/src/Vector_functions.hpp:194-226 */
double orc_dot(int64_t n, const double* x, const double* y);
/* ref/src/Vector_functions.hpp:127-150 */
void orc_waxpby(int64_t n, double alpha, const double* x, double beta, const double* y,
                double* w);

/* ref/src/cg_solve.hpp:69-229.  history (nullable) receives 4 doubles per
 * iteration k=1..num_iters: {rtrans, p_ap_dot, alpha, beta}; entry k-1.
 * times[5] = {waxpy, dot, matvec, matvecdot, total} seconds (nullable).
 * Returns 0, or 1 on numerical breakdown (num_iters = last completed k). */
int orc_cg_solve(int64_t nrows, const int* row_offsets, const int* cols,
                 const double* vals, const double* b, double* x, int max_iter,
                 double tolerance, int* num_iters, double* normr, double* history,
                 double* times);

/* The same CG with OpenMP loops as openmp/src (matvec, waxpby, reduction
 * dot): the multi-core CPU timing baseline.  nthreads <= 0: OpenMP default.
 * Pinned by the serial solve (same iterate up to the dot's summation order). */
int orc_cg_solve_omp(int64_t nrows, const int* row_offsets, const int* cols,
                     const double* vals, const double* b, double* x, int max_iter,
                     double tolerance, int nthreads, int* num_iters, double* normr,
                     double* times);
int orc_max_threads(void);

/* Adjoint element gradient (no reference counterpart): grad[e] =
 * -lambda_e^T K_e(1) utilde_e with K_e(1) = diffusionMatrix_symm at k = 1;
 * per element: for i: t = sum_j K_ij utilde_j; s += lambda_i t; grad = -s.
 * Pinned by central finite differences of the oracle's own solve. */
void orc_element_gradient(const int gbox[6], const double* lambda, const double* utilde,
                          double* grad);
/* ref/fem/analytic_soln.hpp:80-112 */
double orc_soln(double x, double y, double z, int max_p, int max_q);
/* ref/fem/verify_solution.hpp:55-174; returns rc, *max_err the max abs error. */
int orc_verify_solution(const int gbox[6], const double* x, double tolerance,
                        int verify_whole_domain, double* max_err);
/* hot_face 0: the analytic solution mirrored, soln(1 - x, y, z). */
int orc_verify_solution_face(const int gbox[6], const double* x, double tolerance,
                             int verify_whole_domain, int hot_face, double* max_err);

#ifdef __cplusplus
}
#endif
#endif
