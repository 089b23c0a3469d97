// Test-infrastructure only: a driver of our own that calls the reference
// miniFE (ref/) operators, compiled from the reference's headers where they lie
// (see oracle/Makefile).  It is never shipped and never timed as the product.
//
// Modes
//   mesh NX NY NZ OUTDIR MAXITER   run the ref/ assemble-and-solve path
//                                  (driver.hpp:159-305 order) and dump every
//                                  intermediate array as raw little-endian
//                                  binary into OUTDIR, plus the per-iteration
//                                  CG log (cg_solve.hpp:182-204, MINIFE_DEBUG).
//   elem INFILE OUTFILE            for each record of 24 doubles (hex8 node
//                                  coordinates) write Hex8::diffusionMatrix_symm
//                                  (36) and Hex8::sourceVector (8).
//   time NX NY NZ CGITERS          time the phases of the same path (no dumps)
//                                  and print one JSON line.
#include <iostream>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <limits>
#include <string>
#include <vector>
#include <set>
#include <unistd.h>

#include <Box.hpp>
#include <box_utils.hpp>
#include <simple_mesh_description.hpp>
#include <CSRMatrix.hpp>
#include <Vector.hpp>
#include <Parameters.hpp>
#include <generate_matrix_structure.hpp>
#include <assemble_FE_data.hpp>
#include <perform_element_loop.hpp>
#include <SparseMatrix_functions.hpp>
#include <Vector_functions.hpp>
#include <make_local_matrix.hpp>
#include <cg_solve.hpp>
#include <verify_solution.hpp>
#include <Hex8_ElemData.hpp>
#include <outstream.hpp>
#include <mytimer.hpp>

typedef miniFE::CSRMatrix<double, int, int> Mat;
typedef miniFE::Vector<double, int, int> Vec;

template <typename T>
static void dump(const std::string& dir, const char* name, const T* p, size_t n) {
  std::string path = dir + "/" + name + ".bin";
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) { std::perror(path.c_str()); std::exit(2); }
  if (n && std::fwrite(p, sizeof(T), n, f) != n) { std::perror("fwrite"); std::exit(2); }
  std::fclose(f);
}

static void dump_matrix(const std::string& dir, const char* tag, const Mat& A) {
  std::string s(tag);
  dump(dir, (s + "_rows").c_str(), A.rows.data(), A.rows.size());
  dump(dir, (s + "_row_offsets").c_str(), A.row_offsets.data(), A.row_offsets.size());
  dump(dir, (s + "_cols").c_str(), A.packed_cols.data(), A.packed_cols.size());
  dump(dir, (s + "_vals").c_str(), A.packed_coefs.data(), A.packed_coefs.size());
}

static int run_mesh(int nx, int ny, int nz, const std::string& dir, int max_iters) {
  if (chdir(dir.c_str()) != 0) { std::perror("chdir"); return 2; }
#ifdef MINIFE_DEBUG
  outstream().precision(17);
#endif
  Box global_box = {{0, nx, 0, ny, 0, nz}};
  Box my_box = global_box;
  miniFE::Parameters params;
  miniFE::simple_mesh_description<int> mesh(global_box, my_box);

  std::vector<int> bc0(mesh.bc_rows_0.begin(), mesh.bc_rows_0.end());
  std::vector<int> bc1(mesh.bc_rows_1.begin(), mesh.bc_rows_1.end());
  dump(".", "bc_rows_0", bc0.data(), bc0.size());
  dump(".", "bc_rows_1", bc1.data(), bc1.size());

  Mat A;
  if (miniFE::generate_matrix_structure(mesh, A) != 0) return 3;
  int nrows = (int)A.rows.size();
  Vec b(nrows > 0 ? A.rows[0] : -1, nrows);
  Vec x(nrows > 0 ? A.rows[0] : -1, nrows);
  miniFE::assemble_FE_data(mesh, A, b, params);
  dump_matrix(".", "A_prebc", A);
  dump(".", "b_prebc", b.coefs.data(), b.coefs.size());

  miniFE::impose_dirichlet(0.0, A, b, nx + 1, ny + 1, nz + 1, mesh.bc_rows_0);
  dump_matrix(".", "A_bc0", A);
  dump(".", "b_bc0", b.coefs.data(), b.coefs.size());
  miniFE::impose_dirichlet(1.0, A, b, nx + 1, ny + 1, nz + 1, mesh.bc_rows_1);
  dump_matrix(".", "A", A);
  dump(".", "b", b.coefs.data(), b.coefs.size());
  miniFE::make_local_matrix(A);

  int num_iters = 0;
  double normr = 0;
  double tol = std::numeric_limits<double>::epsilon();
  miniFE::timer_type cg_times[miniFE::NUM_TIMERS];
  miniFE::cg_solve(A, b, x, miniFE::matvec_std<Mat, Vec>(), max_iters, tol,
                   num_iters, normr, cg_times);
  dump(".", "x", x.coefs.data(), x.coefs.size());

  int verify_rc = miniFE::verify_solution(mesh, x, 0.06, false);

  FILE* f = std::fopen("meta.json", "w");
  std::fprintf(f,
               "{\"nx\": %d, \"ny\": %d, \"nz\": %d, \"nrows\": %d, \"nnz\": %d, "
               "\"max_iters\": %d, \"num_iters\": %d, \"normr\": \"%a\", "
               "\"normr_dec\": %.17g, \"verify_rc\": %d}\n",
               nx, ny, nz, nrows, (int)A.num_nonzeros(), max_iters, num_iters,
               normr, normr, verify_rc);
  std::fclose(f);
  return 0;
}

static int run_elem(const char* in, const char* out) {
  FILE* fi = std::fopen(in, "rb");
  FILE* fo = std::fopen(out, "wb");
  if (!fi || !fo) { std::perror("elem io"); return 2; }
  miniFE::ElemData<int, double> ed;
  miniFE::compute_gradient_values(ed.grad_vals);
  double coords[24];
  while (std::fread(coords, sizeof(double), 24, fi) == 24) {
    std::memcpy(ed.elem_node_coords, coords, sizeof(coords));
    miniFE::compute_element_matrix_and_vector(ed);
    std::fwrite(ed.elem_diffusion_matrix, sizeof(double), 36, fo);
    std::fwrite(ed.elem_source_vector, sizeof(double), 8, fo);
  }
  std::fclose(fi);
  std::fclose(fo);
  return 0;
}

static int run_time(int nx, int ny, int nz, int cg_iters) {
  Box global_box = {{0, nx, 0, ny, 0, nz}};
  Box my_box = global_box;
  miniFE::Parameters params;
  double t0 = miniFE::mytimer();
  miniFE::simple_mesh_description<int> mesh(global_box, my_box);
  double t_mesh = miniFE::mytimer() - t0;
  Mat A;
  t0 = miniFE::mytimer();
  if (miniFE::generate_matrix_structure(mesh, A) != 0) return 3;
  double t_struct = miniFE::mytimer() - t0;
  int nrows = (int)A.rows.size();
  Vec b(A.rows[0], nrows), x(A.rows[0], nrows);
  t0 = miniFE::mytimer();
  miniFE::assemble_FE_data(mesh, A, b, params);
  double t_asm = miniFE::mytimer() - t0;
  t0 = miniFE::mytimer();
  miniFE::impose_dirichlet(0.0, A, b, nx + 1, ny + 1, nz + 1, mesh.bc_rows_0);
  miniFE::impose_dirichlet(1.0, A, b, nx + 1, ny + 1, nz + 1, mesh.bc_rows_1);
  double t_bc = miniFE::mytimer() - t0;
  miniFE::make_local_matrix(A);
  int num_iters = 0;
  double normr = 0, tol = std::numeric_limits<double>::epsilon();
  miniFE::timer_type cg_times[miniFE::NUM_TIMERS];
  miniFE::cg_solve(A, b, x, miniFE::matvec_std<Mat, Vec>(), cg_iters, tol, num_iters,
                   normr, cg_times);
  std::printf(
      "{\"nx\": %d, \"ny\": %d, \"nz\": %d, \"nrows\": %d, \"nnz\": %zu, \"mesh_s\": %.6f, "
      "\"struct_s\": %.6f, \"assembly_s\": %.6f, \"dirichlet_s\": %.6f, \"cg_iters\": %d, "
      "\"cg_total_s\": %.6f, \"cg_matvec_s\": %.6f, \"cg_dot_s\": %.6f, \"cg_waxpy_s\": %.6f, "
      "\"normr\": %.17g}\n",
      nx, ny, nz, nrows, A.num_nonzeros(), t_mesh, t_struct, t_asm, t_bc, num_iters,
      cg_times[miniFE::TOTAL], cg_times[miniFE::MATVEC], cg_times[miniFE::DOT],
      cg_times[miniFE::WAXPY], normr);
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 7 && !std::strcmp(argv[1], "mesh"))
    return run_mesh(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), argv[5],
                    std::atoi(argv[6]));
  if (argc >= 4 && !std::strcmp(argv[1], "elem")) return run_elem(argv[2], argv[3]);
  if (argc >= 6 && !std::strcmp(argv[1], "time"))
    return run_time(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]),
                    std::atoi(argv[5]));
  std::fprintf(stderr,
               "usage: %s mesh NX NY NZ OUTDIR MAXITER | elem IN OUT | time NX NY NZ ITERS\n",
               argv[0]);
  return 1;
}
