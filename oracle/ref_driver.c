/*
 * Fixture driver for the reference CPU solver (oracle/_ref).
 *
 * Linked against the reference's own sources (everything under src/ except
 * src/main.c) plus a problem plugin; it replaces src/main.c:14-99 with the
 * same call sequence (MPI init, set_problem_config, initialize_config,
 * setup_comm_and_types, initialize_solver, getweights, problem_init,
 * set_background_field, the ring-average set-up when the plugin turns it on,
 * solve(Nt)) and then writes the solver's global arrays
 * as raw little-endian float64 records so tests can read them with numpy.
 *
 * Usage: ref_driver <config.yaml> <out.bin> <dump_geometry 0|1>
 *   The seeded plugin takes its problem from the GK_PROBLEM / GK_SEED
 *   environment; built with a shipped problem file (-DPROBLEM_<name>) the
 *   driver runs that file instead (oracle/Makefile, target "shipped").
 *   Run it in an empty directory: initialize_solver restarts from any
 *   mhd_*.h5 file it finds in the working directory.
 *
 * Record format: char name[32]; int64 n; double data[n]; repeated.
 * This is test infrastructure only; nothing in the product links it.
 */
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "log.h"

#include "config.h"
#include "curvilinear.h"
#include "problem.h"
#include "setup_mpi.h"
#include "solver.h"

#ifdef PROBLEM_GK_SEEDED
int gk_problem_custom_grid(void);
#endif

static FILE *out;

static void put(const char *name, const double *data, int64_t n) {
  char buf[32];
  memset(buf, 0, sizeof buf);
  strncpy(buf, name, sizeof buf - 1);
  fwrite(buf, 1, sizeof buf, out);
  fwrite(&n, sizeof n, 1, out);
  fwrite(data, sizeof(double), (size_t)n, out);
}

static void put_scalar(const char *name, double v) { put(name, &v, 1); }

int main(int argc, char *argv[]) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s <config.yaml> <out.bin> <dump_geometry>\n",
            argv[0]);
    return 1;
  }
  int dump_geometry = argv[3][0] == '1';
  MPI_Init(&argc, &argv);
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  MPI_Comm_size(MPI_COMM_WORLD, &size);
  set_problem_config();
  if (initialize_config(rank, size, argv[1]) != 0) return 2;
  if (setup_comm_and_types(rank) != 0) return 3;
  if (initialize_logging(BASE_LOG_NAME, rank) != 0) return 4;
  if (initialize_solver(BASE_DAT_NAME) != 0) return 5;
#ifdef PROBLEM_GK_SEEDED
  if (gk_problem_custom_grid()) set_geometry_arrays();
#endif
  getweights();
  if (!read_restart) problem_init();
  set_background_field();
  if (doRingAverage == 1) { /* src/main.c:66-70 */
    set_RingAverage_geo_data();
    set_RingAverage_config();
  }
  dtout = 10.0;
  if (solve(Nt) != 0) return 6;

  out = fopen(argv[2], "wb");
  if (!out) return 7;
  size_t n3 = (size_t)config.NI * config.NJ * config.NK;
  put_scalar("ni", config.ni);
  put_scalar("nj", config.nj);
  put_scalar("nk", config.nk);
  put_scalar("dt", dt);
  put_scalar("dt0", dt0);
  put_scalar("time_sim", time_sim);
  put_scalar("divB_max", divB_max);
  put_scalar("CA", CA);
  put("gas", &gas[0][0][0][0][0], (int64_t)(NF_gas_store * n3));
  put("gem", &gem[0][0][0][0], (int64_t)(NF_gem_store * n3));
  put("magtot", &gem[magtot_b1][0][0][0], (int64_t)(3 * n3));
  if (dump_geometry) {
    put("geo", &geo[0][0][0][0], (int64_t)(NF_geo * n3));
    put("rec", &rec[0][0][0][0], (int64_t)(NF_rec * n3));
    put("x1", &x1[0][0][0], (int64_t)n3);
    put("x2", &x2[0][0][0], (int64_t)n3);
    put("x3", &x3[0][0][0], (int64_t)n3);
    put("x1c", &x1c[0][0][0], (int64_t)n3);
    put("x2c", &x2c[0][0][0], (int64_t)n3);
    put("x3c", &x3c[0][0][0], (int64_t)n3);
    put("x1ctr", &x1ctr[0][0][0], (int64_t)n3);
    put("x2ctr", &x2ctr[0][0][0], (int64_t)n3);
    put("x3ctr", &x3ctr[0][0][0], (int64_t)n3);
    put("dx1", &dx1[0][0][0], (int64_t)n3);
    put("dx2", &dx2[0][0][0], (int64_t)n3);
    put("dx3", &dx3[0][0][0], (int64_t)n3);
    put("B0", &gem[B0_b1][0][0][0], (int64_t)((dmagsf_G_x3dir - B0_b1 + 1) * n3));
  }
  fclose(out);
  finalize_solver();
  free_mpi_types();
  fclose(log_file);
  MPI_Finalize();
  return 0;
}
