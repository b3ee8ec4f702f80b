/*
 * Set-up driver for a reference problem file linked against the host facade
 * (test infrastructure; built into oracle/_ref by oracle/Makefile, target
 * "shipped").
 *
 * It runs the set-up half of the reference main() (src/main.c:30-70:
 * set_problem_config, initialize_config, initialize_logging,
 * initialize_solver, getweights, problem_init, set_background_field, the
 * ring-average set-up) on the host, without MPI and without solve(), and
 * writes the host arrays in the record format of oracle/ref_driver.c
 * (char name[32]; int64 n; double data[n]), so a test can compare them with
 * the compiled reference's arrays for the same problem file and config.
 *
 * Usage: facade_setup_<P> <config.yaml> <out.bin>
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "log.h"

#include "config.h"
#include "curvilinear.h"
#include "problem.h"
#include "solver.h"

int setup_comm_and_types(int rank);

static FILE *out;

static void put(const char *name, const double *data, int64_t n) {
  char buf[32];
  memset(buf, 0, sizeof buf);
  strncpy(buf, name, sizeof buf - 1);
  fwrite(buf, 1, sizeof buf, out);
  fwrite(&n, sizeof n, 1, out);
  fwrite(data, sizeof(double), (size_t)n, out);
}

int main(int argc, char *argv[]) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s <config.yaml> <out.bin>\n", argv[0]);
    return 1;
  }
  set_problem_config();
  if (initialize_config(0, 1, argv[1]) != 0) return 2;
  if (setup_comm_and_types(0) != 0) return 3;
  if (initialize_logging(BASE_LOG_NAME, 0) != 0) return 4;
  if (initialize_solver(BASE_DAT_NAME) != 0) return 5;
  getweights();
  if (!read_restart) problem_init();
  set_background_field();
  if (doRingAverage == 1) {
    set_RingAverage_geo_data();
    set_RingAverage_config();
  }
  out = fopen(argv[2], "wb");
  if (!out) return 7;
  const int64_t n3 = (int64_t)config.NI * config.NJ * config.NK;
  double v;
  v = config.ni; put("ni", &v, 1);
  v = config.nj; put("nj", &v, 1);
  v = config.nk; put("nk", &v, 1);
  v = CA; put("CA", &v, 1);
  char name[32];
  for (int f = 0; f < 6; f++) {
    snprintf(name, sizeof name, "gas%d", f);
    put(name, &gas[0][f][0][0][0], n3);
  }
  for (int f = 0; f < 3; f++) {
    snprintf(name, sizeof name, "gem%d", f);
    put(name, &gem[f][0][0][0], n3);
  }
  for (int f = 0; f < NF_geo; f++) {
    snprintf(name, sizeof name, "geo%d", f);
    put(name, &geo[f][0][0][0], n3);
  }
  for (int f = 0; f < NF_rec; f++) {
    snprintf(name, sizeof name, "rec%d", f);
    put(name, &rec[f][0][0][0], n3);
  }
  for (int f = B0_b1; f <= dmagsf_G_x3dir; f++) {
    snprintf(name, sizeof name, "gem%d", f);
    put(name, &gem[f][0][0][0], n3);
  }
  put("x1", &x1[0][0][0], n3);
  put("x2", &x2[0][0][0], n3);
  put("x3", &x3[0][0][0], n3);
  fclose(out);
  finalize_solver();
  return 0;
}
