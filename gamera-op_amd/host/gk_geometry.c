/*
 * Grid coordinates, metric arrays and PDM reconstruction weights for
 * orthogonal Cartesian and spherical grids.
 *
 * The host computes the reference-layout 3-D arrays (problem plugins and the
 * vector-potential initialisation read them) with the reference formulas in
 * the same floating-point order, so initial states match the reference bit
 * for bit:
 *   uniform coordinates       src/solver_init.c:198-222
 *   Cartesian metrics         src/cartesian/cartesian_general.c:42-283
 *   Cartesian weights         src/cartesian/cartesian_general.c:286-367
 *   spherical metrics         src/spherical/spherical_general.c:42-316
 *   spherical weights (LU)    src/spherical/spherical_general.c:319-603,
 *                             math_lib/lu_solver.c:10-106
 * The device only needs per-axis tables; gk_tables.c extracts them.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "config.h"
#include "curvilinear.h"
#include "log.h"
#include "solver.h"

int gk_coord_type = 0;

double H1(double a, double b, double c) {
  (void)a; (void)b; (void)c;
  return 1.0;
}
double H2(double a, double b, double c) {
  (void)b; (void)c;
  return gk_coord_type == 2 ? a : 1.0;
}
double H3(double a, double b, double c) {
  (void)c;
  return gk_coord_type == 2 ? a * sin(b) : 1.0;
}

/* uniform block coordinates (src/solver_init.c:198-222) */
void gk_default_decomposition(void) {
  double dx1_proc = (x1max_global - x1min_global) / config.proc_dims[0];
  double dx2_proc = (x2max_global - x2min_global) / config.proc_dims[1];
  double dx3_proc = (x3max_global - x3min_global) / config.proc_dims[2];
  x1min = x1min_global + dx1_proc * proc_coords[0];
  x1max = x1min_global + dx1_proc * (proc_coords[0] + 1);
  x2min = x2min_global + dx2_proc * proc_coords[1];
  x2max = x2min_global + dx2_proc * (proc_coords[1] + 1);
  x3min = x3min_global + dx3_proc * proc_coords[2];
  x3max = x3min_global + dx3_proc * (proc_coords[2] + 1);
  for (int i = isg; i <= ieg + 1; i++)
    for (int j = jsg; j <= jeg + 1; j++)
      for (int k = ksg; k <= keg + 1; k++) {
        x1[i][j][k] = x1min - (x1max - x1min) * NO2 / config.ni + (i - isg) * (x1max - x1min) / config.ni;
        x2[i][j][k] = x2min - (x2max - x2min) * NO2 / config.nj + (j - jsg) * (x2max - x2min) / config.nj;
        x3[i][j][k] = x3min - (x3max - x3min) * NO2 / config.nk + (k - ksg) * (x3max - x3min) / config.nk;
      }
}

#define LOOP3(i0, i1, j0, j1, k0, k1)                  \
  _Pragma("omp parallel for collapse(2) schedule(static)") \
  for (int i = (i0); i <= (i1); i++)                    \
    for (int j = (j0); j <= (j1); j++)                  \
      for (int k = (k0); k <= (k1); k++)

/* 8-point face sums of the face-volume and edge-area arrays */
static void fac8th_sums(void) {
  LOOP3(is, ie + 1, jsg, jeg, ksg, keg) {
    double sum = 0.0;
    for (int o = -4; o <= 3; o++) sum += geo[vol_center][i + o][j][k] * fac8th[o + 4];
    geo[vol_idir][i][j][k] = sum;
  }
  LOOP3(isg, ieg, js, je + 1, ksg, keg) {
    double sum = 0.0;
    for (int o = -4; o <= 3; o++) sum += geo[vol_center][i][j + o][k] * fac8th[o + 4];
    geo[vol_jdir][i][j][k] = sum;
  }
  LOOP3(isg, ieg, jsg, jeg, ks, ke + 1) {
    double sum = 0.0;
    for (int o = -4; o <= 3; o++) sum += geo[vol_center][i][j][k + o] * fac8th[o + 4];
    geo[vol_kdir][i][j][k] = sum;
  }
  LOOP3(is, ie + 1, js, je + 1, ks, ke) {
    double sI = 0.0, sJ = 0.0;
    for (int o = -4; o <= 3; o++) {
      sI += geo[face_idir][i][j + o][k] * fac8th[o + 4];
      sJ += geo[face_jdir][i + o][j][k] * fac8th[o + 4];
    }
    geo[IfaceAedgeK][i][j][k] = sI;
    geo[JfaceAedgeK][i][j][k] = sJ;
  }
  LOOP3(is, ie, js, je + 1, ks, ke + 1) {
    double sJ = 0.0, sK = 0.0;
    for (int o = -4; o <= 3; o++) {
      sJ += geo[face_jdir][i][j][k + o] * fac8th[o + 4];
      sK += geo[face_kdir][i][j + o][k] * fac8th[o + 4];
    }
    geo[JfaceAedgeI][i][j][k] = sJ;
    geo[KfaceAedgeI][i][j][k] = sK;
  }
  LOOP3(is, ie + 1, js, je, ks, ke + 1) {
    double sK = 0.0, sI = 0.0;
    for (int o = -4; o <= 3; o++) {
      sK += geo[face_kdir][i + o][j][k] * fac8th[o + 4];
      sI += geo[face_idir][i][j][k + o] * fac8th[o + 4];
    }
    geo[KfaceAedgeJ][i][j][k] = sK;
    geo[IfaceAedgeJ][i][j][k] = sI;
  }
}

static void centres(void) {
  LOOP3(isg, ieg, jsg, jeg + 1, ksg, keg + 1) x1c[i][j][k] = 0.5 * (x1[i][j][k] + x1[i + 1][j][k]);
  LOOP3(isg, ieg + 1, jsg, jeg, ksg, keg + 1) x2c[i][j][k] = 0.5 * (x2[i][j][k] + x2[i][j + 1][k]);
  LOOP3(isg, ieg + 1, jsg, jeg + 1, ksg, keg) x3c[i][j][k] = 0.5 * (x3[i][j][k] + x3[i][j][k + 1]);
}

static void cartesian_geometry(void) {
  centres();
  LOOP3(isg, ieg, jsg, jeg + 1, ksg, keg + 1) {
    dx1[i][j][k] = x1[i + 1][j][k] - x1[i][j][k];
    geo[edge_idir][i][j][k] = dx1[i][j][k];
  }
  LOOP3(isg, ieg + 1, jsg, jeg, ksg, keg + 1) {
    dx2[i][j][k] = x2[i][j + 1][k] - x2[i][j][k];
    geo[edge_jdir][i][j][k] = dx2[i][j][k];
  }
  LOOP3(isg, ieg + 1, jsg, jeg + 1, ksg, keg) {
    dx3[i][j][k] = x3[i][j][k + 1] - x3[i][j][k];
    geo[edge_kdir][i][j][k] = dx3[i][j][k];
  }
  LOOP3(isg, ieg, jsg, jeg, ksg, keg) {
    geo[dxi][i][j][k] = dx1[i][j][k];
    geo[dxj][i][j][k] = dx2[i][j][k];
    geo[dxk][i][j][k] = dx3[i][j][k];
  }
  LOOP3(isg, ieg + 1, jsg, jeg, ksg, keg) geo[face_idir][i][j][k] = dx2[i][j][k] * dx3[i][j][k];
  LOOP3(isg, ieg, jsg, jeg + 1, ksg, keg) geo[face_jdir][i][j][k] = dx1[i][j][k] * dx3[i][j][k];
  LOOP3(isg, ieg, jsg, jeg, ksg, keg + 1) geo[face_kdir][i][j][k] = dx1[i][j][k] * dx2[i][j][k];
  LOOP3(isg, ieg, jsg, jeg, ksg, keg) geo[vol_center][i][j][k] = dx1[i][j][k] * dx2[i][j][k] * dx3[i][j][k];
  LOOP3(isg, ieg + 1, jsg, jeg, ksg, keg) geo[GF_face_idir][i][j][k] = geo[face_idir][i][j][k];
  LOOP3(isg, ieg, jsg, jeg + 1, ksg, keg) geo[GF_face_jdir][i][j][k] = geo[face_jdir][i][j][k];
  LOOP3(isg, ieg, jsg, jeg, ksg, keg + 1) geo[GF_face_kdir][i][j][k] = geo[face_kdir][i][j][k];
  LOOP3(isg, ieg, jsg, jeg, ksg, keg) {
    geo[GF_vol_idir][i][j][k] = geo[vol_center][i][j][k];
    geo[GF_vol_jdir][i][j][k] = geo[vol_center][i][j][k];
    geo[GF_vol_kdir][i][j][k] = geo[vol_center][i][j][k];
    x1ctr[i][j][k] = x1c[i][j][k];
    x2ctr[i][j][k] = x2c[i][j][k];
    x3ctr[i][j][k] = x3c[i][j][k];
  }
  fac8th_sums();
}

static void spherical_geometry(void) {
  centres();
  LOOP3(isg, ieg, jsg, jeg + 1, ksg, keg + 1) {
    dx1[i][j][k] = x1[i + 1][j][k] - x1[i][j][k];
    geo[edge_idir][i][j][k] = dx1[i][j][k];
  }
  LOOP3(isg, ieg + 1, jsg, jeg, ksg, keg + 1) {
    dx2[i][j][k] = x2[i][j + 1][k] - x2[i][j][k];
    geo[edge_jdir][i][j][k] = x1[i][j][k] * dx2[i][j][k];
  }
  LOOP3(isg, ieg + 1, jsg, jeg + 1, ksg, keg) {
    dx3[i][j][k] = x3[i][j][k + 1] - x3[i][j][k];
    geo[edge_kdir][i][j][k] = x1[i][j][k] * sin(x2[i][j][k]) * dx3[i][j][k];
  }
  LOOP3(isg, ieg, jsg, jeg, ksg, keg) {
    geo[dxi][i][j][k] = dx1[i][j][k];
    geo[dxj][i][j][k] = x1c[i][j][k] * dx2[i][j][k];
    geo[dxk][i][j][k] = x1c[i][j][k] * sin(x2c[i][j][k]) * dx3[i][j][k];
  }
  LOOP3(isg, ieg + 1, jsg, jeg, ksg, keg)
  geo[face_idir][i][j][k] = -x1[i][j][k] * x1[i][j][k] * (cos(x2[i][j + 1][k]) - cos(x2[i][j][k])) * dx3[i][j][k];
  LOOP3(isg, ieg, jsg, jeg + 1, ksg, keg)
  geo[face_jdir][i][j][k] = 0.5 * (x1[i + 1][j][k] * x1[i + 1][j][k] - x1[i][j][k] * x1[i][j][k]) * sin(x2[i][j][k]) * dx3[i][j][k];
  LOOP3(isg, ieg, jsg, jeg, ksg, keg + 1)
  geo[face_kdir][i][j][k] = 0.5 * (x1[i + 1][j][k] * x1[i + 1][j][k] - x1[i][j][k] * x1[i][j][k]) * dx2[i][j][k];
  LOOP3(isg, ieg, jsg, jeg, ksg, keg)
  geo[vol_center][i][j][k] = fabs(1.0 / 3.0 * (pow(x1[i + 1][j][k], 3) - pow(x1[i][j][k], 3)) *
                                  (cos(x2[i][j + 1][k]) - cos(x2[i][j][k])) * dx3[i][j][k]);
  LOOP3(isg, ieg + 1, jsg, jeg, ksg, keg) geo[GF_face_idir][i][j][k] = x1[i][j][k] * x1[i][j][k];
  LOOP3(isg, ieg, jsg, jeg + 1, ksg, keg) geo[GF_face_jdir][i][j][k] = sin(x2[i][j][k]);
  LOOP3(isg, ieg, jsg, jeg, ksg, keg + 1) geo[GF_face_kdir][i][j][k] = 1.0;
  LOOP3(isg, ieg, jsg, jeg, ksg, keg) {
    geo[GF_vol_idir][i][j][k] = 1.0 / 3.0 * (pow(x1[i + 1][j][k], 3) - pow(x1[i][j][k], 3));
    geo[GF_vol_jdir][i][j][k] = -(cos(x2[i][j + 1][k]) - cos(x2[i][j][k]));
    geo[GF_vol_kdir][i][j][k] = dx3[i][j][k];
  }
  /* volume centroids */
  LOOP3(isg, ieg, jsg, jeg + 1, ksg, keg + 1)
  x1ctr[i][j][k] = x1c[i][j][k] + 2.0 * x1c[i][j][k] * dx1[i][j][k] * dx1[i][j][k] /
                                      (12.0 * x1c[i][j][k] * x1c[i][j][k] + dx1[i][j][k] * dx1[i][j][k]);
  LOOP3(isg, ieg + 1, jsg, jeg, ksg, keg + 1)
  x2ctr[i][j][k] = (sin(x2[i][j + 1][k]) - sin(x2[i][j][k]) -
                    (x2[i][j + 1][k] * cos(x2[i][j + 1][k]) - x2[i][j][k] * cos(x2[i][j][k]))) /
                   (-cos(x2[i][j + 1][k]) + cos(x2[i][j][k]));
  LOOP3(isg, ieg + 1, jsg, jeg + 1, ksg, keg) x3ctr[i][j][k] = x3c[i][j][k];
  fac8th_sums();
}

int set_geometry_arrays() {
  if (gk_coord_type == 2)
    spherical_geometry();
  else
    cartesian_geometry();
  return 0;
}

/* ------------------------------------------------------------------------ */
/* reconstruction weights */
static const double pdm7_left[7] = {-1.0 / 140.0, 5.0 / 84.0, -101.0 / 420.0, 319.0 / 420.0,
                                    107.0 / 210.0, -19.0 / 210.0, 1.0 / 105.0};

static void cartesian_weights(void) {
  /* w = facPDMU7 * V / (8-point face volume); right weights mirror the left */
  LOOP3(is, ie + 1, jsg, jeg + 1, ksg, keg + 1) {
    for (int n = 0; n < 7; n++) {
      rec[w1l_first_direction + n][i][j][k] = pdm7_left[n] * geo[vol_center][i][j][k] / geo[vol_idir][i][j][k];
      rec[w1r_first_direction + n][i][j][k] = pdm7_left[6 - n] * geo[vol_center][i][j][k] / geo[vol_idir][i][j][k];
    }
  }
  LOOP3(isg, ieg + 1, js, je + 1, ksg, keg + 1) {
    for (int n = 0; n < 7; n++) {
      rec[w1l_second_direction + n][i][j][k] = pdm7_left[n] * geo[vol_center][i][j][k] / geo[vol_jdir][i][j][k];
      rec[w1r_second_direction + n][i][j][k] = pdm7_left[6 - n] * geo[vol_center][i][j][k] / geo[vol_jdir][i][j][k];
    }
  }
  LOOP3(isg, ieg + 1, jsg, jeg + 1, ks, ke + 1) {
    for (int n = 0; n < 7; n++) {
      rec[w1l_third_direction + n][i][j][k] = pdm7_left[n] * geo[vol_center][i][j][k] / geo[vol_kdir][i][j][k];
      rec[w1r_third_direction + n][i][j][k] = pdm7_left[6 - n] * geo[vol_center][i][j][k] / geo[vol_kdir][i][j][k];
    }
  }
}

/* partial-pivot LU, solved in place (math_lib/lu_solver.c:10-106) */
static int lu5(double A[5][5], const double b[5], double x[5]) {
  int perm[5];
  for (int i = 0; i < 5; i++) perm[i] = i;
  for (int k = 0; k < 4; k++) {
    double max_val = fabs(A[k][k]);
    int max_row = k;
    for (int i = k + 1; i < 5; i++)
      if (fabs(A[i][k]) > max_val) {
        max_val = fabs(A[i][k]);
        max_row = i;
      }
    if (max_val < 1e-12) return -1;
    if (max_row != k) {
      for (int j = 0; j < 5; j++) {
        double t = A[k][j];
        A[k][j] = A[max_row][j];
        A[max_row][j] = t;
      }
      int tp = perm[k];
      perm[k] = perm[max_row];
      perm[max_row] = tp;
    }
    for (int i = k + 1; i < 5; i++) {
      double factor = A[i][k] / A[k][k];
      A[i][k] = factor;
      for (int j = k + 1; j < 5; j++) A[i][j] -= factor * A[k][j];
    }
  }
  if (fabs(A[4][4]) < 1e-12) return -1;
  double y[5];
  for (int i = 0; i < 5; i++) {
    y[i] = b[perm[i]];
    for (int j = 0; j < i; j++) y[i] -= A[i][j] * y[j];
  }
  for (int i = 4; i >= 0; i--) {
    x[i] = y[i];
    for (int j = i + 1; j < 5; j++) x[i] -= A[i][j] * x[j];
    x[i] /= A[i][i];
  }
  return 0;
}

double factorial(int k) {
  if (k < 0) return 0.0;
  double r = 1.0;
  for (int i = 2; i <= k; ++i) r *= i;
  return r;
}

double nchoosek(int j, int k) {
  if (k < 0 || k > j) return 0.0;
  if (k == 0 || k == j) return 1.0;
  k = (k < j - k) ? k : j - k;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r *= (j - k + i) / (double)i;
  return r;
}

static void store_w5(int first_rec, int idx_axis, int idx, const double *x) {
  /* 5-point weights in slots 2..6 of the 7-slot layout, 1 and 7 zero */
  int i0 = isg, i1 = ieg + 1, j0 = jsg, j1 = jeg + 1, k0 = ksg, k1 = keg + 1;
  if (idx_axis == 0) i0 = i1 = idx;
  if (idx_axis == 1) j0 = j1 = idx;
  for (int i = i0; i <= i1; i++)
    for (int j = j0; j <= j1; j++)
      for (int k = k0; k <= k1; k++) {
        rec[first_rec][i][j][k] = 0.0;
        for (int n = 0; n < 5; n++) rec[first_rec + 1 + n][i][j][k] = x ? x[n] : 0.0;
        rec[first_rec + 6][i][j][k] = 0.0;
      }
}

/* radial weights: 5-point conservative reconstruction in r^2 dr measure */
static void spherical_weights_r(void) {
  int n = config.ni + 2 * NO2 + 1 + 2 * NO2;
  double *xs = (double *)calloc((size_t)n, sizeof(double));
  for (int i = isg + NO2; i <= ieg + NO2 + 1; i++) xs[i] = x1[i - NO2][0][0];
  for (int i = isg + NO2 - 1; i >= 0; i--) xs[i] = xs[i + 1] - dx1[0][0][0];
  for (int i = ieg + NO2 + 2; i <= ieg + 2 * NO2 + 1; i++) xs[i] = xs[i - 1] + dx1[ieg][0][0];
  const int Order = 5, m = 2, iL = -Order / 2, iR = iL + Order - 1;
  for (int side = 0; side < 2; side++) {
#pragma omp parallel for schedule(dynamic)
    for (int idx = isg; idx <= ieg + 1; idx++) {
      double A[5][5], b[5], x[5];
      for (int i = iL; i <= iR; i++) {
        double xm = xs[idx + NO2 + i - 1 + side], xp = xs[idx + NO2 + i + side];
        for (int j = 0; j <= Order - 1; j++)
          A[j][i + abs(iL)] = (double)(m + 1) / (double)(j + m + 1) *
                              (pow(xp, j + m + 1) - pow(xm, j + m + 1)) / (pow(xp, m + 1) - pow(xm, m + 1));
        b[i + abs(iL)] = pow(xs[idx + NO2 + 0], i + abs(iL));
      }
      int ok = lu5(A, b, x) == 0;
      store_w5(side == 0 ? w1l_first_direction : w1r_first_direction, 0, idx, ok ? x : NULL);
    }
  }
  free(xs);
}

/* polar weights: 5-point conservative reconstruction in sin(theta) measure */
static void spherical_weights_theta(void) {
  int n = config.nj + 2 * NO2 + 1 + 2 * NO2;
  double *xs = (double *)calloc((size_t)n, sizeof(double));
  for (int j = jsg + NO2; j <= jeg + NO2 + 1; j++) xs[j] = x2[0][j - NO2][0];
  for (int j = jsg + NO2 - 1; j >= 0; j--) xs[j] = xs[j + 1] - dx2[0][0][0];
  for (int j = jeg + NO2 + 2; j <= jeg + 2 * NO2 + 1; j++) xs[j] = xs[j - 1] + dx2[0][jeg][0];
  const int Order = 5, iL = -Order / 2, iR = iL + Order - 1;
  for (int side = 0; side < 2; side++) {
    int last = side == 0 ? jeg + 1 : jeg;
#pragma omp parallel for schedule(dynamic)
    for (int idx = jsg; idx <= last; idx++) {
      double A[5][5], b[5], x[5];
      for (int i = iL; i <= iR; i++) {
        double xm = xs[idx + NO2 + i - 1 + side], xp = xs[idx + NO2 + i + side];
        double sum1 = 0.0;
        for (int j = 0; j <= Order - 1; j++) {
          for (int k = 0; k <= j; k++)
            sum1 = sum1 + factorial(k) * nchoosek(j, k) *
                              (pow(xm, j - k) * cos(xm + k * PI / 2.0) - pow(xp, j - k) * cos(xp + k * PI / 2.0));
          A[j][i + abs(iL)] = sum1 / (cos(xm) - cos(xp));
          sum1 = 0;
        }
        b[i + abs(iL)] = pow(xs[idx + NO2 + 0], i + abs(iL));
      }
      if (lu5(A, b, x) == 0) {
        store_w5(side == 0 ? w1l_second_direction : w1r_second_direction, 1, idx, x);
      } else {
        /* the reference zeroes the slots at i = isg on failure
           (src/spherical/spherical_general.c:493-500); keep that behaviour */
        for (int jj = jsg; jj <= jeg + 1; jj++)
          for (int kk = ksg; kk <= keg + 1; kk++)
            for (int q = 0; q < 7; q++)
              rec[(side == 0 ? w1l_second_direction : w1r_second_direction) + q][isg][jj][kk] = 0.0;
        store_w5(side == 0 ? w1l_second_direction : w1r_second_direction, 1, idx, x);
      }
    }
  }
  free(xs);
}

static void spherical_weights_phi(void) {
  static const double wl[7] = {0.0, 1.0 / 30.0, -13.0 / 60.0, 47.0 / 60.0, 9.0 / 20.0, -1.0 / 20.0, 0.0};
  static const double wr[7] = {0.0, -1.0 / 20.0, 9.0 / 20.0, 47.0 / 60.0, -13.0 / 60.0, 1.0 / 30.0, 0.0};
  LOOP3(isg, ieg + 1, jsg, jeg + 1, ksg, keg + 1) {
    for (int n = 0; n < 7; n++) {
      rec[w1l_third_direction + n][i][j][k] = wl[n];
      rec[w1r_third_direction + n][i][j][k] = wr[n];
    }
  }
}

void getweights() {
  if (gk_coord_type == 2) {
    spherical_weights_r();
    spherical_weights_theta();
    spherical_weights_phi();
  } else {
    cartesian_weights();
  }
}

/* ------------------------------------------------------------------------ */
/* coordinate transforms (src/cartesian/cartesian_general.c:397-425,
   src/spherical/spherical_general.c:731-795) */
void cartesian_to_curvilinear_coord(double x, double y, double z, double *a, double *b, double *c) {
  if (gk_coord_type != 2) {
    *a = x;
    *b = y;
    *c = z;
    return;
  }
  *a = sqrt(x * x + y * y + z * z);
  if (*a == 0.0) {
    *b = 0.0;
    *c = 0.0;
    return;
  }
  *b = acos(z / *a);
  *c = atan2(y, x);
  if (*c < 0.0) *c += 2.0 * PI;
}

void curvilinear_to_cartesian_coord(double a, double b, double c, double *x, double *y, double *z) {
  if (gk_coord_type != 2) {
    *x = a;
    *y = b;
    *z = c;
    return;
  }
  *x = a * sin(b) * cos(c);
  *y = a * sin(b) * sin(c);
  *z = a * cos(b);
}

void cartesian_to_curvilinear_vector(double x, double y, double z, double vx, double vy, double vz,
                                     double *v1, double *v2, double *v3) {
  if (gk_coord_type != 2) {
    *v1 = vx;
    *v2 = vy;
    *v3 = vz;
    return;
  }
  double r, th, ph;
  cartesian_to_curvilinear_coord(x, y, z, &r, &th, &ph);
  double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
  *v1 = vx * st * cp + vy * st * sp + vz * ct;
  *v2 = vx * ct * cp + vy * ct * sp - vz * st;
  *v3 = -vx * sp + vy * cp;
}

void curvilinear_to_cartesian_vector(double a, double b, double c, double v1, double v2, double v3,
                                     double *vx, double *vy, double *vz) {
  if (gk_coord_type != 2) {
    *vx = v1;
    *vy = v2;
    *vz = v3;
    return;
  }
  double st = sin(b), ct = cos(b), sp = sin(c), cp = cos(c);
  (void)a;
  *vx = v1 * st * cp + v2 * ct * cp - v3 * sp;
  *vy = v1 * st * sp + v2 * ct * sp + v3 * cp;
  *vz = v1 * ct - v2 * st;
}
