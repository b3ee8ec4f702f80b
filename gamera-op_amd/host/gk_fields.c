/*
 * Host state arrays, magnetic-field initialisation, background field and the
 * boundary-condition library of the host facade.
 *
 *   allocation (reference layout)   src/array.c:87-164
 *   AinitB / GaussianLineIntegral   src/solver_func.c:15-125
 *   background field quadratures    src/solver_func.c:1623-2077
 *   boundary-condition library      src/boundary_conditions.c:9-774
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "common.h"
#include "config.h"
#include "curvilinear.h"
#include "gk_host.h"
#include "log.h"
#include "solver.h"

double Field_Default(double x, double y, double z) {
  (void)x; (void)y; (void)z;
  return 0.0;
}
Field A1 = Field_Default, A2 = Field_Default, A3 = Field_Default;
Field B10 = Field_Default, B20 = Field_Default, B30 = Field_Default;
Field Source1_B0 = Field_Default, Source2_B0 = Field_Default, Source3_B0 = Field_Default;

double B10_uniform(double x, double y, double z) { (void)x; (void)y; (void)z; return 1 / sqrt(2); }
double B20_uniform(double x, double y, double z) { (void)x; (void)y; (void)z; return 1 / sqrt(2); }
double B30_uniform(double x, double y, double z) { (void)x; (void)y; (void)z; return 1 / sqrt(2); }

/* ------------------------------------------------------------------------ */
/* reference-layout allocation: a field is a contiguous [NI][NJ][NK] block
   reachable through [i][j][k] pointer tables */
static double ***view3(double *data, int ni, int nj, int nk) {
  double ***a = (double ***)malloc(sizeof(double **) * (size_t)ni);
  double **rows = (double **)malloc(sizeof(double *) * (size_t)ni * nj);
  if (!a || !rows) {
    free(a);
    free(rows);
    return NULL;
  }
  for (int i = 0; i < ni; i++) {
    a[i] = rows + (size_t)i * nj;
    for (int j = 0; j < nj; j++) a[i][j] = data + ((size_t)i * nj + j) * nk;
  }
  return a;
}

static void free_view3(double ***a) {
  if (!a) return;
  free(a[0]);
  free(a);
}

static size_t n3(void) { return (size_t)config.NI * config.NJ * config.NK; }

double ***gk_alloc3(void) {
  double *d = (double *)calloc(n3(), sizeof(double));
  if (!d) return NULL;
  double ***v = view3(d, config.NI, config.NJ, config.NK);
  if (!v) free(d);
  return v;
}

void gk_free3(double ***a) {
  if (!a) return;
  free(a[0][0]);
  free_view3(a);
}

/* 4-D table over nf fields; only fields with keep[f] get storage (one
   contiguous block for runs of kept fields so [0..17] stays contiguous) */
double ****gk_alloc4(int nf, const unsigned char *keep) {
  double ****a = (double ****)calloc((size_t)nf, sizeof(double ***));
  if (!a) return NULL;
  int f = 0;
  while (f < nf) {
    if (keep && !keep[f]) {
      f++;
      continue;
    }
    int g = f;
    while (g < nf && (!keep || keep[g])) g++;
    double *d = (double *)calloc(n3() * (size_t)(g - f), sizeof(double));
    if (!d) return a; /* caller checks the kept fields */
    for (int q = f; q < g; q++) a[q] = view3(d + n3() * (size_t)(q - f), config.NI, config.NJ, config.NK);
    f = g;
  }
  return a;
}

void gk_free4(double ****a, int nf) {
  if (!a) return;
  /* free each contiguous run once: the first field of a run owns the data */
  double *prev_end = NULL;
  for (int f = 0; f < nf; f++) {
    if (!a[f]) {
      prev_end = NULL;
      continue;
    }
    double *d = a[f][0][0];
    if (d != prev_end) free(d);
    prev_end = d + n3();
    free_view3(a[f]);
  }
  free(a);
}

/* ------------------------------------------------------------------------ */
/* 12-point Gauss-Legendre line average of fx from a to b (the sum over +-A
   nodes halved), src/solver_func.c:97-125 */
double GaussianLineIntegral(double (*fx)(double, double, double), double xa, double ya, double za,
                            double xb, double yb, double zb) {
  static const double A[] = {0.1252334085, 0.3678314989, 0.5873179542,
                             0.7699026741, 0.9041172563, 0.9815606342};
  static const double WT[] = {0.2491470458, 0.2334925365, 0.2031674267,
                              0.1600783285, 0.1069393259, 0.0471753363};
  double dx = (xb - xa) / 2.0, dy = (yb - ya) / 2.0, dz = (zb - za) / 2.0;
  double xbar = (xa + xb) / 2.0, ybar = (ya + yb) / 2.0, zbar = (za + zb) / 2.0;
  double sum = 0;
  for (int k = 0; k < 6; k++)
    sum += WT[k] * (fx(xbar + A[k] * dx, ybar + A[k] * dy, zbar + A[k] * dz) +
                    fx(xbar - A[k] * dx, ybar - A[k] * dy, zbar - A[k] * dz));
  sum /= 2.0;
  return sum;
}

/* face B from the curl of an edge vector potential (Stokes), so the discrete
   div B of the initial state is zero to roundoff; src/solver_func.c:15-95 */
void AinitB(double ***magi, double ***magj, double ***magk, Field FA1, Field FA2, Field FA3,
            Field FB10, Field FB20, Field FB30) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = isg; i <= ieg; i++)
    for (int j = jsg; j <= jeg + 1; j++)
      for (int k = ksg; k <= keg + 1; k++)
        gem[LAi][i][j][k] = geo[edge_idir][i][j][k] *
                            GaussianLineIntegral(*FA1, x1[i + 1][j][k], x2[i + 1][j][k], x3[i + 1][j][k],
                                                 x1[i][j][k], x2[i][j][k], x3[i][j][k]);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = isg; i <= ieg + 1; i++)
    for (int j = jsg; j <= jeg; j++)
      for (int k = ksg; k <= keg + 1; k++)
        gem[LAj][i][j][k] = geo[edge_jdir][i][j][k] *
                            GaussianLineIntegral(*FA2, x1[i][j + 1][k], x2[i][j + 1][k], x3[i][j + 1][k],
                                                 x1[i][j][k], x2[i][j][k], x3[i][j][k]);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = isg; i <= ieg + 1; i++)
    for (int j = jsg; j <= jeg + 1; j++)
      for (int k = ksg; k <= keg; k++)
        gem[LAk][i][j][k] = geo[edge_kdir][i][j][k] *
                            GaussianLineIntegral(*FA3, x1[i][j][k + 1], x2[i][j][k + 1], x3[i][j][k + 1],
                                                 x1[i][j][k], x2[i][j][k], x3[i][j][k]);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = isg; i <= ieg + 1; i++)
    for (int j = jsg; j <= jeg; j++)
      for (int k = ksg; k <= keg; k++) {
        if (geo[face_idir][i][j][k] < 1e-10)
          magi[i][j][k] = GaussianLineIntegral(FB10, x1[i][j][k], x2[i][j][k], x3[i][j][k], x1[i][j][k + 1],
                                               x2[i][j][k + 1], x3[i][j][k + 1]);
        else
          magi[i][j][k] = (gem[LAj][i][j][k] - gem[LAj][i][j][k + 1] + gem[LAk][i][j + 1][k] -
                           gem[LAk][i][j][k]) / geo[face_idir][i][j][k];
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = isg; i <= ieg; i++)
    for (int j = jsg; j <= jeg + 1; j++)
      for (int k = ksg; k <= keg; k++) {
        if (geo[face_jdir][i][j][k] < 1e-10)
          magj[i][j][k] = GaussianLineIntegral(FB20, x1[i][j][k], x2[i][j][k], x3[i][j][k], x1[i + 1][j][k],
                                               x2[i + 1][j][k], x3[i + 1][j][k]);
        else
          magj[i][j][k] = -(gem[LAi][i][j][k] - gem[LAi][i][j][k + 1] + gem[LAk][i + 1][j][k] -
                            gem[LAk][i][j][k]) / geo[face_jdir][i][j][k];
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = isg; i <= ieg; i++)
    for (int j = jsg; j <= jeg; j++)
      for (int k = ksg; k <= keg + 1; k++) {
        if (geo[face_kdir][i][j][k] < 1e-10)
          magk[i][j][k] = GaussianLineIntegral(FB30, x1[i][j][k], x2[i][j][k], x3[i][j][k], x1[i][j + 1][k],
                                               x2[i][j + 1][k], x3[i][j + 1][k]);
        else
          magk[i][j][k] = (gem[LAi][i][j][k] - gem[LAi][i][j + 1][k] + gem[LAj][i + 1][j][k] -
                           gem[LAj][i][j][k]) / geo[face_kdir][i][j][k];
      }
}

/* ------------------------------------------------------------------------ */
/* background field B0 (src/solver_func.c:1623-2077) */
static const double g_a12[12] = {0.1252334085, 0.3678314989, 0.5873179542, 0.7699026741,
                                 0.9041172563, 0.9815606342, -0.1252334085, -0.3678314989,
                                 -0.5873179542, -0.7699026741, -0.9041172563, -0.9815606342};
static const double g_w12[12] = {0.2491470458, 0.2334925365, 0.2031674267, 0.1600783285,
                                 0.1069393259, 0.0471753363, 0.2491470458, 0.2334925365,
                                 0.2031674267, 0.1600783285, 0.1069393259, 0.0471753363};

/* 12x12 face quadrature of B0 averages (g2int_ortho, src/solver_func.c:1623-1825).
   x*[0..3]: face corners 0, 1, 2, 3 (eta along 0->1, psi along 0->2). */
static void face_quad(int dir, double face_area, const double xa[4], const double ya[4],
                      const double za[4], double out[8]) {
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (fabs(face_area) >= 1e-10) {
    for (int i = 0; i < 12; i++)
      for (int j = 0; j < 12; j++) {
        double eta = 0.5 * (1.0 + g_a12[i]), psi = 0.5 * (1.0 + g_a12[j]);
        double w2 = g_w12[i] * g_w12[j];
        double p1 = xa[0] + psi * (xa[2] - xa[0]) + eta * (xa[1] - xa[0]);
        double p2 = ya[0] + psi * (ya[2] - ya[0]) + eta * (ya[1] - ya[0]);
        double p3 = za[0] + psi * (za[2] - za[0]) + eta * (za[1] - za[0]);
        double d1 = fmax(fmax(fabs(xa[1] - xa[0]), fabs(xa[2] - xa[0])), fabs(xa[3] - xa[0]));
        double d2 = fmax(fmax(fabs(ya[1] - ya[0]), fabs(ya[2] - ya[0])), fabs(ya[3] - ya[0]));
        double d3 = fmax(fmax(fabs(za[1] - za[0]), fabs(za[2] - za[0])), fabs(za[3] - za[0]));
        double b1 = B10(p1, p2, p3), b2 = B20(p1, p2, p3), b3 = B30(p1, p2, p3);
        double h1 = H1(p1, p2, p3), h2 = H2(p1, p2, p3), h3 = H3(p1, p2, p3);
        double bsq = b1 * b1 + b2 * b2 + b3 * b3;
        double bn, bn1, bn2, bn3, wgt, ha, hb;
        if (dir == 0) {
          bn = b1; bn1 = b1 * b1; bn2 = b1 * b2; bn3 = b1 * b3;
          wgt = 0.25 * d2 * d3 * w2; ha = h2; hb = h3;
        } else if (dir == 1) {
          bn = b2; bn1 = b2 * b1; bn2 = b2 * b2; bn3 = b2 * b3;
          wgt = 0.25 * d1 * d3 * w2; ha = h1; hb = h3;
        } else {
          bn = b3; bn1 = b3 * b1; bn2 = b3 * b2; bn3 = b3 * b3;
          wgt = 0.25 * d1 * d2 * w2; ha = h1; hb = h2;
        }
        s[0] += wgt * b1 * ha * hb;
        s[1] += wgt * b2 * ha * hb;
        s[2] += wgt * b3 * ha * hb;
        s[3] += wgt * (dir == 0 ? b1 : bn) * ha * hb;
        s[4] += wgt * bsq * ha * hb;
        s[5] += wgt * bn1 * ha * hb;
        s[6] += wgt * bn2 * ha * hb;
        s[7] += wgt * bn3 * ha * hb;
      }
    for (int q = 0; q < 8; q++) out[q] = s[q] / face_area;
  } else {
    double f1 = GaussianLineIntegral(B10, xa[0], ya[0], za[0], xa[3], ya[3], za[3]);
    double f2 = GaussianLineIntegral(B20, xa[0], ya[0], za[0], xa[3], ya[3], za[3]);
    double f3 = GaussianLineIntegral(B30, xa[0], ya[0], za[0], xa[3], ya[3], za[3]);
    double fn = dir == 0 ? f1 : (dir == 1 ? f2 : f3);
    out[0] = f1; out[1] = f2; out[2] = f3; out[3] = fn;
    out[4] = f1 * f1 + f2 * f2 + f3 * f3;
    out[5] = fn * f1; out[6] = fn * f2; out[7] = fn * f3;
  }
}

static void face_corners(int dir, int i, int j, int k, double xa[4], double ya[4], double za[4]) {
  int di[4], dj[4], dk[4];
  if (dir == 0) { int a[4] = {0, 0, 0, 0}, b[4] = {0, 1, 0, 1}, c[4] = {0, 0, 1, 1}; memcpy(di, a, sizeof a); memcpy(dj, b, sizeof b); memcpy(dk, c, sizeof c); }
  else if (dir == 1) { int a[4] = {0, 0, 1, 1}, b[4] = {0, 0, 0, 0}, c[4] = {0, 1, 0, 1}; memcpy(di, a, sizeof a); memcpy(dj, b, sizeof b); memcpy(dk, c, sizeof c); }
  else { int a[4] = {0, 1, 0, 1}, b[4] = {0, 0, 1, 1}, c[4] = {0, 0, 0, 0}; memcpy(di, a, sizeof a); memcpy(dj, b, sizeof b); memcpy(dk, c, sizeof c); }
  for (int q = 0; q < 4; q++) {
    xa[q] = x1[i + di[q]][j + dj[q]][k + dk[q]];
    ya[q] = x2[i + di[q]][j + dj[q]][k + dk[q]];
    za[q] = x3[i + di[q]][j + dj[q]][k + dk[q]];
  }
}

static void face_b0_at(int dir, int i, int j, int k, double out[8]) {
  double xa[4], ya[4], za[4];
  face_corners(dir, i, j, k, xa, ya, za);
  face_quad(dir, geo[face_idir + dir][i][j][k], xa, ya, za, out);
}

static void edge_b0_at(int dir, int i, int j, int k, double out[3]) {
  int a = dir == 0, b = dir == 1, c = dir == 2;
  Field fs[3] = {B10, B20, B30};
  for (int m = 0; m < 3; m++)
    out[m] = GaussianLineIntegral(*fs[m], x1[i + a][j + b][k + c], x2[i + a][j + b][k + c],
                                  x3[i + a][j + b][k + c], x1[i][j][k], x2[i][j][k], x3[i][j][k]);
}

/* 12^3 volume quadrature of the B0 source functions (g3int_ortho) */
static void vol_src_at(int i, int j, int k, double S[3]) {
  double s1 = 0, s2 = 0, s3 = 0;
  double x1n = x1[i][j][k], x1p = x1[i + 1][j][k];
  double x2n = x2[i][j][k], x2p = x2[i][j + 1][k];
  double x3n = x3[i][j][k], x3p = x3[i][j][k + 1];
  double d1 = x1p - x1n, d2 = x2p - x2n, d3 = x3p - x3n;
  for (int a = 0; a < 12; a++)
    for (int b = 0; b < 12; b++)
      for (int c = 0; c < 12; c++) {
        double p1 = x1n + 0.5 * (1.0 + g_a12[a]) * d1;
        double p2 = x2n + 0.5 * (1.0 + g_a12[b]) * d2;
        double p3 = x3n + 0.5 * (1.0 + g_a12[c]) * d3;
        double w3 = g_w12[a] * g_w12[b] * g_w12[c];
        double h = H1(p1, p2, p3), hh = H2(p1, p2, p3), hhh = H3(p1, p2, p3);
        s1 += 0.125 * d1 * d2 * d3 * w3 * Source1_B0(p1, p2, p3) * h * hh * hhh;
        s2 += 0.125 * d1 * d2 * d3 * w3 * Source2_B0(p1, p2, p3) * h * hh * hhh;
        s3 += 0.125 * d1 * d2 * d3 * w3 * Source3_B0(p1, p2, p3) * h * hh * hhh;
      }
  double V = geo[vol_center][i][j][k];
  S[0] = s1 / V;
  S[1] = s2 / V;
  S[2] = s3 / V;
}

int gk_b0_active(void) {
  return !(B10 == Field_Default && B20 == Field_Default && B30 == Field_Default &&
           Source1_B0 == Field_Default && Source2_B0 == Field_Default && Source3_B0 == Field_Default);
}

/* full background-field arrays over the reference ranges; heavy (the volume
   quadrature is 1728 source evaluations per cell), used for the checker and
   for small grids */
static void b0_full(void) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i <= ieg; i++)
    for (int j = 0; j <= jeg; j++)
      for (int k = 0; k <= keg; k++) {
        gem[B0_b1][i][j][k] = B10(x1ctr[i][j][k], x2ctr[i][j][k], x3ctr[i][j][k]);
        gem[B0_b2][i][j][k] = B20(x1ctr[i][j][k], x2ctr[i][j][k], x3ctr[i][j][k]);
        gem[B0_b3][i][j][k] = B30(x1ctr[i][j][k], x2ctr[i][j][k], x3ctr[i][j][k]);
      }
  for (int dir = 0; dir < 3; dir++) {
    int iu = ieg + (dir == 0), ju = jeg + (dir == 1), ku = keg + (dir == 2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i <= iu; i++)
      for (int j = 0; j <= ju; j++)
        for (int k = 0; k <= ku; k++) {
          double o[8];
          face_b0_at(dir, i, j, k, o);
          for (int q = 0; q < 8; q++) gem[B0AfaceI_b1 + 8 * dir + q][i][j][k] = o[q];
        }
  }
  for (int dir = 0; dir < 3; dir++) {
    int iu = ie + (dir != 0), ju = je + (dir != 1), ku = ke + (dir != 2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = is; i <= iu; i++)
      for (int j = js; j <= ju; j++)
        for (int k = ks; k <= ku; k++) {
          double o[3];
          edge_b0_at(dir, i, j, k, o);
          for (int q = 0; q < 3; q++) gem[B0AedgeI_b1 + 3 * dir + q][i][j][k] = o[q];
        }
  }
  int have_src = !(Source1_B0 == Field_Default && Source2_B0 == Field_Default && Source3_B0 == Field_Default);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i <= ieg; i++)
    for (int j = 0; j <= jeg; j++)
      for (int k = 0; k <= keg; k++) {
        double S[3] = {0.0, 0.0, 0.0};
        if (have_src) vol_src_at(i, j, k, S);
        /* x1-direction */
        double IX1 = 0.5 * gem[B0AfaceI_bsq][i + 1][j][k] - gem[B0AfaceI_bn1][i + 1][j][k];
        double IX0 = 0.5 * gem[B0AfaceI_bsq][i][j][k] - gem[B0AfaceI_bn1][i][j][k];
        double JX1 = -gem[B0AfaceJ_bn1][i][j + 1][k], JX0 = -gem[B0AfaceJ_bn1][i][j][k];
        double KX1 = -gem[B0AfaceK_bn1][i][j][k + 1], KX0 = -gem[B0AfaceK_bn1][i][j][k];
        gem[dmagsf_G_x1dir][i][j][k] =
            (IX1 * geo[face_idir][i + 1][j][k] - IX0 * geo[face_idir][i][j][k] +
             JX1 * geo[face_jdir][i][j + 1][k] - JX0 * geo[face_jdir][i][j][k] +
             KX1 * geo[face_kdir][i][j][k + 1] - KX0 * geo[face_kdir][i][j][k]) /
                geo[vol_center][i][j][k] - S[0];
        double IY1 = -gem[B0AfaceI_bn2][i + 1][j][k], IY0 = -gem[B0AfaceI_bn2][i][j][k];
        double JY1 = 0.5 * gem[B0AfaceJ_bsq][i][j + 1][k] - gem[B0AfaceJ_bn2][i][j + 1][k];
        double JY0 = 0.5 * gem[B0AfaceJ_bsq][i][j][k] - gem[B0AfaceJ_bn2][i][j][k];
        double KY1 = -gem[B0AfaceK_bn2][i][j][k + 1], KY0 = -gem[B0AfaceK_bn2][i][j][k];
        gem[dmagsf_G_x2dir][i][j][k] =
            (IY1 * geo[face_idir][i + 1][j][k] - IY0 * geo[face_idir][i][j][k] +
             JY1 * geo[face_jdir][i][j + 1][k] - JY0 * geo[face_jdir][i][j][k] +
             KY1 * geo[face_kdir][i][j][k + 1] - KY0 * geo[face_kdir][i][j][k]) /
                geo[vol_center][i][j][k] - S[1];
        double IZ1 = -gem[B0AfaceI_bn3][i + 1][j][k], IZ0 = -gem[B0AfaceI_bn3][i][j][k];
        double JZ1 = -gem[B0AfaceJ_bn3][i][j + 1][k], JZ0 = -gem[B0AfaceJ_bn3][i][j][k];
        double KZ1 = 0.5 * gem[B0AfaceK_bsq][i][j][k + 1] - gem[B0AfaceK_bn3][i][j][k + 1];
        double KZ0 = 0.5 * gem[B0AfaceK_bsq][i][j][k] - gem[B0AfaceK_bn3][i][j][k];
        gem[dmagsf_G_x3dir][i][j][k] =
            (IZ1 * geo[face_idir][i + 1][j][k] - IZ0 * geo[face_idir][i][j][k] +
             JZ1 * geo[face_jdir][i][j + 1][k] - JZ0 * geo[face_jdir][i][j][k] +
             KZ1 * geo[face_kdir][i][j][k + 1] - KZ0 * geo[face_kdir][i][j][k]) /
                geo[vol_center][i][j][k] - S[2];
      }
}

/* Device representation of B0.  The reference's arrays (face and edge
   quadratures, cell values, the Lorentz force of the background stress) are
   computed in full whenever a background field is set.  When every array
   holds one value, bit for bit, over the points the device reads (the
   active cells and the faces and edges around them), the device takes those
   values as constants (mode 1); otherwise it takes the arrays (mode 2): a
   field that varies in space, or a uniform one whose quadratures differ in
   the last bits from face to face (grid widths that are not powers of two),
   which the device must see exactly as the reference does. */
gk_b0_info gk_b0;

/* one value over the box, bit for bit (then *val is it) */
static int same_over(int f, int i0, int i1, int j0, int j1, int k0, int k1, double *val) {
  const double v = gem[f][i0][j0][k0];
  for (int i = i0; i <= i1; i++)
    for (int j = j0; j <= j1; j++)
      for (int k = k0; k <= k1; k++)
        if (memcmp(&gem[f][i][j][k], &v, sizeof v) != 0) return 0;
  *val = v;
  return 1;
}

static int b0_constants(void) {
  int ok = 1;
  for (int m = 0; m < 3; m++) {
    ok &= same_over(B0_b1 + m, is, ie, js, je, ks, ke, &gk_b0.cell[m]);
    ok &= same_over(dmagsf_G_x1dir + m, is, ie, js, je, ks, ke, &gk_b0.lorentz[m]);
  }
  for (int d = 0; d < 3; d++) {
    /* faces of direction d: [s, e+1] along d, the active cells across; the
       device reads b1 b2 b3 bn (bsq, bn1..bn3 feed only dmagsf_G) */
    const int fa = d == 0, fb = d == 1, fc = d == 2;
    for (int q = 0; q < 8; q++) {
      double v = gem[B0AfaceI_b1 + 8 * d + q][is][js][ks];
      if (q < 4) ok &= same_over(B0AfaceI_b1 + 8 * d + q, is, ie + fa, js, je + fb, ks, ke + fc, &v);
      gk_b0.face[d][q] = v;
    }
    /* edges of direction d: [s, e] along d, [s, e+1] across; the device
       reads the two components across the edge */
    for (int q = 0; q < 3; q++) {
      double v = gem[B0AedgeI_b1 + 3 * d + q][is][js][ks];
      if (q != d)
        ok &= same_over(B0AedgeI_b1 + 3 * d + q, is, ie + !fa, js, je + !fb, ks, ke + !fc, &v);
      gk_b0.edge[d][q] = v;
    }
  }
  return ok;
}

int gk_b0_full_arrays = 1; /* kept for the bindings; the arrays are always computed */

void set_background_field() {
  log_info("Initializing Background Fields...");
  memset(&gk_b0, 0, sizeof gk_b0);
  if (!gk_b0_active()) return; /* mode 0: no background field */
  b0_full();
  gk_b0.mode = b0_constants() ? 1 : 2;
  if (gk_b0.mode == 2) {
    memset(gk_b0.cell, 0, sizeof gk_b0.cell);
    memset(gk_b0.face, 0, sizeof gk_b0.face);
    memset(gk_b0.edge, 0, sizeof gk_b0.edge);
    memset(gk_b0.lorentz, 0, sizeof gk_b0.lorentz);
    log_info("background field varies over the grid: the device reads it per point");
  }
}

/* ------------------------------------------------------------------------ */
/* boundary-condition library: apply on the host arrays and record the kind */
int gk_bc_kind[2][6]; /* [gas|gem][i_lo,i_hi,j_lo,j_hi,k_lo,k_hi], GK_BCK_* */
int gk_bc_record_only = 0;

static int dim_periodic(int dim) {
  Boundary_condition_t b = dim == 0 ? problem_config.Boundary_i
                                    : (dim == 1 ? problem_config.Boundary_j : problem_config.Boundary_k);
  return b == BC_PERIODIC;
}

static void mirror(int is_gem, int dim, int side, int negate_normal) {
  int N[3] = {config.NI, config.NJ, config.NK};
  int n = N[dim];
  int t1 = dim == 0 ? 1 : 0, t2 = dim == 2 ? 1 : 2; /* transverse axes */
  for (int f = 0; f < 6; f++) {
    double ***a = is_gem ? gem[f] : gas[0][f];
    int o = is_gem ? (dim == 0 ? gem_onface_i[f] : dim == 1 ? gem_onface_j[f] : gem_onface_k[f])
                   : (dim == 0 ? gas_onface_i[f] : dim == 1 ? gas_onface_j[f] : gas_onface_k[f]);
    double sgn = (negate_normal && !is_gem && f == gas_v1 + dim) ? -1.0 : 1.0;
    for (int g = 0; g < NG; g++) {
      int dst = side == 0 ? g : n - 1 - NG + g + o;
      int src = side == 0 ? 2 * NG - 1 - g + o : n - 1 - NG - 1 - g;
      for (int u = 0; u < N[t1]; u++)
        for (int v = 0; v < N[t2]; v++) {
          int d3[3], s3[3];
          d3[dim] = dst;
          s3[dim] = src;
          d3[t1] = s3[t1] = u;
          d3[t2] = s3[t2] = v;
          a[d3[0]][d3[1]][d3[2]] = sgn * a[s3[0]][s3[1]][s3[2]];
        }
    }
  }
}

/* zero-gradient ghosts: every ghost layer copies the first active layer
   (low side) or the last active layer / boundary face (high side) */
static void extrapolate(int is_gem, int dim, int side) {
  int N[3] = {config.NI, config.NJ, config.NK};
  int n = N[dim];
  int t1 = dim == 0 ? 1 : 0, t2 = dim == 2 ? 1 : 2;
  for (int f = 0; f < 6; f++) {
    double ***a = is_gem ? gem[f] : gas[0][f];
    int o = is_gem ? (dim == 0 ? gem_onface_i[f] : dim == 1 ? gem_onface_j[f] : gem_onface_k[f])
                   : (dim == 0 ? gas_onface_i[f] : dim == 1 ? gas_onface_j[f] : gas_onface_k[f]);
    int src = side == 0 ? NG : n - 1 - NG - 1 + o;
    for (int g = 0; g < NG; g++) {
      int dst = side == 0 ? g : n - 1 - NG + g + o;
      for (int u = 0; u < N[t1]; u++)
        for (int v = 0; v < N[t2]; v++) {
          int d3[3], s3[3];
          d3[dim] = dst;
          s3[dim] = src;
          d3[t1] = s3[t1] = u;
          d3[t2] = s3[t2] = v;
          a[d3[0]][d3[1]][d3[2]] = a[s3[0]][s3[1]][s3[2]];
        }
    }
  }
}

/* fixed ghosts: gas rho = 1, v = 0, p = 1 (p_S untouched), gem fields 0 */
static void fixed_fill(int is_gem, int dim, int side) {
  int N[3] = {config.NI, config.NJ, config.NK};
  int n = N[dim];
  int t1 = dim == 0 ? 1 : 0, t2 = dim == 2 ? 1 : 2;
  for (int f = 0; f < (is_gem ? 6 : 5); f++) {
    double ***a = is_gem ? gem[f] : gas[0][f];
    int o = is_gem ? (dim == 0 ? gem_onface_i[f] : dim == 1 ? gem_onface_j[f] : gem_onface_k[f]) : 0;
    double v = is_gem ? 0.0 : (f == gas_rho || f == gas_p ? 1.0 : 0.0);
    for (int g = 0; g < NG; g++) {
      int dst = side == 0 ? g : n - 1 - NG + g + o;
      for (int u = 0; u < N[t1]; u++)
        for (int w = 0; w < N[t2]; w++) {
          int d3[3];
          d3[dim] = dst;
          d3[t1] = u;
          d3[t2] = w;
          a[d3[0]][d3[1]][d3[2]] = v;
        }
    }
  }
}

static void bc_request(int is_gem, int dim, int side, int kind) {
  if (dim_periodic(dim)) return; /* MPI neighbour exists: nothing to do */
  gk_bc_kind[is_gem][2 * dim + side] = kind;
  if (gk_bc_record_only || !gas || !gem) return;
  if (kind == GK_BCK_SYMMETRIC) mirror(is_gem, dim, side, 0);
  if (kind == GK_BCK_REFLECTIVE) mirror(is_gem, dim, side, 1);
  if (kind == GK_BCK_EXTRAPOLATED) extrapolate(is_gem, dim, side);
  if (kind == GK_BCK_FIXED) fixed_fill(is_gem, dim, side);
}

#define GK_BC_DEF(arr, isg_, kind, KIND)                                       \
  void arr##_bc_##kind##_i_low() { bc_request(isg_, 0, 0, KIND); }  \
  void arr##_bc_##kind##_i_high() { bc_request(isg_, 0, 1, KIND); } \
  void arr##_bc_##kind##_j_low() { bc_request(isg_, 1, 0, KIND); }  \
  void arr##_bc_##kind##_j_high() { bc_request(isg_, 1, 1, KIND); } \
  void arr##_bc_##kind##_k_low() { bc_request(isg_, 2, 0, KIND); }  \
  void arr##_bc_##kind##_k_high() { bc_request(isg_, 2, 1, KIND); }

GK_BC_DEF(gas, 0, symmetric, GK_BCK_SYMMETRIC)
GK_BC_DEF(gas, 0, reflective, GK_BCK_REFLECTIVE)
GK_BC_DEF(gas, 0, fixed, GK_BCK_FIXED)
GK_BC_DEF(gas, 0, extrapolated, GK_BCK_EXTRAPOLATED)
GK_BC_DEF(gas, 0, none, GK_BCK_NONE)
GK_BC_DEF(gem, 1, symmetric, GK_BCK_SYMMETRIC)
/* the reference's gem reflective routines are empty (src/boundary_conditions.c:595-600) */
GK_BC_DEF(gem, 1, reflective, GK_BCK_NONE)
GK_BC_DEF(gem, 1, fixed, GK_BCK_FIXED)
GK_BC_DEF(gem, 1, extrapolated, GK_BCK_EXTRAPOLATED)
GK_BC_DEF(gem, 1, none, GK_BCK_NONE)

/* Pole boundary on a j side (src/spherical/spherical_boundary.c:23-138):
   ghost rows from the row mirrored across the pole, rotated by pi in phi, for
   every i and k of the box; v2, v3 and B_j, B_k, b2, b3 change sign. */
static void pole_mirror(int is_gem, int side) {
  const int NJ = config.NJ, nk = config.nk_global;
  for (int f = 0; f < 6; f++) {
    double ***a = is_gem ? gem[f] : gas[0][f];
    int o = is_gem ? gem_onface_j[f] : gas_onface_j[f];
    int neg = is_gem ? !(f == mag_bi || f == mags_b1) : (f == gas_v2 || f == gas_v3);
    for (int i = 0; i < config.NI; i++)
      for (int g = 0; g < NG; g++) {
        int dst = side == 0 ? g : NJ - 1 - NG + g + o;
        int src = side == 0 ? 2 * NG - 1 - g + o : NJ - 1 - NG - 1 - g;
        for (int k = 0; k < config.NK; k++) {
          double v = a[i][src][NG + (k - NG + nk / 2) % nk];
          a[i][dst][k] = neg ? -v : v;
        }
      }
  }
}

static void bc_pole_request(int is_gem, int side) {
  if (dim_periodic(1)) return;
  gk_bc_kind[is_gem][2 + side] = GK_BCK_POLE;
  if (gk_bc_record_only || !gas || !gem) return;
  pole_mirror(is_gem, side);
}

void gas_bc_pole_j_low() { bc_pole_request(0, 0); }
void gas_bc_pole_j_high() { bc_pole_request(0, 1); }
void gem_bc_pole_j_low() { bc_pole_request(1, 0); }
void gem_bc_pole_j_high() { bc_pole_request(1, 1); }
