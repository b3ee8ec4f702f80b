/*
 * Per-axis device tables from the host geometry.
 *
 * On orthogonal grids every metric quantity the step uses is a product of
 * one-dimensional factors, f(i) g(j) h(k), and the reconstruction weights and
 * limiter coefficients depend on the face index along the sweep axis only.
 * The device therefore reads O(N) tables instead of the reference's 70
 * full-size geometry arrays (src/config.h:161-291); per-cell HBM traffic is
 * state only.  Factors are taken from the host arrays along the first active
 * line of each axis, so they equal the reference values to roundoff.
 *
 * Cartesian reconstruction weights vary with the transverse cell widths at
 * roundoff level; tables are kept per distinct width (see below) so the
 * device reproduces the reference arrays exactly.
 *
 * PDM limiter coefficients (src/solver_func.c:236-246): with e = 1/(2 PDMB+1)
 *   c1 = (q - e A_i) / (e A_i),  c2 = (e A_im1 - q) / (e A_i),  q = V_g / dx
 * evaluated per face for the left state (upwind cell f-1, neighbour face f-1)
 * and the right state (upwind cell f, neighbour face f+1).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "log.h"

#include "config.h"
#include "curvilinear.h"
#include "gk_host.h"

static double *zalloc(int n) { return (double *)calloc((size_t)n, sizeof(double)); }

int gk_tables_build(gk_tables *t) {
  memset(t, 0, sizeof *t);
  const int N[3] = {config.NI, config.NJ, config.NK};
  const int S[3] = {is, js, ks}, E[3] = {ie, je, ke};
  double *xf[3], *dx[3], *xc[3], *xctr[3];
  for (int a = 0; a < 3; a++) {
    xf[a] = zalloc(N[a]);
    dx[a] = zalloc(N[a]);
    xc[a] = zalloc(N[a]);
    xctr[a] = zalloc(N[a]);
  }
  for (int n = 0; n < N[0]; n++) xf[0][n] = x1[n][js][ks];
  for (int n = 0; n < N[1]; n++) xf[1][n] = x2[is][n][ks];
  for (int n = 0; n < N[2]; n++) xf[2][n] = x3[is][js][n];
  for (int a = 0; a < 3; a++)
    for (int n = 0; n + 1 < N[a]; n++) {
      dx[a][n] = xf[a][n + 1] - xf[a][n];
      xc[a][n] = 0.5 * (xf[a][n] + xf[a][n + 1]);
    }
  for (int n = 0; n + 1 < N[0]; n++) xctr[0][n] = x1ctr[n][js][ks];
  for (int n = 0; n + 1 < N[1]; n++) xctr[1][n] = x2ctr[is][n][ks];
  for (int n = 0; n + 1 < N[2]; n++) xctr[2][n] = x3ctr[is][js][n];

  for (int q = 0; q < GK_G_COUNT; q++)
    for (int a = 0; a < 3; a++) {
      t->geom[q][a] = zalloc(N[a]);
      for (int n = 0; n < N[a]; n++) t->geom[q][a][n] = 1.0;
    }
  const int sph = gk_coord_type == 2;
  for (int n = 0; n < N[0]; n++) {
    double r = xf[0][n], r1 = n + 1 < N[0] ? xf[0][n + 1] : r;
    double d1 = dx[0][n];
    t->geom[GK_G_EDGE_I][0][n] = d1;
    t->geom[GK_G_FACE_J][0][n] = sph ? 0.5 * (r1 * r1 - r * r) : d1;
    t->geom[GK_G_FACE_K][0][n] = sph ? 0.5 * (r1 * r1 - r * r) : d1;
    t->geom[GK_G_VOL][0][n] = sph ? 1.0 / 3.0 * (pow(r1, 3) - pow(r, 3)) : d1;
    t->geom[GK_G_DXI][0][n] = d1;
    if (sph) {
      t->geom[GK_G_EDGE_J][0][n] = r;
      t->geom[GK_G_EDGE_K][0][n] = r;
      t->geom[GK_G_FACE_I][0][n] = r * r;
      t->geom[GK_G_DXJ][0][n] = xc[0][n];
      t->geom[GK_G_DXK][0][n] = xc[0][n];
    }
  }
  for (int n = 0; n < N[1]; n++) {
    double th = xf[1][n], th1 = n + 1 < N[1] ? xf[1][n + 1] : th;
    double d2 = dx[1][n];
    t->geom[GK_G_EDGE_J][1][n] = d2;
    t->geom[GK_G_FACE_I][1][n] = sph ? -(cos(th1) - cos(th)) : d2;
    t->geom[GK_G_FACE_K][1][n] = d2;
    t->geom[GK_G_VOL][1][n] = sph ? -(cos(th1) - cos(th)) : d2;
    t->geom[GK_G_DXJ][1][n] = d2;
    if (sph) {
      t->geom[GK_G_EDGE_K][1][n] = sin(th);
      t->geom[GK_G_FACE_J][1][n] = sin(th);
      t->geom[GK_G_DXK][1][n] = sin(xc[1][n]);
    }
  }
  for (int n = 0; n < N[2]; n++) {
    double d3 = dx[2][n];
    t->geom[GK_G_EDGE_K][2][n] = d3;
    t->geom[GK_G_FACE_I][2][n] = d3;
    t->geom[GK_G_FACE_J][2][n] = d3;
    t->geom[GK_G_VOL][2][n] = d3;
    t->geom[GK_G_DXK][2][n] = d3;
  }

  /* cell-centre B from the two faces (src/solver_func.c:310-322); the
     device evaluates (B[n]*d0)/dx + (B[n+1]*d1)/dx exactly as the reference */
  for (int a = 0; a < 3; a++) {
    for (int q = 0; q < 3; q++) t->cc[a][q] = zalloc(N[a]);
    for (int n = 0; n + 1 < N[a]; n++) {
      t->cc[a][0][n] = xf[a][n + 1] - xctr[a][n];
      t->cc[a][1][n] = xctr[a][n] - xf[a][n];
      t->cc[a][2][n] = dx[a][n];
    }
  }

  /* Reconstruction weights and limiter coefficients per face.  On Cartesian
     grids the reference's weight arrays depend on the transverse cell widths
     through roundoff (w = c V / sum(fac8th V) with V = dx1 dx2 dx3), so each
     distinct transverse width gets its own table row; on spherical grids they
     are exactly one-dimensional. */
  const double e = 1.0 / (2 * PDMB + 1.0);
  double ***dxa[3] = {dx1, dx2, dx3};
  t->exact = 1;
  for (int a = 0; a < 3; a++) {
    const int tax[2] = {a == 0 ? 1 : 0, a == 2 ? 1 : 2};
    int rep[2][64];
    for (int q = 0; q < 2; q++) {
      int b = tax[q];
      t->cls[a][q] = (int *)calloc((size_t)N[b], sizeof(int));
      t->ncls[a][q] = 1;
      rep[q][0] = S[b];
      if (sph) continue;
      double vals[64];
      int nv = 0, overflow = 0;
      for (int n = 0; n < N[b]; n++) {
        int c = -1;
        double v = n + 1 < N[b] ? dx[b][n] : dx[b][n - 1];
        for (int u = 0; u < nv; u++)
          if (vals[u] == v) c = u;
        if (c < 0) {
          if (nv == 64) {
            overflow = 1;
            break;
          }
          vals[nv] = v;
          rep[q][nv] = n + 1 < N[b] ? n : n - 1;
          c = nv++;
        }
        t->cls[a][q][n] = c;
      }
      if (overflow) { /* strongly non-uniform: one representative row */
        memset(t->cls[a][q], 0, sizeof(int) * (size_t)N[b]);
        t->ncls[a][q] = 1;
        rep[q][0] = S[b];
        t->exact = 0;
      } else {
        t->ncls[a][q] = nv;
      }
    }
    const int C0 = t->ncls[a][0], C1 = t->ncls[a][1];
    t->recw[a] = zalloc(N[a] * C0 * C1 * GK_RECW_STRIDE);
    for (int f = S[a]; f <= E[a] + 1; f++)
      for (int c0 = 0; c0 < C0; c0++)
        for (int c1 = 0; c1 < C1; c1++) {
          int p[3];
          p[tax[0]] = rep[0][c0];
          p[tax[1]] = rep[1][c1];
          double *T = t->recw[a] + (((size_t)f * C0 + c0) * C1 + c1) * GK_RECW_STRIDE;
#define AT(arr, off) (p[a] = f + (off), (arr)[p[0]][p[1]][p[2]])
          for (int n = 0; n < 7; n++) {
            T[n] = AT(rec[w1l_first_direction + 7 * a + n], 0);
            T[7 + n] = AT(rec[w1r_first_direction + 7 * a + (6 - n)], 0);
          }
          double Af = AT(geo[GF_face_idir + a], 0);
          double Am = AT(geo[GF_face_idir + a], -1);
          double Ap = AT(geo[GF_face_idir + a], 1);
          double qL = AT(geo[GF_vol_idir + a], -1) / AT(dxa[a], -1);
          double qR = AT(geo[GF_vol_idir + a], 0) / AT(dxa[a], 0);
#undef AT
          T[14] = (qL - e * Af) / (e * Af);
          T[15] = (e * Am - qL) / (e * Af);
          T[16] = (qR - e * Af) / (e * Af);
          T[17] = (e * Ap - qR) / (e * Af);
        }
  }

  for (int a = 0; a < 3; a++) {
    free(xf[a]);
    free(dx[a]);
    free(xc[a]);
    free(xctr[a]);
  }
  return 0;
}

void gk_tables_free(gk_tables *t) {
  for (int q = 0; q < GK_G_COUNT; q++)
    for (int a = 0; a < 3; a++) free(t->geom[q][a]);
  for (int a = 0; a < 3; a++) {
    for (int q = 0; q < 3; q++) free(t->cc[a][q]);
    free(t->recw[a]);
    free(t->cls[a][0]);
    free(t->cls[a][1]);
  }
  memset(t, 0, sizeof *t);
}

static int dev_bc(int dim, int side, int arr) {
  Boundary_condition_t b = dim == 0 ? problem_config.Boundary_i
                                    : (dim == 1 ? problem_config.Boundary_j : problem_config.Boundary_k);
  if (b == BC_PERIODIC) return GK_BC_PERIODIC;
  switch (gk_bc_kind[arr][2 * dim + side]) {
    case GK_BCK_SYMMETRIC: return GK_BC_SYMMETRIC;
    case GK_BCK_POLE: return dim == 1 ? GK_BC_POLE : -1;
    case GK_BCK_REFLECTIVE: return arr == 0 ? GK_BC_REFLECTIVE : -1;
    case GK_BCK_EXTRAPOLATED: return GK_BC_EXTRAPOLATED;
    case GK_BCK_FIXED: return GK_BC_FIXED;
    case GK_BCK_UNSET:
    case GK_BCK_NONE: return GK_BC_NONE;
    default: return -1;
  }
}

int gk_fill_desc(gk_grid_desc *d, const gk_tables *t, int precision) {
  memset(d, 0, sizeof *d);
  d->ni = config.ni;
  d->nj = config.nj;
  d->nk = config.nk;
  d->coord = gk_coord_type;
  d->precision = precision;
  d->CA = CA;
  for (int arr = 0; arr < 2; arr++)
    for (int dim = 0; dim < 3; dim++)
      for (int side = 0; side < 2; side++) {
        int k = dev_bc(dim, side, arr);
        if (k < 0) return GK_ERR_ARG;
        d->bc[arr][2 * dim + side] = k;
      }
  for (int q = 0; q < GK_G_COUNT; q++)
    for (int a = 0; a < 3; a++) d->geom[q][a] = t->geom[q][a];
  for (int a = 0; a < 3; a++) {
    for (int q = 0; q < 3; q++) d->cell_coef[a][q] = t->cc[a][q];
    d->recw[a] = t->recw[a];
    for (int q = 0; q < 2; q++) {
      d->cls[a][q] = t->cls[a][q];
      d->ncls[a][q] = t->ncls[a][q];
    }
  }
  if (gk_b0.mode < 0) return GK_ERR_ARG;
  d->b0_mode = gk_b0.mode;
  memcpy(d->b0_cell, gk_b0.cell, sizeof d->b0_cell);
  memcpy(d->b0_face, gk_b0.face, sizeof d->b0_face);
  memcpy(d->b0_edge, gk_b0.edge, sizeof d->b0_edge);
  memcpy(d->b0_lorentz, gk_b0.lorentz, sizeof d->b0_lorentz);
  if (gk_b0.mode == 2) { /* the reference's arrays, in the order of gamera_op.h */
    int q = 0;
    for (int f = B0AfaceI_b1; f <= B0AfaceK_bn3; f++)
      if ((f - B0AfaceI_b1) % 8 < 4) d->b0_field[q++] = &gem[f][0][0][0];
    const int edge[6] = {B0AedgeI_b2, B0AedgeI_b3, B0AedgeJ_b3, B0AedgeJ_b1, B0AedgeK_b1, B0AedgeK_b2};
    for (int e = 0; e < 6; e++) d->b0_field[q++] = &gem[edge[e]][0][0][0];
    for (int m = 0; m < 3; m++) d->b0_field[q++] = &gem[B0_b1 + m][0][0][0];
    for (int m = 0; m < 3; m++) d->b0_field[q++] = &gem[dmagsf_G_x1dir + m][0][0][0];
  }
  /* the problem's floors and gravity (globals a plugin may set) */
  d->rho_floor = rho_floor;
  d->p_floor = p_floor;
  d->gravity[0] = G01;
  d->gravity[1] = G02;
  d->gravity[2] = G03;
  if (!(rho_floor > 0.0) || !(p_floor > 0.0)) {
    log_error("device solver: floors must be positive (rho_floor %g, p_floor %g)", rho_floor,
              p_floor);
    return GK_ERR_ARG;
  }
  if (doRingAverage == 1) {
    d->ring = 1;
    d->ring_rows = gk_ring_table(config.nk_global, d->ring_nchunk);
    if (d->ring_rows < 0) return GK_ERR_ARG;
    d->ring_pole[0] = fabs(x2min_global) < 1e-10;
    d->ring_pole[1] = fabs(x2max_global - PI) < 1e-10;
  }
  return GK_OK;
}

int gk_ring_table(int nk, int nchunk[16]) {
  static const int t[16] = {8, 16, 32, 32, 64, 64, 64, 64, 128, 128, 128, 128, 128, 128, 128, 128};
  int n = nk == 256 ? 16 : nk == 128 ? 8 : nk == 64 ? 4 : nk == 32 ? 2 : -1;
  for (int r = 0; r < 16; r++) nchunk[r] = r < n ? t[r] : 0;
  return n;
}

/* Ring-average set-up of the reference main() (src/main.c:66-70).  The face
   areas of the ring rows are the ordinary ones (the reference copies them,
   src/spherical/spherical_RingAverage.c:32-51), so only the CFL width
   geo[dxk_ring] and the pole flags are set here (:54-147). */
void set_RingAverage_geo_data() {}

void set_RingAverage_config() {
  int nch[16];
  NAVR = gk_ring_table(config.nk_global, nch);
  if (NAVR < 0) {
    log_error("config.nk_global must be 32, 64, 128, or 256");
    return;
  }
  for (int r = 0; r < NAVR; r++) Nchunk[r] = nch[r];
  for (int i = isg; i <= ieg; i++)
    for (int j = jsg; j <= jeg; j++)
      for (int k = ksg; k <= keg; k++) geo[dxk_ring][i][j][k] = geo[dxk][i][j][k];
  doRingAverage_thisrank_positivepole = fabs(x2min_global) <= 1e-10;
  doRingAverage_thisrank_negativepole = fabs(x2max_global - PI) <= 1e-10;
  for (int p = 0; p < 2; p++) {
    if (!(p == 0 ? doRingAverage_thisrank_positivepole : doRingAverage_thisrank_negativepole)) continue;
    for (int r = 0; r < NAVR && r < config.nj; r++) {
      int Nm = config.nk_global / Nchunk[r];
      int j = p == 0 ? js + r : je - r;
      for (int i = is; i <= ie; i++)
        for (int k = ks; k <= ke; k++) geo[dxk_ring][i][j][k] = Nm * geo[dxk][i][j][k];
    }
  }
}
