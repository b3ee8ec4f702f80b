/*
 * CPU oracle for the GAMERA-OP per-step MHD update -- TEST INFRASTRUCTURE.
 *
 * This file is the checker, not the product: only tests/, the smoke() entry
 * and bench.py's cpu_baseline leg load it.  It restates the reference time
 * step with the reference's loop ranges and floating-point operation order so
 * that, built strict IEEE (-O2 -ffp-contract=off), it reproduces the reference
 * CPU solver bit for bit.  Each function cites the reference lines it follows.
 * MPI is replaced by in-process ghost fills that reproduce the index maps of
 * the reference's MPI datatypes (src/setup_mpi.c:62-253) for a single rank.
 */
#include "gamera_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NG 4

/* field indices: the reference layout contract (src/config.h:161-499) */
enum {
  edge_idir = 0, edge_jdir, edge_kdir, dxi, dxj, dxk, dxi_ring, dxj_ring,
  dxk_ring, face_idir, face_jdir, face_kdir, vol_idir, vol_jdir, vol_kdir,
  vol_center, GF_face_idir, GF_face_jdir, GF_face_kdir, GF_vol_idir,
  GF_vol_jdir, GF_vol_kdir, IfaceAedgeJ, IfaceAedgeK, JfaceAedgeK,
  JfaceAedgeI, KfaceAedgeI, KfaceAedgeJ, NF_geo
};
enum { NF_rec = 42 };
enum {
  mag_bi = 0, mag_bj, mag_bk, mags_b1, mags_b2, mags_b3, mag_bi_h, mag_bj_h,
  mag_bk_h, mag_bi_p, mag_bj_p, mag_bk_p, mags_b1_h, mags_b2_h, mags_b3_h,
  mags_b1_p, mags_b2_p, mags_b3_p, bi_avg, bj_avg, bk_avg, bil, bir, bjl, bjr,
  bkl, bkr, dvzz, alf_ratio, perp_ratio, dv_alf, mag0_bi, mag0_bj, mag0_bk,
  magsl_b1, magsl_b2, magsl_b3, magsr_b1, magsr_b2, magsr_b3, magsf_x1dir,
  magsf_x2dir, magsf_x3dir, dmagsf_x1dir, dmagsf_x2dir, dmagsf_x3dir,
  efield_ei, efield_ej, efield_ek, LAi, LAj, LAk, B0_b1, B0_b2, B0_b3,
  B0AfaceI_b1, B0AfaceI_b2, B0AfaceI_b3, B0AfaceI_bn, B0AfaceI_bsq,
  B0AfaceI_bn1, B0AfaceI_bn2, B0AfaceI_bn3, B0AfaceJ_b1, B0AfaceJ_b2,
  B0AfaceJ_b3, B0AfaceJ_bn, B0AfaceJ_bsq, B0AfaceJ_bn1, B0AfaceJ_bn2,
  B0AfaceJ_bn3, B0AfaceK_b1, B0AfaceK_b2, B0AfaceK_b3, B0AfaceK_bn,
  B0AfaceK_bsq, B0AfaceK_bn1, B0AfaceK_bn2, B0AfaceK_bn3, B0AedgeI_b1,
  B0AedgeI_b2, B0AedgeI_b3, B0AedgeJ_b1, B0AedgeJ_b2, B0AedgeJ_b3,
  B0AedgeK_b1, B0AedgeK_b2, B0AedgeK_b3, dmagsf_G_x1dir, dmagsf_G_x2dir,
  dmagsf_G_x3dir, S1_B, S2_B, S3_B, magtot_b1, magtot_b2, magtot_b3, NF_gem
};
enum {
  gas_rho = 0, gas_v1, gas_v2, gas_v3, gas_p, gas_p_S, gas_rho_h, gas_v1_h,
  gas_v2_h, gas_v3_h, gas_p_h, gas_p_S_h, gas_rho_p, gas_v1_p, gas_v2_p,
  gas_v3_p, gas_p_p, gas_p_S_p, v1_interp_l, v2_interp_l, v3_interp_l,
  v1_interp_r, v2_interp_r, v3_interp_r, v1_interp, v2_interp, v3_interp,
  v1_avg, v2_avg, v3_avg, gasl_rho, gasl_v1, gasl_v2, gasl_v3, gasl_p,
  gasl_p_S, gasr_rho, gasr_v1, gasr_v2, gasr_v3, gasr_p, gasr_p_S, gasc_rho,
  gasc_rhov1, gasc_rhov2, gasc_rhov3, gasc_eng, gasc_entropy, gascf_rho,
  gascf_rhov1, gascf_rhov2, gascf_rhov3, gascf_eng, gascf_entropy, gasc0_rho,
  gasc0_rhov1, gasc0_rhov2, gasc0_rhov3, gasc0_eng, gasc0_entropy, dgascf_rho,
  dgascf_rhov1, dgascf_rhov2, dgascf_rhov3, dgascf_eng, dgascf_entropy,
  v1_tmp, v2_tmp, v3_tmp, v1_0, v2_0, v3_0, rhov1_fluid, rhov2_fluid,
  rhov3_fluid, rhov1_h, rhov2_h, rhov3_h, S1_HD, S2_HD, S3_HD, bdotv, NF_gas
};
/* rec weights: w{1..7}{l,r}_{first,second,third}_direction
   (src/config.h:235-291): left block dir*7 + n, right block 21 + dir*7 + n */
#define W_L(dir, n) ((dir) * 7 + (n))
#define W_R(dir, n) (21 + (dir) * 7 + (n))

/* physical constants (src/config.c:82-106) */
static const double gamma_val = 5.0 / 3.0;
static const double CFL = 0.3;
static const double PDMB = 1.0;
/* floors and gravity are mutable globals a problem plugin may set
   (src/config.c:89-97); these are their initial values */
static const double rho_floor_default = 0.00316227;
static const double p_floor_default = 0.000316227;

struct go_solver {
  int ni, nj, nk, NI, NJ, NK;
  int is, ie, isg, ieg, js, je, jsg, jeg, ks, ke, ksg, keg;
  int coord;
  int bc[6]; /* i_low, i_high, j_low, j_high, k_low, k_high */
  double CA;
  int64_t n3;
  double *gas, *gem, *geo, *rec;
  double *crd[12]; /* x1 x2 x3 x1c x2c x3c x1ctr x2ctr x3ctr dx1 dx2 dx3 */
  double dt, dt0, time_sim, divB_max;
  double *dbgF; /* optional per-direction copies of the 9 face fluxes */
  double rho_floor, p_floor, G0[3];
  int ring, pole[2], navr, nchunk[16]; /* pole ring averaging */
};

#define IDX(i, j, k) ((((int64_t)(i)) * s->NJ + (j)) * s->NK + (k))
#define GAS(f, i, j, k) s->gas[(int64_t)(f) * s->n3 + IDX(i, j, k)]
#define GEM(f, i, j, k) s->gem[(int64_t)(f) * s->n3 + IDX(i, j, k)]
#define GEO(f, i, j, k) s->geo[(int64_t)(f) * s->n3 + IDX(i, j, k)]
#define REC(f, i, j, k) s->rec[(int64_t)(f) * s->n3 + IDX(i, j, k)]
#define X1(i, j, k) s->crd[0][IDX(i, j, k)]
#define X2(i, j, k) s->crd[1][IDX(i, j, k)]
#define X3(i, j, k) s->crd[2][IDX(i, j, k)]
#define X1C(i, j, k) s->crd[3][IDX(i, j, k)]
#define X1CTR(i, j, k) s->crd[6][IDX(i, j, k)]
#define X2CTR(i, j, k) s->crd[7][IDX(i, j, k)]
#define X3CTR(i, j, k) s->crd[8][IDX(i, j, k)]
#define DX1(i, j, k) s->crd[9][IDX(i, j, k)]
#define DX2(i, j, k) s->crd[10][IDX(i, j, k)]
#define DX3(i, j, k) s->crd[11][IDX(i, j, k)]

go_solver *go_create(int ni, int nj, int nk, int coord, const int bc[6],
                     double CA) {
  go_solver *s = (go_solver *)calloc(1, sizeof(go_solver));
  if (!s) return NULL;
  s->ni = ni; s->nj = nj; s->nk = nk;
  s->NI = ni + 2 * NG + 1; s->NJ = nj + 2 * NG + 1; s->NK = nk + 2 * NG + 1;
  /* grid indices (src/config.c:267-280) */
  s->isg = 0; s->is = NG; s->ie = s->is + ni - 1; s->ieg = s->ie + NG;
  s->jsg = 0; s->js = NG; s->je = s->js + nj - 1; s->jeg = s->je + NG;
  s->ksg = 0; s->ks = NG; s->ke = s->ks + nk - 1; s->keg = s->ke + NG;
  s->coord = coord;
  for (int n = 0; n < 6; n++) s->bc[n] = bc[n];
  s->CA = CA;
  s->rho_floor = rho_floor_default;
  s->p_floor = p_floor_default;
  s->n3 = (int64_t)s->NI * s->NJ * s->NK;
  s->gas = (double *)calloc((size_t)(NF_gas * s->n3), sizeof(double));
  s->gem = (double *)calloc((size_t)(NF_gem * s->n3), sizeof(double));
  s->geo = (double *)calloc((size_t)(NF_geo * s->n3), sizeof(double));
  s->rec = (double *)calloc((size_t)(NF_rec * s->n3), sizeof(double));
  int ok = s->gas && s->gem && s->geo && s->rec;
  for (int n = 0; n < 12; n++) {
    s->crd[n] = (double *)calloc((size_t)s->n3, sizeof(double));
    ok = ok && s->crd[n];
  }
  if (!ok) {
    go_destroy(s);
    return NULL;
  }
  return s;
}

void go_destroy(go_solver *s) {
  if (!s) return;
  free(s->gas); free(s->gem); free(s->geo); free(s->rec); free(s->dbgF);
  for (int n = 0; n < 12; n++) free(s->crd[n]);
  free(s);
}

double *go_array(go_solver *s, int which) {
  switch (which) {
    case GO_GAS: return s->gas;
    case GO_GEM: return s->gem;
    case GO_GEO: return s->geo;
    case GO_REC: return s->rec;
    default:
      if (which >= GO_X1 && which <= GO_DX3) return s->crd[which - GO_X1];
      return NULL;
  }
}

int64_t go_npoints(const go_solver *s) { return s->n3; }

double go_scalar(const go_solver *s, int which) {
  switch (which) {
    case 0: return s->dt;
    case 1: return s->dt0;
    case 2: return s->time_sim;
    case 4: return s->rho_floor;
    case 5: return s->p_floor;
    case 6: case 7: case 8: return s->G0[which - 6];
    default: return s->divB_max;
  }
}

void go_set_scalar(go_solver *s, int which, double v) {
  switch (which) {
    case 0: s->dt = v; break;
    case 1: s->dt0 = v; break;
    case 2: s->time_sim = v; break;
    case 4: s->rho_floor = v; break;
    case 5: s->p_floor = v; break;
    case 6: case 7: case 8: s->G0[which - 6] = v; break;
    default: s->divB_max = v; break;
  }
}

/* ---------------------------------------------------------------------- */
/* Ghost fill: periodic exchange with self-neighbours
   (src/boundary_exchange.c:7-165, datatypes src/setup_mpi.c:62-253).
   dim 0/1/2 = i/j/k; o = onface flag of the field along dim. */
static void exchange_dim(go_solver *s, double *a, int dim, int o) {
  int N[3] = {s->NI, s->NJ, s->NK};
  int n = N[dim];
  /* high ghost <- low active, then low ghost <- high active */
  for (int pass = 0; pass < 2; pass++) {
    int src0 = pass == 0 ? NG + o : n - 2 * NG - 1;
    int dst0 = pass == 0 ? n - NG - 1 + o : 0;
    for (int g = 0; g < NG; g++) {
      int src = src0 + g, dst = dst0 + g;
      if (dim == 0) {
        memcpy(&a[IDX(dst, 0, 0)], &a[IDX(src, 0, 0)],
               sizeof(double) * (size_t)s->NJ * s->NK);
      } else if (dim == 1) {
        for (int i = 0; i < s->NI; i++)
          memcpy(&a[IDX(i, dst, 0)], &a[IDX(i, src, 0)],
                 sizeof(double) * (size_t)s->NK);
      } else {
        for (int i = 0; i < s->NI; i++)
          for (int j = 0; j < s->NJ; j++) a[IDX(i, j, dst)] = a[IDX(i, j, src)];
      }
    }
  }
}

/* boundary_exchange_4d(gem) + boundary_exchange_5d(gas): fields 0..5, i then
   j then k; gem face-normal B carries onface=1 along its own axis
   (src/config.c:250-260). */
static void boundary_exchange(go_solver *s) {
  for (int dim = 0; dim < 3; dim++) {
    if (s->bc[2 * dim] != GO_BC_PERIODIC) continue; /* MPI_PROC_NULL */
    for (int f = 0; f < 6; f++)
      exchange_dim(s, &s->gem[(int64_t)f * s->n3], dim, f == dim ? 1 : 0);
  }
  for (int dim = 0; dim < 3; dim++) {
    if (s->bc[2 * dim] != GO_BC_PERIODIC) continue;
    for (int f = 0; f < 6; f++)
      exchange_dim(s, &s->gas[(int64_t)f * s->n3], dim, 0);
  }
}

/* symmetric mirror for one (dim, side) of fields 0..5
   (src/boundary_conditions.c:9-112 gas, :505-593 gem); every row of the
   other two axes is mirrored, ghosts included.  neg: the reflective variant
   (gas, :114-245), the velocity normal to the side negated. */
static void bc_mirror(go_solver *s, double *base, int dim, int side, int is_gem, int neg) {
  int N[3] = {s->NI, s->NJ, s->NK};
  int n = N[dim];
  for (int f = 0; f < 6; f++) {
    double *a = &base[(int64_t)f * s->n3];
    int o = (is_gem && f == dim) ? 1 : 0;
    for (int g = 0; g < NG; g++) {
      int dst = side == 0 ? g : n - 1 - NG + g + o;
      int src = side == 0 ? 2 * NG - 1 - g + o : n - 1 - NG - 1 - g;
      if (neg && f == gas_v1 + dim) {
        for (int i = 0; i < s->NI; i++)
          for (int j = 0; j < s->NJ; j++)
            for (int k = 0; k < s->NK; k++) {
              int d3[3] = {i, j, k}, s3[3] = {i, j, k};
              if (d3[dim] != dst) continue;
              s3[dim] = src;
              a[IDX(i, j, k)] = -a[IDX(s3[0], s3[1], s3[2])];
            }
        continue;
      }
      if (dim == 0) {
        memcpy(&a[IDX(dst, 0, 0)], &a[IDX(src, 0, 0)],
               sizeof(double) * (size_t)s->NJ * s->NK);
      } else if (dim == 1) {
        for (int i = 0; i < s->NI; i++)
          memcpy(&a[IDX(i, dst, 0)], &a[IDX(i, src, 0)],
                 sizeof(double) * (size_t)s->NK);
      } else {
        for (int i = 0; i < s->NI; i++)
          for (int j = 0; j < s->NJ; j++) a[IDX(i, j, dst)] = a[IDX(i, j, src)];
      }
    }
  }
}

/* pole boundary for one j side of fields 0..5 (src/spherical/
   spherical_boundary.c:23-138): the ghost rows take the row mirrored across
   the pole, rotated by pi in phi (k -> NG + (k - NG + nk/2) mod nk), for every
   i and k of the box; v2, v3 (gas) and B_j, B_k, b2, b3 (gem) change sign. */
static void bc_pole(go_solver *s, double *base, int side, int is_gem) {
  for (int f = 0; f < 6; f++) {
    double *a = &base[(int64_t)f * s->n3];
    int o = (is_gem && f == 1) ? 1 : 0;
    double sgn = is_gem ? ((f == 0 || f == 3) ? 1.0 : -1.0)
                        : ((f == 2 || f == 3) ? -1.0 : 1.0);
    for (int i = 0; i < s->NI; i++)
      for (int g = 0; g < NG; g++) {
        int dst = side == 0 ? g : s->NJ - 1 - NG + g + o;
        int src = side == 0 ? 2 * NG - 1 - g + o : s->NJ - 1 - NG - 1 - g;
        for (int k = 0; k < s->NK; k++) {
          int kr = NG + (k - NG + s->nk / 2) % s->nk;
          a[IDX(i, dst, k)] = sgn * a[IDX(i, src, kr)];
        }
      }
  }
}

/* extrapolated (zero-gradient) ghosts for one (dim, side) of fields 0..5
   (src/boundary_conditions.c:404-496 gas, :687-770 gem): every ghost layer
   copies the first active layer (low) or the last active layer or boundary
   face (high), over every row of the other two axes. */
static void bc_extrapolated(go_solver *s, double *base, int dim, int side, int is_gem) {
  int N[3] = {s->NI, s->NJ, s->NK};
  int n = N[dim];
  for (int f = 0; f < 6; f++) {
    double *a = &base[(int64_t)f * s->n3];
    int o = (is_gem && f == dim) ? 1 : 0;
    int src = side == 0 ? NG : n - 1 - NG - 1 + o;
    for (int g = 0; g < NG; g++) {
      int dst = side == 0 ? g : n - 1 - NG + g + o;
      for (int i = 0; i < s->NI; i++)
        for (int j = 0; j < s->NJ; j++)
          for (int k = 0; k < s->NK; k++) {
            int d3[3] = {i, j, k}, s3[3] = {i, j, k};
            if (d3[dim] != dst) continue;
            s3[dim] = src;
            a[IDX(i, j, k)] = a[IDX(s3[0], s3[1], s3[2])];
          }
    }
  }
}

/* fixed ghosts for one (dim, side) (src/boundary_conditions.c:247-403 gas,
   :602-686 gem): gas rho = 1, v1..v3 = 0, p = 1 with p_S left as it is; gem
   fields 0..5 = 0; over every row of the other two axes. */
static void bc_fixed(go_solver *s, double *base, int dim, int side, int is_gem) {
  int N[3] = {s->NI, s->NJ, s->NK};
  int n = N[dim];
  for (int f = 0; f < (is_gem ? 6 : 5); f++) {
    double *a = &base[(int64_t)f * s->n3];
    int o = (is_gem && f == dim) ? 1 : 0;
    double v = is_gem ? 0.0 : ((f == gas_rho || f == gas_p) ? 1.0 : 0.0);
    for (int g = 0; g < NG; g++) {
      int dst = side == 0 ? g : n - 1 - NG + g + o;
      for (int i = 0; i < s->NI; i++)
        for (int j = 0; j < s->NJ; j++)
          for (int k = 0; k < s->NK; k++) {
            int d3[3] = {i, j, k};
            if (d3[dim] == dst) a[IDX(i, j, k)] = v;
          }
    }
  }
}

/* boundary_conditions(): gas i low/high, j low/high, k low/high, then gem in
   the same order -- the order the shipped plugins call the BC library. */
static void boundary_conditions(go_solver *s) {
  for (int arr = 0; arr < 2; arr++)
    for (int dim = 0; dim < 3; dim++)
      for (int side = 0; side < 2; side++) {
        if (s->bc[2 * dim + side] == GO_BC_SYMMETRIC)
          bc_mirror(s, arr ? s->gem : s->gas, dim, side, arr, 0);
        if (s->bc[2 * dim + side] == GO_BC_REFLECTIVE && !arr)
          bc_mirror(s, s->gas, dim, side, 0, 1);
        if (s->bc[2 * dim + side] == GO_BC_EXTRAPOLATED)
          bc_extrapolated(s, arr ? s->gem : s->gas, dim, side, arr);
        if (s->bc[2 * dim + side] == GO_BC_FIXED)
          bc_fixed(s, arr ? s->gem : s->gas, dim, side, arr);
        if (s->bc[2 * dim + side] == GO_BC_POLE && dim == 1)
          bc_pole(s, arr ? s->gem : s->gas, side, arr);
      }
}

/* ---------------------------------------------------------------------- */
/* get_derived_variables (src/solver_func.c:308-383) */
static void get_derived_variables(go_solver *s) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->isg; i <= s->ieg; i++)
    for (int j = s->jsg; j <= s->jeg; j++)
      for (int k = s->ksg; k <= s->keg; k++) {
        GEM(mags_b1, i, j, k) = GEM(mag_bi, i, j, k) * (X1(i + 1, j, k) - X1CTR(i, j, k)) / DX1(i, j, k)
            + GEM(mag_bi, i + 1, j, k) * (X1CTR(i, j, k) - X1(i, j, k)) / DX1(i, j, k);
        GEM(mags_b2, i, j, k) = GEM(mag_bj, i, j, k) * (X2(i, j + 1, k) - X2CTR(i, j, k)) / DX2(i, j, k)
            + GEM(mag_bj, i, j + 1, k) * (X2CTR(i, j, k) - X2(i, j, k)) / DX2(i, j, k);
        GEM(mags_b3, i, j, k) = GEM(mag_bk, i, j, k) * (X3(i, j, k + 1) - X3CTR(i, j, k)) / DX3(i, j, k)
            + GEM(mag_bk, i, j, k + 1) * (X3CTR(i, j, k) - X3(i, j, k)) / DX3(i, j, k);
      }

  double dmax = 0.0;
#pragma omp parallel for collapse(2) schedule(static) reduction(max : dmax)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke; k++) {
        double div_b =
            ((GEM(mag_bi, i + 1, j, k) + GEM(B0AfaceI_b1, i + 1, j, k)) * GEO(face_idir, i + 1, j, k)
             - (GEM(mag_bi, i, j, k) + GEM(B0AfaceI_b1, i, j, k)) * GEO(face_idir, i, j, k)) / GEO(vol_center, i, j, k) +
            ((GEM(mag_bj, i, j + 1, k) + GEM(B0AfaceJ_b2, i, j + 1, k)) * GEO(face_jdir, i, j + 1, k)
             - (GEM(mag_bj, i, j, k) + GEM(B0AfaceJ_b2, i, j, k)) * GEO(face_jdir, i, j, k)) / GEO(vol_center, i, j, k) +
            ((GEM(mag_bk, i, j, k + 1) + GEM(B0AfaceK_b3, i, j, k + 1)) * GEO(face_kdir, i, j, k + 1)
             - (GEM(mag_bk, i, j, k) + GEM(B0AfaceK_b3, i, j, k)) * GEO(face_kdir, i, j, k)) / GEO(vol_center, i, j, k);
        if (fabs(div_b) > dmax) dmax = fabs(div_b);
      }
  s->divB_max = dmax;

#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->isg; i <= s->ieg; i++)
    for (int j = s->jsg; j <= s->jeg; j++)
      for (int k = s->ksg; k <= s->keg; k++) {
        GEM(magtot_b1, i, j, k) = GEM(mags_b1, i, j, k) + GEM(B0_b1, i, j, k);
        GEM(magtot_b2, i, j, k) = GEM(mags_b2, i, j, k) + GEM(B0_b2, i, j, k);
        GEM(magtot_b3, i, j, k) = GEM(mags_b3, i, j, k) + GEM(B0_b3, i, j, k);
      }

#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke; k++) {
        double rho = GAS(gas_rho, i, j, k);
        double v1 = GAS(gas_v1, i, j, k), v2 = GAS(gas_v2, i, j, k), v3 = GAS(gas_v3, i, j, k);
        GAS(gasc_rho, i, j, k) = rho;
        GAS(gasc_rhov1, i, j, k) = rho * v1;
        GAS(gasc_rhov2, i, j, k) = rho * v2;
        GAS(gasc_rhov3, i, j, k) = rho * v3;
        GAS(gasc_eng, i, j, k) = 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3) +
                                 GAS(gas_p, i, j, k) / (gamma_val - 1.0);
        GAS(gasc_entropy, i, j, k) = GAS(gas_p_S, i, j, k) / pow(rho, gamma_val - 1.0);
      }
}

/* get_dt_normal (src/solver_func.c:255-298) */
static double get_dt(go_solver *s) {
  double CA = s->CA;
  double m = INFINITY;
#pragma omp parallel for collapse(2) schedule(static) reduction(min : m)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke; k++) {
        double b1 = GEM(magtot_b1, i, j, k), b2 = GEM(magtot_b2, i, j, k), b3 = GEM(magtot_b3, i, j, k);
        double B_tot_sq = b1 * b1 + b2 * b2 + b3 * b3;
        double rho = GAS(gas_rho, i, j, k);
        double Va = sqrt(B_tot_sq / rho);
        double Va_eff = Va * CA / (sqrt(Va * Va + CA * CA));
        double v1 = GAS(gas_v1, i, j, k), v2 = GAS(gas_v2, i, j, k), v3 = GAS(gas_v3, i, j, k);
        double Vfluid = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
        double Vsound = sqrt(gamma_val * GAS(gas_p, i, j, k) / rho);
        double VCFL = Vfluid + sqrt(Vsound * Vsound + Va_eff * Va_eff);
        /* get_dt_ring (src/spherical/spherical_general.c:700-735) differs
           only in the azimuthal width */
        double dxk3 = s->ring ? GEO(dxk_ring, i, j, k) : GEO(dxk, i, j, k);
        double dtCFL = CFL / (VCFL / GEO(dxi, i, j, k) + VCFL / GEO(dxj, i, j, k) +
                              VCFL / dxk3);
        m = fmin(m, dtCFL);
      }
  return m;
}

/* get_AB2_variables (src/solver_func.c:385-492) */
static void get_AB2_variables(go_solver *s) {
  double c = 0.5 * s->dt / s->dt0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->isg; i <= s->ieg; i++)
    for (int j = s->jsg; j <= s->jeg; j++)
      for (int k = s->ksg; k <= s->keg; k++) {
        for (int m = 0; m < 6; m++)
          GAS(gas_rho_h + m, i, j, k) = GAS(m, i, j, k) + c * (GAS(m, i, j, k) - GAS(gas_rho_p + m, i, j, k));
        for (int m = 0; m < 3; m++)
          GEM(mags_b1_h + m, i, j, k) = GEM(mags_b1 + m, i, j, k) + c * (GEM(mags_b1 + m, i, j, k) - GEM(mags_b1_p + m, i, j, k));
      }
  for (int m = 0; m < 3; m++) {
    int iu = s->ieg + (m == 0), ju = s->jeg + (m == 1), ku = s->keg + (m == 2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = s->isg; i <= iu; i++)
      for (int j = s->jsg; j <= ju; j++)
        for (int k = s->ksg; k <= ku; k++)
          GEM(mag_bi_h + m, i, j, k) = GEM(mag_bi + m, i, j, k) + c * (GEM(mag_bi + m, i, j, k) - GEM(mag_bi_p + m, i, j, k));
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->isg; i <= s->ieg; i++)
    for (int j = s->jsg; j <= s->jeg; j++)
      for (int k = s->ksg; k <= s->keg; k++) {
        for (int m = 0; m < 6; m++) GAS(gas_rho_p + m, i, j, k) = GAS(m, i, j, k);
        for (int m = 0; m < 3; m++) GEM(mags_b1_p + m, i, j, k) = GEM(mags_b1 + m, i, j, k);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->isg; i <= s->ieg + 1; i++)
    for (int j = s->jsg; j <= s->jeg + 1; j++)
      for (int k = s->ksg; k <= s->keg + 1; k++)
        for (int m = 0; m < 3; m++) GEM(mag_bi_p + m, i, j, k) = GEM(mag_bi + m, i, j, k);
  s->dt0 = s->dt;
}

/* PDMU7 limiter (src/solver_func.c:215-253) */
static inline double fsign(double x) {
  return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0);
}
static inline double PDMU7(const double fh0[7], const double fp0[7],
                           const double w[7], double faceim1, double facei,
                           double vol, double dx) {
  const int im1 = 2, i0 = 3, ip1 = 4;
  double f_itp = 0.0;
  for (int n = 0; n < 7; n++) f_itp += fh0[n] * w[n];
  double maxf = fmax(fp0[i0], fp0[ip1]);
  double minf = fmin(fp0[i0], fp0[ip1]);
  double f = fmax(minf, fmin(f_itp, maxf));
  double df0 = PDMB * (fp0[i0] - fp0[im1]);
  double df1 = PDMB * (fp0[ip1] - fp0[i0]);
  double s0 = fsign(df0), s1 = fsign(df1);
  df0 = fabs(df0);
  double q0 = fabs(s0 + s1);
  double df_left = f - fp0[i0];
  double q = vol / dx;
  double e = 1.0 / (2 * PDMB + 1.0);
  double df0c = fabs((q - e * facei) / (e * facei) * fp0[i0] +
                     (e * faceim1 - q) / (e * facei) * fp0[im1]) / 2.0;
  double a1 = PDMB / (PDMB + 1.0);
  double a2 = 1.0 - a1;
  df0c = a1 * df0c + a2 * df0;
  return f - s1 * fmax(0, fabs(df_left) - q0 * df0c);
}

/* exported for unit tests: one PDMU7 evaluation */
double go_pdmu7(const double fh0[7], const double fp0[7], const double w[7], double faceim1,
                double facei, double vol, double dx) {
  return PDMU7(fh0, fp0, w, faceim1, facei, vol, dx);
}

/* reconstruct_3dv_{i,j,k} (src/solver_func.c:503-649) for one field array
   base (gas or gem), source field m, limiter field m+pdm, outputs ml, mr. */
static void reconstruct(go_solver *s, double *base, int dir, int pdm, int m,
                        int ml, int mr) {
  int i0 = s->isg, i1 = s->ieg, j0 = s->jsg, j1 = s->jeg, k0 = s->ksg, k1 = s->keg;
  int GFf = GF_face_idir + dir, GFv = GF_vol_idir + dir;
  if (dir == 0) { i0 = s->is; i1 = s->ie + 1; }
  if (dir == 1) { j0 = s->js; j1 = s->je + 1; }
  if (dir == 2) { k0 = s->ks; k1 = s->ke + 1; }
  int64_t st = dir == 0 ? (int64_t)s->NJ * s->NK : (dir == 1 ? s->NK : 1);
  double *dxa = s->crd[9 + dir];
  const double *F = &base[(int64_t)m * s->n3];
  const double *P = &base[(int64_t)(m + pdm) * s->n3];
  double *L = &base[(int64_t)ml * s->n3];
  double *R = &base[(int64_t)mr * s->n3];
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = i0; i <= i1; i++)
    for (int j = j0; j <= j1; j++)
      for (int k = k0; k <= k1; k++) {
        int64_t c = IDX(i, j, k);
        double fh0[7], fp0[7], w[7];
        for (int n = 0; n < 7; n++) w[n] = s->rec[(int64_t)W_L(dir, n) * s->n3 + c];
        for (int off = -4, n = 0; off <= 2; off++, n++) {
          fh0[n] = F[c + off * st];
          fp0[n] = P[c + off * st];
        }
        L[c] = PDMU7(fh0, fp0, w, s->geo[(int64_t)GFf * s->n3 + c - st],
                     s->geo[(int64_t)GFf * s->n3 + c], s->geo[(int64_t)GFv * s->n3 + c - st],
                     dxa[c - st]);
        for (int n = 0; n < 7; n++) w[6 - n] = s->rec[(int64_t)W_R(dir, n) * s->n3 + c];
        for (int off = 3, n = 0; off >= -3; off--, n++) {
          fh0[n] = F[c + off * st];
          fp0[n] = P[c + off * st];
        }
        R[c] = PDMU7(fh0, fp0, w, s->geo[(int64_t)GFf * s->n3 + c + st],
                     s->geo[(int64_t)GFf * s->n3 + c], s->geo[(int64_t)GFv * s->n3 + c],
                     dxa[c]);
      }
}

/* averaging loops of center2corner_* (src/solver_func.c:686-778) */
static void avg_lr(go_solver *s, int mout, int ml, int mr, int i0, int i1,
                   int j0, int j1, int k0, int k1) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = i0; i <= i1; i++)
    for (int j = j0; j <= j1; j++)
      for (int k = k0; k <= k1; k++)
        GAS(mout, i, j, k) = 0.5 * (GAS(ml, i, j, k) + GAS(mr, i, j, k));
}

/* center2corner_ij / _jk / _ki (src/solver_func.c:686-778): first sweep
   along d1 limited against X^n, second along d2 self-limited. */
static void center2corner(go_solver *s, int d1, int d2, int m0, int m_interp,
                          int ml, int mr, int m_avg) {
  reconstruct(s, s->gas, d1, gas_rho_p - gas_rho_h, m0, ml, mr);
  int lo[3] = {s->is, s->js, s->ks}, hi[3] = {s->ie, s->je, s->ke};
  int glo[3] = {s->isg, s->jsg, s->ksg}, ghi[3] = {s->ieg, s->jeg, s->keg};
  int a0[3], a1[3];
  for (int d = 0; d < 3; d++) { a0[d] = lo[d]; a1[d] = hi[d]; }
  a1[d1] = hi[d1] + 1;
  a0[d2] = glo[d2]; a1[d2] = ghi[d2];
  avg_lr(s, m_interp, ml, mr, a0[0], a1[0], a0[1], a1[1], a0[2], a1[2]);
  reconstruct(s, s->gas, d2, 0, m_interp, ml, mr);
  a0[d2] = lo[d2]; a1[d2] = hi[d2] + 1;
  avg_lr(s, m_avg, ml, mr, a0[0], a1[0], a0[1], a1[1], a0[2], a1[2]);
}

/* get_e_fields_{i,j,k} (src/solver_func.c:779-901) */
static void get_e_fields(go_solver *s) {
  const double CA = s->CA;
  const int pdm_gem = mag_bi_p - mag_bi_h;
  /* E_i */
  center2corner(s, 1, 2, gas_v2_h, v2_interp, v2_interp_l, v2_interp_r, v2_avg);
  center2corner(s, 1, 2, gas_v3_h, v3_interp, v3_interp_l, v3_interp_r, v3_avg);
  reconstruct(s, s->gem, 2, pdm_gem, mag_bj_h, bjl, bjr);
  reconstruct(s, s->gem, 1, pdm_gem, mag_bk_h, bkl, bkr);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je + 1; j++)
      for (int k = s->ks; k <= s->ke + 1; k++) {
        double bja = 0.5 * (GEM(bjl, i, j, k) + GEM(bjr, i, j, k)) + GEM(B0AedgeI_b2, i, j, k);
        double bka = 0.5 * (GEM(bkl, i, j, k) + GEM(bkr, i, j, k)) + GEM(B0AedgeI_b3, i, j, k);
        GEM(bj_avg, i, j, k) = bja;
        GEM(bk_avg, i, j, k) = bka;
        double v2a = GAS(v2_avg, i, j, k), v3a = GAS(v3_avg, i, j, k);
        double e = -(v2a * bka - v3a * bja);
        double dv = sqrt((bja * bja + bka * bka) /
                         (0.25 * (GAS(gas_rho_h, i, j - 1, k - 1) + GAS(gas_rho_h, i, j - 1, k) +
                                  GAS(gas_rho_h, i, j, k - 1) + GAS(gas_rho_h, i, j, k))));
        dv = dv * CA / sqrt(dv * dv + CA * CA) + sqrt(v2a * v2a + v3a * v3a);
        GEM(dvzz, i, j, k) = dv;
        e += 0.5 * dv * (GEM(bjl, i, j, k) + GEM(bkr, i, j, k) - GEM(bjr, i, j, k) - GEM(bkl, i, j, k));
        GEM(efield_ei, i, j, k) = e;
      }
  /* E_j */
  center2corner(s, 2, 0, gas_v3_h, v3_interp, v3_interp_l, v3_interp_r, v3_avg);
  center2corner(s, 2, 0, gas_v1_h, v1_interp, v1_interp_l, v1_interp_r, v1_avg);
  reconstruct(s, s->gem, 0, pdm_gem, mag_bk_h, bkl, bkr);
  reconstruct(s, s->gem, 2, pdm_gem, mag_bi_h, bil, bir);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie + 1; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke + 1; k++) {
        double bka = 0.5 * (GEM(bkl, i, j, k) + GEM(bkr, i, j, k)) + GEM(B0AedgeJ_b3, i, j, k);
        double bia = 0.5 * (GEM(bil, i, j, k) + GEM(bir, i, j, k)) + GEM(B0AedgeJ_b1, i, j, k);
        GEM(bk_avg, i, j, k) = bka;
        GEM(bi_avg, i, j, k) = bia;
        double v1a = GAS(v1_avg, i, j, k), v3a = GAS(v3_avg, i, j, k);
        double e = -(v3a * bia - v1a * bka);
        double dv = sqrt((bka * bka + bia * bia) /
                         (0.25 * (GAS(gas_rho_h, i - 1, j, k - 1) + GAS(gas_rho_h, i - 1, j, k) +
                                  GAS(gas_rho_h, i, j, k - 1) + GAS(gas_rho_h, i, j, k))));
        dv = dv * CA / sqrt(dv * dv + CA * CA) + sqrt(v1a * v1a + v3a * v3a);
        GEM(dvzz, i, j, k) = dv;
        e += 0.5 * dv * (GEM(bkl, i, j, k) + GEM(bir, i, j, k) - GEM(bkr, i, j, k) - GEM(bil, i, j, k));
        GEM(efield_ej, i, j, k) = e;
      }
  /* E_k */
  center2corner(s, 0, 1, gas_v1_h, v1_interp, v1_interp_l, v1_interp_r, v1_avg);
  center2corner(s, 0, 1, gas_v2_h, v2_interp, v2_interp_l, v2_interp_r, v2_avg);
  reconstruct(s, s->gem, 1, pdm_gem, mag_bi_h, bil, bir);
  reconstruct(s, s->gem, 0, pdm_gem, mag_bj_h, bjl, bjr);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie + 1; i++)
    for (int j = s->js; j <= s->je + 1; j++)
      for (int k = s->ks; k <= s->ke; k++) {
        double bia = 0.5 * (GEM(bil, i, j, k) + GEM(bir, i, j, k)) + GEM(B0AedgeK_b1, i, j, k);
        double bja = 0.5 * (GEM(bjl, i, j, k) + GEM(bjr, i, j, k)) + GEM(B0AedgeK_b2, i, j, k);
        GEM(bi_avg, i, j, k) = bia;
        GEM(bj_avg, i, j, k) = bja;
        double v1a = GAS(v1_avg, i, j, k), v2a = GAS(v2_avg, i, j, k);
        double e = -(v1a * bja - v2a * bia);
        double dv = sqrt((bia * bia + bja * bja) /
                         (0.25 * (GAS(gas_rho_h, i - 1, j - 1, k) + GAS(gas_rho_h, i - 1, j, k) +
                                  GAS(gas_rho_h, i, j - 1, k) + GAS(gas_rho_h, i, j, k))));
        dv = dv * CA / sqrt(dv * dv + CA * CA) + sqrt(v2a * v2a + v1a * v1a);
        GEM(dvzz, i, j, k) = dv;
        e += 0.5 * dv * (GEM(bil, i, j, k) + GEM(bjr, i, j, k) - GEM(bir, i, j, k) - GEM(bjl, i, j, k));
        GEM(efield_ek, i, j, k) = e;
      }
}

/* get_fluid_flux_rusanov_{i,j,k} (src/solver_func.c:934-1251) followed by
   get_magnetic_stress_rusanov_{i,j,k} (src/solver_func.c:1268-1515). */
static void face_fluxes(go_solver *s, int dir) {
  const double CA = s->CA;
  int i1 = s->ie + (dir == 0), j1 = s->je + (dir == 1), k1 = s->ke + (dir == 2);
  int B0f = B0AfaceI_b1 + 8 * dir; /* b1 b2 b3 bn bsq bn1 bn2 bn3 */
  int magn = mag_bi + dir;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= i1; i++)
    for (int j = s->js; j <= j1; j++)
      for (int k = s->ks; k <= k1; k++) {
        double rl = GAS(gasl_rho, i, j, k), rr = GAS(gasr_rho, i, j, k);
        double pl = GAS(gasl_p, i, j, k), pr = GAS(gasr_p, i, j, k);
        double psl = GAS(gasl_p_S, i, j, k), psr = GAS(gasr_p_S, i, j, k);
        double vl[3] = {GAS(gasl_v1, i, j, k), GAS(gasl_v2, i, j, k), GAS(gasl_v3, i, j, k)};
        double vr[3] = {GAS(gasr_v1, i, j, k), GAS(gasr_v2, i, j, k), GAS(gasr_v3, i, j, k)};
        double bl[3] = {GEM(magsl_b1, i, j, k), GEM(magsl_b2, i, j, k), GEM(magsl_b3, i, j, k)};
        double br[3] = {GEM(magsr_b1, i, j, k), GEM(magsr_b2, i, j, k), GEM(magsr_b3, i, j, k)};
        double b0[3] = {GEM(B0f, i, j, k), GEM(B0f + 1, i, j, k), GEM(B0f + 2, i, j, k)};
        double b0n = GEM(B0f + 3, i, j, k);
        double Vsound_l = sqrt(gamma_val * pl / rl);
        double Vsound_r = sqrt(gamma_val * pr / rr);
        double Vse_l = sqrt(gamma_val * psl / rl);
        double Vse_r = sqrt(gamma_val * psr / rr);
        double Btl = (bl[0] + b0[0]) * (bl[0] + b0[0]) + (bl[1] + b0[1]) * (bl[1] + b0[1]) +
                     (bl[2] + b0[2]) * (bl[2] + b0[2]);
        double Btr = (br[0] + b0[0]) * (br[0] + b0[0]) + (br[1] + b0[1]) * (br[1] + b0[1]) +
                     (br[2] + b0[2]) * (br[2] + b0[2]);
        double Va_l = sqrt(Btl / rl), Va_r = sqrt(Btr / rr);
        double Vae_l = Va_l * CA / (sqrt(Va_l * Va_l + CA * CA));
        double Vae_r = Va_r * CA / (sqrt(Va_r * Va_r + CA * CA));
        double vnl = vl[dir], vnr = vr[dir];
        double VF = fmax(fabs(vnl) + sqrt(Vsound_l * Vsound_l + Vae_l * Vae_l),
                         fabs(vnr) + sqrt(Vsound_r * Vsound_r + Vae_r * Vae_r));
        double VFe = fmax(fabs(vnl) + sqrt(Vse_l * Vse_l + Vae_l * Vae_l),
                          fabs(vnr) + sqrt(Vse_r * Vse_r + Vae_r * Vae_r));
        double rvl[3], rvr[3];
        for (int m = 0; m < 3; m++) { rvl[m] = rl * vl[m]; rvr[m] = rr * vr[m]; }
        double eng_l = 0.5 * rl * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]) + pl / (gamma_val - 1.0);
        double eng_r = 0.5 * rr * (vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]) + pr / (gamma_val - 1.0);
        double S_l = psl / pow(rl, gamma_val - 1.0);
        double S_r = psr / pow(rr, gamma_val - 1.0);
        GAS(gascf_rho, i, j, k) = 0.5 * (rvl[dir] + rvr[dir] - VF * (rr - rl));
        for (int m = 0; m < 3; m++) {
          double fl;
          if (m == dir)
            fl = 0.5 * (rvl[m] * vnl + pl + rvr[m] * vnr + pr - VF * (rvr[m] - rvl[m]));
          else
            fl = 0.5 * (rvl[m] * vnl + rvr[m] * vnr - VF * (rvr[m] - rvl[m]));
          GAS(gascf_rhov1 + m, i, j, k) = fl;
        }
        GAS(gascf_eng, i, j, k) = 0.5 * (eng_l * vnl + pl * vnl + eng_r * vnr + pr * vnr - VF * (eng_r - eng_l));
        GAS(gascf_entropy, i, j, k) = 0.5 * (S_l * vnl + S_r * vnr - VFe * (S_r - S_l));

        /* magnetic stress */
        double Bn = GEM(magn, i, j, k);
        double bsq_l = (bl[0] * bl[0] + bl[1] * bl[1] + bl[2] * bl[2]) +
                       2 * (bl[0] * b0[0] + bl[1] * b0[1] + bl[2] * b0[2]);
        double bsq_r = (br[0] * br[0] + br[1] * br[1] + br[2] * br[2]) +
                       2 * (br[0] * b0[0] + br[1] * b0[1] + br[2] * b0[2]);
        double ms[3];
        for (int m = 0; m < 3; m++) {
          double sp, sn;
          if (m == dir) {
            sp = 0.5 * (0.5 * bsq_l - (bl[m] + b0[m]) * Bn - bl[m] * b0n);
            sn = 0.5 * (0.5 * bsq_r - (br[m] + b0[m]) * Bn - br[m] * b0n);
          } else {
            sp = 0.5 * (-(bl[m] + b0[m]) * Bn - bl[m] * b0n);
            sn = 0.5 * (-(br[m] + b0[m]) * Bn - br[m] * b0n);
          }
          ms[m] = sp + sn;
        }
        double B_sq = (Btl + Btr) / 4 / CA;
        for (int m = 0; m < 3; m++)
          GEM(magsf_x1dir + m, i, j, k) = ms[m] - B_sq * (vr[m] - vl[m]);
      }
}

/* exported for unit tests: the face fluxes of one face (dir 0/1/2) from
   left/right primitive states, left/right cell B, the B0 face block
   (b1 b2 b3 bn bsq bn1 bn2 bn3) and the face-normal B; out[9] = rho, rho v1..3,
   energy, entropy, magnetic stress 1..3.  Runs face_fluxes on a scratch
   solver of the minimum size. */
void go_face_flux(int dir, const double L[6], const double R[6], const double bl[3],
                  const double br[3], const double b0[8], double Bn, double CA, double out[9]) {
  int bc[6] = {0, 0, 0, 0, 0, 0};
  go_solver *s = go_create(5, 5, 5, 0, bc, CA);
  if (!s) return;
  int i = s->is, j = s->js, k = s->ks;
  for (int m = 0; m < 6; m++) {
    GAS(gasl_rho + m, i, j, k) = L[m];
    GAS(gasr_rho + m, i, j, k) = R[m];
  }
  for (int m = 0; m < 3; m++) {
    GEM(magsl_b1 + m, i, j, k) = bl[m];
    GEM(magsr_b1 + m, i, j, k) = br[m];
  }
  for (int q = 0; q < 8; q++) GEM(B0AfaceI_b1 + 8 * dir + q, i, j, k) = b0[q];
  GEM(mag_bi + dir, i, j, k) = Bn;
  /* neighbouring faces of the loop range stay zero-filled; only (is,js,ks) is read back */
  for (int m = 0; m < 6; m++)
    for (int n = 0; n < s->n3; n++) {
      if (s->gas[(int64_t)(gasl_rho + m) * s->n3 + n] == 0.0) s->gas[(int64_t)(gasl_rho + m) * s->n3 + n] = L[m];
      if (s->gas[(int64_t)(gasr_rho + m) * s->n3 + n] == 0.0) s->gas[(int64_t)(gasr_rho + m) * s->n3 + n] = R[m];
    }
  face_fluxes(s, dir);
  for (int m = 0; m < 6; m++) out[m] = GAS(gascf_rho + m, i, j, k);
  for (int m = 0; m < 3; m++) out[6 + m] = GEM(magsf_x1dir + m, i, j, k);
  go_destroy(s);
}

/* spherical Source_Term_i / _j (src/spherical/spherical_general.c:605-688) */
static void source_terms(go_solver *s, int dir) {
  if (s->coord != 2 || dir == 2) return;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke; k++) {
        double V = GEO(vol_center, i, j, k);
        if (dir == 0) {
          double M22 = GAS(gas_p_h, i, j, k) + GAS(gas_rho_h, i, j, k) * GAS(gas_v2_h, i, j, k) * GAS(gas_v2_h, i, j, k);
          double M33 = GAS(gas_p_h, i, j, k) + GAS(gas_rho_h, i, j, k) * GAS(gas_v3_h, i, j, k) * GAS(gas_v3_h, i, j, k);
          GAS(S1_HD, i, j, k) = 0.5 * (M22 + M33) * (GEO(face_idir, i + 1, j, k) - GEO(face_idir, i, j, k)) / V;
          double s21 = -(GAS(gascf_rhov2, i, j, k) * GEO(face_idir, i, j, k) + GAS(gascf_rhov2, i + 1, j, k) * GEO(face_idir, i + 1, j, k)) *
                       DX1(i, j, k) / X1C(i, j, k) / 2.0 / V;
          double s22 = (GEO(face_jdir, i, j + 1, k) - GEO(face_jdir, i, j, k)) / V * M33;
          GAS(S2_HD, i, j, k) = s21 + s22;
          GAS(S3_HD, i, j, k) = -(GAS(gascf_rhov3, i, j, k) * GEO(face_idir, i, j, k) + GAS(gascf_rhov3, i + 1, j, k) * GEO(face_idir, i + 1, j, k)) *
                                DX1(i, j, k) / X1C(i, j, k) / 2.0 / V;
          double b1 = GEM(mags_b1_h, i, j, k), b2 = GEM(mags_b2_h, i, j, k), b3 = GEM(mags_b3_h, i, j, k);
          double c1 = GEM(B0_b1, i, j, k), c2 = GEM(B0_b2, i, j, k), c3 = GEM(B0_b3, i, j, k);
          double Bsq = b1 * b1 + b2 * b2 + b3 * b3 + 2 * (b1 * c1 + b2 * c2 + b3 * c3);
          double M22b = (0.5 * Bsq - (b2 + c2) * (b2) - b2 * c2);
          double M33b = (0.5 * Bsq - (b3 + c3) * (b3) - b3 * c3);
          GEM(S1_B, i, j, k) = 0.5 * (M22b + M33b) * (GEO(face_idir, i + 1, j, k) - GEO(face_idir, i, j, k)) / V;
          double b21 = -(GEM(magsf_x2dir, i, j, k) * GEO(face_idir, i, j, k) + GEM(magsf_x2dir, i + 1, j, k) * GEO(face_idir, i + 1, j, k)) *
                       DX1(i, j, k) / X1C(i, j, k) / 2.0 / V;
          double b22 = (GEO(face_jdir, i, j + 1, k) - GEO(face_jdir, i, j, k)) / V * M33b;
          GEM(S2_B, i, j, k) = b21 + b22;
          GEM(S3_B, i, j, k) = -(GEM(magsf_x3dir, i, j, k) * GEO(face_idir, i, j, k) + GEM(magsf_x3dir, i + 1, j, k) * GEO(face_idir, i + 1, j, k)) *
                               DX1(i, j, k) / X1C(i, j, k) / 2.0 / V;
        } else {
          double sd = sin(X2(i, j + 1, k)) - sin(X2(i, j, k));
          double ss = sin(X2(i, j + 1, k)) + sin(X2(i, j, k));
          GAS(S3_HD, i, j, k) += -(GAS(gascf_rhov3, i, j, k) * GEO(face_jdir, i, j, k) + GAS(gascf_rhov3, i, j + 1, k) * GEO(face_jdir, i, j + 1, k)) *
                                 sd / ss / V;
          GEM(S3_B, i, j, k) += -(GEM(magsf_x3dir, i, j, k) * GEO(face_jdir, i, j, k) + GEM(magsf_x3dir, i, j + 1, k) * GEO(face_jdir, i, j + 1, k)) *
                                sd / ss / V;
        }
      }
}

/* flux divergence (src/solver_solve.c:178-255) */
static void flux_divergence(go_solver *s, int dir) {
  int64_t st = dir == 0 ? (int64_t)s->NJ * s->NK : (dir == 1 ? s->NK : 1);
  int fA = face_idir + dir;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke; k++) {
        int64_t c = IDX(i, j, k);
        double A1 = s->geo[(int64_t)fA * s->n3 + c + st], A0 = s->geo[(int64_t)fA * s->n3 + c];
        double V = s->geo[(int64_t)vol_center * s->n3 + c];
        for (int m = 0; m < 6; m++) {
          double *F = &s->gas[(int64_t)(gascf_rho + m) * s->n3];
          double d = (F[c + st] * A1 - F[c] * A0) / V;
          if (dir == 0) GAS(dgascf_rho + m, i, j, k) = d;
          else GAS(dgascf_rho + m, i, j, k) += d;
        }
        for (int m = 0; m < 3; m++) {
          double *F = &s->gem[(int64_t)(magsf_x1dir + m) * s->n3];
          double d = (F[c + st] * A1 - F[c] * A0) / V;
          if (dir == 0) GEM(dmagsf_x1dir + m, i, j, k) = d;
          else GEM(dmagsf_x1dir + m, i, j, k) += d;
        }
      }
}

/* one pass of the solve() loop body (src/solver_solve.c:103-508) */

/* ---------------------------------------------------------------------- */
/* Pole ring averaging (src/spherical/spherical_RingAverage.c, helpers
   src/solver_func.c:2136-2257).  Single rank: the gather / broadcast copies
   (src/spherical/spherical_mpi.c) are the identity, so the routines work on
   gas / gem in place; the ring E fields live in a scratch block. */

/* blocks per ring row by azimuthal resolution
   (src/spherical/spherical_RingAverage.c:55-79) */
int go_set_ring(go_solver *s, int on, int pole_lo, int pole_hi) {
  static const int t256[16] = {8, 16, 32, 32, 64, 64, 64, 64, 128, 128, 128, 128, 128, 128, 128, 128};
  s->ring = on;
  s->pole[0] = on && pole_lo;
  s->pole[1] = on && pole_hi;
  if (!on) return 0;
  int n;
  if (s->nk == 256) n = 16;
  else if (s->nk == 128) n = 8;
  else if (s->nk == 64) n = 4;
  else if (s->nk == 32) n = 2;
  else return -1;
  s->navr = n;
  for (int r = 0; r < n; r++) s->nchunk[r] = t256[r];
  /* geo[dxk_ring] (src/spherical/spherical_RingAverage.c:81-147) */
  for (int i = s->isg; i <= s->ieg; i++)
    for (int j = s->jsg; j <= s->jeg; j++)
      for (int k = s->ksg; k <= s->keg; k++) GEO(dxk_ring, i, j, k) = GEO(dxk, i, j, k);
  for (int p = 0; p < 2; p++) {
    if (!s->pole[p]) continue;
    for (int r = 0; r < n; r++) {
      int Nm = s->nk / s->nchunk[r];
      int j = p == 0 ? s->js + r : s->je - r;
      for (int i = s->is; i <= s->ie; i++)
        for (int k = s->ks; k <= s->ke; k++) GEO(dxk_ring, i, j, k) = Nm * GEO(dxk, i, j, k);
    }
  }
  return n;
}

/* PPM (src/solver_func.c:2148-2185) with minmod (:2139-2141) */
static double ring_minmod(double x, double y) {
  return 0.5 * (fsign(x) + fsign(y)) * fmin(fabs(x), fabs(y));
}
static void ring_ppm(double fm2, double fm1, double f, double fp1, double fp2, double *vm,
                     double *vp) {
  double dvm2 = fm1 - fm2, dvm1 = f - fm1, dvp1 = fp1 - f, dvp2 = fp2 - fp1;
  double dvc = 0.5 * (dvm2 + dvm1);
  double SM = 2.0 * ring_minmod(dvm2, dvm1);
  double Sm1 = ring_minmod(dvc, SM);
  dvc = 0.5 * (dvm1 + dvp1);
  SM = 2.0 * ring_minmod(dvm1, dvp1);
  double Sp1 = ring_minmod(dvc, SM);
  dvc = 0.5 * (dvp2 + dvp1);
  SM = 2.0 * ring_minmod(dvp2, dvp1);
  double Sp2 = ring_minmod(dvc, SM);
  *vp = 0.5 * (f + fp1) - (Sp2 - Sp1) / 6.0;
  *vm = 0.5 * (f + fm1) - (Sp1 - Sm1) / 6.0;
  double ap = *vp - f, am = *vm - f;
  if (ap * am >= 0.0) {
    ap = 0.0;
    am = 0.0;
  } else {
    if (fabs(ap) >= 2.0 * fabs(am)) ap = -2.0 * am;
    if (fabs(am) >= 2.0 * fabs(ap)) am = -2.0 * ap;
  }
  *vp = f + ap;
  *vm = f + am;
}

/* FilterBaseMode (src/solver_func.c:2187-2212): Q0 := the m = 0, 1 Fourier
   part of Q, Q -= Q0 */
static void ring_filter(double *Q, int N, double *Q0) {
  const double scl = 2.0 / N, dtheta = 2.0 * 3.14159265358979323846 / N;
  double a0 = 0.0, a1 = 0.0, b1 = 0.0;
  for (int i = 0; i < N; i++) {
    double th = dtheta * (i + 0.5);
    a0 += Q[i];
    a1 += Q[i] * cos(th);
    b1 += Q[i] * sin(th);
  }
  a1 *= scl;
  b1 *= scl;
  a0 *= 0.5 * scl;
  for (int i = 0; i < N; i++) {
    double th = dtheta * (i + 0.5);
    Q0[i] = a0 + a1 * cos(th) + b1 * sin(th);
    Q[i] -= Q0[i];
  }
}

/* RingBlockAvg (src/solver_func.c:2214-2257): block means of nch blocks,
   periodic PPM edge values per block, parabola integrated over each cell */
static void ring_block_avg(double *Q, int N, int nch) {
  int ncell = N / nch;
  double blk[128], ext[132];
  for (int ib = 0; ib < nch; ib++) {
    double sum = 0.0;
    for (int ic = 0; ic < ncell; ic++) sum += Q[ib * ncell + ic];
    blk[ib] = sum / ncell;
  }
  ext[0] = blk[nch - 2];
  ext[1] = blk[nch - 1];
  for (int ib = 0; ib < nch; ib++) ext[ib + 2] = blk[ib];
  ext[nch + 2] = blk[0];
  ext[nch + 3] = blk[1];
  for (int ib = 2; ib < nch + 2; ib++) {
    int ig = (ib - 2) * ncell;
    double fL, fR;
    ring_ppm(ext[ib - 2], ext[ib - 1], ext[ib], ext[ib + 1], ext[ib + 2], &fL, &fR);
    double a = 3.0 * (fL + fR - 2.0 * ext[ib]);
    double b = 2.0 * (3.0 * ext[ib] - fR - 2.0 * fL);
    double c = fL;
    for (int ic = 0; ic < ncell; ic++) {
      double kl = (double)ic + 1.0;
      Q[ig + ic] = (a / 3.0 / ncell / ncell) * (3 * kl * kl - 3 * kl + 1) + 0.5 * b * (2 * kl - 1) / ncell + c;
    }
  }
}

void go_ring_filter(double *Q, int N, double *Q0) { ring_filter(Q, N, Q0); }
void go_ring_block_avg(double *Q, int N, int nch) { ring_block_avg(Q, N, nch); }

/* tackle_Efield_pole (src/spherical/spherical_RingAverage.c:153-201):
   E_i on the pole edges := its ring mean over nk + 1 edges; E_k := 0 */
static void ring_efield_pole(go_solver *s) {
  for (int p = 0; p < 2; p++) {
    if (!s->pole[p]) continue;
    int jp = p == 0 ? s->js : s->je + 1;
    for (int i = s->is; i <= s->ie; i++) {
      double m = 0.0;
      for (int k = s->ks; k <= s->ke + 1; k++) m = m + GEM(efield_ei, i, jp, k);
      m = m / (s->nk + 1);
      for (int k = s->ks; k <= s->ke + 1; k++) GEM(efield_ei, i, jp, k) = m;
    }
    for (int i = s->is; i <= s->ie + 1; i++)
      for (int k = s->ks; k <= s->ke; k++) GEM(efield_ek, i, jp, k) = 0.0;
  }
}

/* tackle_Magfield_pole (src/spherical/spherical_RingAverage.c:203-235):
   B_j on the pole face from the next face row and its antipode in phi */
static void ring_magfield_pole(go_solver *s) {
  for (int p = 0; p < 2; p++) {
    if (!s->pole[p]) continue;
    int jp = p == 0 ? s->js : s->je + 1, jn = p == 0 ? s->js + 1 : s->je;
    for (int i = s->is; i <= s->ie; i++)
      for (int k = s->ks; k <= s->ke; k++)
        GEM(mag_bj, i, jp, k) = 0.5 * (GEM(mag_bj, i, jn, k) -
                                       GEM(mag_bj, i, jn, NG + (k - NG + s->nk / 2) % s->nk));
  }
}

/* HydroRingAverage_base (src/spherical/spherical_RingAverage.c:237-377):
   rho and p block-averaged; each velocity component keeps its m = 0, 1 part
   and block-averages the rest */
static void ring_hydro(go_solver *s) {
  const int N = s->nk;
  for (int p = 0; p < 2; p++) {
    if (!s->pole[p]) continue;
#pragma omp parallel for schedule(static)
    for (int i = s->is; i <= s->ie; i++) {
      double Q[256], base[256], rem[256];
      for (int r = 0; r < s->navr; r++) {
        int j = p == 0 ? s->js + r : s->je - r;
        for (int f = gas_rho; f <= gas_p; f++) {
          for (int k = 0; k < N; k++) Q[k] = GAS(f, i, j, s->ks + k);
          if (f == gas_rho || f == gas_p) {
            ring_block_avg(Q, N, s->nchunk[r]);
          } else {
            ring_filter(Q, N, base);
            memcpy(rem, Q, sizeof(double) * N);
            ring_block_avg(rem, N, s->nchunk[r]);
            for (int k = 0; k < N; k++) Q[k] = rem[k] + base[k];
          }
          for (int k = 0; k < N; k++) GAS(f, i, j, s->ks + k) = Q[k];
        }
      }
    }
  }
}

/* MagneticRingAverage_base (src/spherical/spherical_RingAverage.c:380-677):
   the face-B flux through each ring row is block-averaged (m = 0, 1 kept);
   the change is written as edge E fields inside each block (cumulative sums
   of the flux differences) and applied by a curl update, so div B is
   unchanged; then the pole-face B_j copy and the zero-mean pole-row B_k. */
static void ring_magnetic(go_solver *s, double *E) {
  const int N = s->nk;
  double *Ei = E, *Ej = E + s->n3;
  memset(E, 0, sizeof(double) * 2 * (size_t)s->n3);
  for (int p = 0; p < 2; p++) {
    if (!s->pole[p]) continue;
    for (int comp = 0; comp < 2; comp++) {
      /* comp 0: B_j flux on face rows next to the cell rows, E_i (negated);
         comp 1: B_i flux on the cell rows, E_j */
      int i1 = comp == 0 ? s->ie : s->ie + 1;
#pragma omp parallel for schedule(static)
      for (int i = s->is; i <= i1; i++) {
        double fl[256], base[256], rem[256];
        for (int r = 0; r < s->navr; r++) {
          int j = comp == 0 ? (p == 0 ? s->js + r + 1 : s->je - r) : (p == 0 ? s->js + r : s->je - r);
          for (int k = 0; k < N; k++)
            fl[k] = comp == 0 ? GEM(mag_bj, i, j, s->ks + k) * GEO(face_jdir, i, j, s->ks + k)
                              : GEM(mag_bi, i, j, s->ks + k) * GEO(face_idir, i, j, s->ks + k);
          ring_filter(fl, N, base);
          memcpy(rem, fl, sizeof(double) * N);
          ring_block_avg(rem, N, s->nchunk[r]);
          int ncell = N / s->nchunk[r];
          for (int blk = 0; blk < s->nchunk[r]; blk++) {
            int start = blk * ncell;
            double sum = 0.0;
            for (int m = 0; m < ncell - 1; m++) {
              sum += rem[start + m] - fl[start + m];
              if (comp == 0) Ei[IDX(i, j, start + s->ks + 1 + m)] = -sum;
              else Ej[IDX(i, j, start + s->ks + 1 + m)] = sum;
            }
          }
        }
      }
    }
  }
  /* curl update over the active faces (E_k is zero) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie + 1; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke; k++) {
        double c = 1.0 * ((0.0 - 0.0) - (Ej[IDX(i, j, k + 1)] - Ej[IDX(i, j, k)]));
        GEM(mag_bi, i, j, k) -= c / GEO(face_idir, i, j, k);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je + 1; j++)
      for (int k = s->ks; k <= s->ke; k++) {
        double c = 1.0 * ((Ei[IDX(i, j, k + 1)] - Ei[IDX(i, j, k)]) - (0.0 - 0.0));
        GEM(mag_bj, i, j, k) -= c / GEO(face_jdir, i, j, k);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke + 1; k++) {
        double c = 1.0 * ((Ej[IDX(i + 1, j, k)] - Ej[IDX(i, j, k)]) - (Ei[IDX(i, j + 1, k)] - Ei[IDX(i, j, k)]));
        GEM(mag_bk, i, j, k) -= c / GEO(face_kdir, i, j, k);
      }
  for (int p = 0; p < 2; p++) {
    if (!s->pole[p]) continue;
    int jp = p == 0 ? s->js : s->je + 1, jn = p == 0 ? s->js + 1 : s->je;
    for (int i = s->is; i <= s->ie; i++)
      for (int k = s->ks; k <= s->ke; k++) GEM(mag_bj, i, jp, k) = GEM(mag_bj, i, jn, k);
  }
  for (int p = 0; p < 2; p++) {
    if (!s->pole[p]) continue;
    int jr = p == 0 ? s->js : s->je;
    for (int i = s->is; i <= s->ie; i++) {
      double m = 0.0;
      for (int k = s->ks; k <= s->ke + 1; k++) m += GEM(mag_bk, i, jr, k);
      m = m / (s->nk + 1);
      for (int k = s->ks; k <= s->ke + 1; k++) GEM(mag_bk, i, jr, k) -= m;
    }
  }
}

static void step_once(go_solver *s) {
  const double CA = s->CA;
  double dt = get_dt(s);
  s->dt = dt;
  s->time_sim += dt;
  get_AB2_variables(s);

  /* stage saves (src/solver_solve.c:119-162) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke; k++) {
        for (int m = 0; m < 6; m++) GAS(gasc0_rho + m, i, j, k) = GAS(gasc_rho + m, i, j, k);
        GAS(v1_0, i, j, k) = GAS(gas_v1, i, j, k);
        GAS(v2_0, i, j, k) = GAS(gas_v2, i, j, k);
        GAS(v3_0, i, j, k) = GAS(gas_v3, i, j, k);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->isg; i <= s->ieg + 1; i++)
    for (int j = s->jsg; j <= s->jeg + 1; j++)
      for (int k = s->ksg; k <= s->keg + 1; k++)
        for (int m = 0; m < 3; m++) GEM(mag0_bi + m, i, j, k) = GEM(mag_bi + m, i, j, k) * 1.0;

  get_e_fields(s);

  for (int dir = 0; dir < 3; dir++) {
    for (int m = 0; m < 6; m++)
      reconstruct(s, s->gas, dir, gas_rho_p - gas_rho_h, gas_rho_h + m, gasl_rho + m, gasr_rho + m);
    for (int m = 0; m < 3; m++)
      reconstruct(s, s->gem, dir, mag_bi_p - mag_bi_h, mags_b1_h + m, magsl_b1 + m, magsr_b1 + m);
    face_fluxes(s, dir);
    if (s->dbgF) {
      for (int m = 0; m < 6; m++)
        memcpy(&s->dbgF[(int64_t)(dir * 9 + m) * s->n3], &s->gas[(int64_t)(gascf_rho + m) * s->n3],
               sizeof(double) * (size_t)s->n3);
      for (int m = 0; m < 3; m++)
        memcpy(&s->dbgF[(int64_t)(dir * 9 + 6 + m) * s->n3], &s->gem[(int64_t)(magsf_x1dir + m) * s->n3],
               sizeof(double) * (size_t)s->n3);
      for (int m = 0; m < 6; m++)
        memcpy(&s->dbgF[(int64_t)(27 + dir * 6 + m) * s->n3], &s->gem[(int64_t)(magsl_b1 + m) * s->n3],
               sizeof(double) * (size_t)s->n3);
    }
    source_terms(s, dir);
    flux_divergence(s, dir);
  }

  if (s->ring) ring_efield_pole(s); /* src/solver_solve.c:258-260 */

  /* CT update (src/solver_solve.c:263-306) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie + 1; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke; k++)
        GEM(mag_bi, i, j, k) -= dt * ((GEM(efield_ek, i, j + 1, k) * GEO(edge_kdir, i, j + 1, k) -
                                       GEM(efield_ek, i, j, k) * GEO(edge_kdir, i, j, k)) -
                                      (GEM(efield_ej, i, j, k + 1) * GEO(edge_jdir, i, j, k + 1) -
                                       GEM(efield_ej, i, j, k) * GEO(edge_jdir, i, j, k))) /
                                GEO(face_idir, i, j, k);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je + 1; j++)
      for (int k = s->ks; k <= s->ke; k++)
        GEM(mag_bj, i, j, k) -= dt * ((GEM(efield_ei, i, j, k + 1) * GEO(edge_idir, i, j, k + 1) -
                                       GEM(efield_ei, i, j, k) * GEO(edge_idir, i, j, k)) -
                                      (GEM(efield_ek, i + 1, j, k) * GEO(edge_kdir, i + 1, j, k) -
                                       GEM(efield_ek, i, j, k) * GEO(edge_kdir, i, j, k))) /
                                GEO(face_jdir, i, j, k);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke + 1; k++)
        GEM(mag_bk, i, j, k) -= dt * ((GEM(efield_ej, i + 1, j, k) * GEO(edge_jdir, i + 1, j, k) -
                                       GEM(efield_ej, i, j, k) * GEO(edge_jdir, i, j, k)) -
                                      (GEM(efield_ei, i, j + 1, k) * GEO(edge_idir, i, j + 1, k) -
                                       GEM(efield_ei, i, j, k) * GEO(edge_idir, i, j, k))) /
                                GEO(face_kdir, i, j, k);

  if (s->ring) ring_magfield_pole(s); /* src/solver_solve.c:308-310 */

  /* B0 Lorentz add, conservative update, cons->prim, positivity, cell B,
     magtot and the Boris momentum update (src/solver_solve.c:314-484,
     src/solver_func.c:2117-2135) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->is; i <= s->ie; i++)
    for (int j = s->js; j <= s->je; j++)
      for (int k = s->ks; k <= s->ke; k++) {
        for (int m = 0; m < 3; m++) GEM(dmagsf_x1dir + m, i, j, k) += GEM(dmagsf_G_x1dir + m, i, j, k);
        GAS(gasc_rho, i, j, k) -= dt * GAS(dgascf_rho, i, j, k);
        for (int m = 0; m < 3; m++)
          GAS(gasc_rhov1 + m, i, j, k) = GAS(gasc_rhov1 + m, i, j, k) - dt * GAS(dgascf_rhov1 + m, i, j, k) +
                                         dt * GAS(S1_HD + m, i, j, k);
        GAS(gasc_eng, i, j, k) -= dt * GAS(dgascf_eng, i, j, k);
        GAS(gasc_entropy, i, j, k) -= dt * GAS(dgascf_entropy, i, j, k);

        double rho = GAS(gasc_rho, i, j, k);
        GAS(gas_rho, i, j, k) = rho;
        double v1 = GAS(gasc_rhov1, i, j, k) / rho, v2 = GAS(gasc_rhov2, i, j, k) / rho,
               v3 = GAS(gasc_rhov3, i, j, k) / rho;
        GAS(gas_v1, i, j, k) = v1;
        GAS(gas_v2, i, j, k) = v2;
        GAS(gas_v3, i, j, k) = v3;
        GAS(gas_p, i, j, k) = (GAS(gasc_eng, i, j, k) - 0.5 * (v1 * v1 + v2 * v2 + v3 * v3) * rho) * (gamma_val - 1.0);
        GAS(gas_p_S, i, j, k) = GAS(gasc_entropy, i, j, k) * pow(rho, gamma_val - 1.0);
        if (GAS(gas_rho, i, j, k) < s->rho_floor) {
          GAS(gas_rho, i, j, k) = s->rho_floor;
          GAS(gas_p, i, j, k) = s->p_floor;
        }
        if (GAS(gas_p, i, j, k) < s->p_floor) GAS(gas_p, i, j, k) = s->p_floor;

        GEM(mags_b1, i, j, k) = GEM(mag_bi, i, j, k) * (X1(i + 1, j, k) - X1CTR(i, j, k)) / DX1(i, j, k)
            + GEM(mag_bi, i + 1, j, k) * (X1CTR(i, j, k) - X1(i, j, k)) / DX1(i, j, k);
        GEM(mags_b2, i, j, k) = GEM(mag_bj, i, j, k) * (X2(i, j + 1, k) - X2CTR(i, j, k)) / DX2(i, j, k)
            + GEM(mag_bj, i, j + 1, k) * (X2CTR(i, j, k) - X2(i, j, k)) / DX2(i, j, k);
        GEM(mags_b3, i, j, k) = GEM(mag_bk, i, j, k) * (X3(i, j, k + 1) - X3CTR(i, j, k)) / DX3(i, j, k)
            + GEM(mag_bk, i, j, k + 1) * (X3CTR(i, j, k) - X3(i, j, k)) / DX3(i, j, k);
        for (int m = 0; m < 3; m++)
          GEM(magtot_b1 + m, i, j, k) = 0.5 * (GEM(mags_b1 + m, i, j, k) + GEM(mags_b1_p + m, i, j, k)) + GEM(B0_b1 + m, i, j, k);

        double mt[3] = {GEM(magtot_b1, i, j, k), GEM(magtot_b2, i, j, k), GEM(magtot_b3, i, j, k)};
        double alf = (mt[0] * mt[0] + mt[1] * mt[1] + mt[2] * mt[2]) / GAS(gas_rho, i, j, k) / (CA * CA);
        double perp = 1 / (1 + alf);
        double drho = -dt * GAS(dgascf_rho, i, j, k);
        double dva = alf * drho;
        GEM(alf_ratio, i, j, k) = alf;
        GEM(perp_ratio, i, j, k) = perp;
        GEM(dv_alf, i, j, k) = dva;
        double vt[3], v0[3];
        for (int m = 0; m < 3; m++) {
          v0[m] = GAS(v1_0 + m, i, j, k);
          vt[m] = alf * (-dt * GAS(dgascf_rhov1 + m, i, j, k) + dt * GAS(S1_HD + m, i, j, k) - (drho) * v0[m]);
          GAS(v1_tmp + m, i, j, k) = vt[m];
        }
        double bdv = mt[0] * vt[0] + mt[1] * vt[1] + mt[2] * vt[2];
        bdv = bdv / (mt[0] * mt[0] + mt[1] * mt[1] + mt[2] * mt[2] + 1e-10);
        GAS(bdotv, i, j, k) = bdv;
        for (int m = 0; m < 3; m++) {
          double rv = GAS(gasc0_rhov1 + m, i, j, k) +
                      perp * ((-dt * GAS(dgascf_rhov1 + m, i, j, k) + dt * GAS(S1_HD + m, i, j, k)) +
                              (-dt * GEM(dmagsf_x1dir + m, i, j, k) + dt * GEM(S1_B + m, i, j, k)) +
                              dva * v0[m] + mt[m] * bdv);
          rv += dt * GAS(gasc_rho, i, j, k) * s->G0[m];
          GAS(gasc_rhov1 + m, i, j, k) = rv;
        }
        for (int m = 0; m < 3; m++)
          GAS(gas_v1 + m, i, j, k) = GAS(gasc_rhov1 + m, i, j, k) / GAS(gasc_rho, i, j, k);
      }

  if (s->ring) { /* src/solver_solve.c:486-491 */
    ring_hydro(s);
    double *E = (double *)malloc(sizeof(double) * 2 * (size_t)s->n3);
    if (E) ring_magnetic(s, E);
    free(E);
  }

  boundary_exchange(s);
  boundary_conditions(s);
  get_derived_variables(s);
}

void go_prologue(go_solver *s) {
  boundary_exchange(s);
  boundary_conditions(s);
  get_derived_variables(s);
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->isg; i <= s->ieg; i++)
    for (int j = s->jsg; j <= s->jeg; j++)
      for (int k = s->ksg; k <= s->keg; k++) {
        for (int m = 0; m < 6; m++) GAS(gas_rho_p + m, i, j, k) = GAS(m, i, j, k);
        for (int m = 0; m < 3; m++) GEM(mags_b1_p + m, i, j, k) = GEM(mags_b1 + m, i, j, k);
      }
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = s->isg; i <= s->ieg + 1; i++)
    for (int j = s->jsg; j <= s->jeg + 1; j++)
      for (int k = s->ksg; k <= s->keg + 1; k++)
        for (int m = 0; m < 3; m++) GEM(mag_bi_p + m, i, j, k) = GEM(mag_bi + m, i, j, k);
  s->dt0 = get_dt(s);
}

double *go_debug_fluxes(go_solver *s) {
  if (!s->dbgF) s->dbgF = (double *)calloc((size_t)(45 * s->n3), sizeof(double));
  return s->dbgF;
}

void go_step(go_solver *s, int n) {
  for (int t = 0; t < n; t++) step_once(s);
}
