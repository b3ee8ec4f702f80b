/*
 * Solver entry points of the host facade: initialize_solver, solve,
 * finalize_solver (reference src/solver_init.c:352-450 and
 * src/solver_solve.c:23-536), backed by the gamera_op device solver.
 *
 * solve(nt) keeps the reference's semantics on a single GPU: the prologue
 * (ghost fill, boundary conditions, derived variables, X_p := X, dt0) and then
 * nt AB2 steps, stopping early once time_sim exceeds time_stop.  The state
 * lives in device memory between calls; the host arrays are refreshed after
 * solve() returns and whenever a dump is due.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <unistd.h>

#include "common.h"
#include "config.h"
#include "curvilinear.h"
#include "gk_host.h"
#include "log.h"
#include "problem.h"
#include "solver.h"

/* problem selection of the seeded plugin (problems/gk_seeded.c) */
extern int gk_problem_kind;
extern uint64_t gk_problem_seed;
extern int gk_problem_seed_set;

/* optional problem hook: rewrite x1/x2/x3 before the metrics are computed */
int gk_problem_custom_grid(void) __attribute__((weak));

static gk_solver *g_dev = NULL;
static gk_tables g_tab;
static int g_precision = 0;
static int g_dev_fresh = 1; /* prologue pending */

int gk_host_dump_hook(const char *base) __attribute__((weak));

static unsigned char gas_keep[NF_gas], gem_keep[NF_gem];

int allocate_solver() {
  memset(gas_keep, 0, sizeof gas_keep);
  memset(gem_keep, 0, sizeof gem_keep);
  for (int f = 0; f < NF_gas_store; f++) gas_keep[f] = 1;
  for (int f = 0; f < NF_gem_store; f++) gem_keep[f] = 1;
  for (int f = LAi; f <= dmagsf_G_x3dir; f++) gem_keep[f] = 1;
  for (int f = magtot_b1; f <= magtot_b3; f++) gem_keep[f] = 1;
  double ****arrs[12];
  (void)arrs;
  x1 = gk_alloc3(); x2 = gk_alloc3(); x3 = gk_alloc3();
  x1c = gk_alloc3(); x2c = gk_alloc3(); x3c = gk_alloc3();
  x1ctr = gk_alloc3(); x2ctr = gk_alloc3(); x3ctr = gk_alloc3();
  dx1 = gk_alloc3(); dx2 = gk_alloc3(); dx3 = gk_alloc3();
  geo = gk_alloc4(NF_geo, NULL);
  rec = gk_alloc4(NF_rec, NULL);
  gem = gk_alloc4(NF_gem, gem_keep);
  gas = (double *****)calloc(NS1, sizeof(double ****));
  if (gas) gas[0] = gk_alloc4(NF_gas, gas_keep);
  if (!x1 || !x2 || !x3 || !x1c || !x2c || !x3c || !x1ctr || !x2ctr || !x3ctr || !dx1 || !dx2 ||
      !dx3 || !geo || !rec || !gem || !gas || !gas[0])
    return -1;
  for (int f = 0; f < NF_geo; f++)
    if (!geo[f]) return -1;
  for (int f = 0; f < NF_rec; f++)
    if (!rec[f]) return -1;
  for (int f = 0; f < NF_gem; f++)
    if (gem_keep[f] && !gem[f]) return -1;
  for (int f = 0; f < NF_gas; f++)
    if (gas_keep[f] && !gas[0][f]) return -1;
  return 0;
}

/* dumps and restart lookup: on for the reference's main() sequence, off for
   the programmatic set-up (gkh_setup), which never touches the disk */
int gk_io_enabled = 1;

/* The HDF5 layer lives in libgamera_h5.so next to this library (it links the
   HDF5 runtime); it is loaded on first use.  Without it the facade runs
   without dumps, restarts or grid files and says so once. */
typedef struct {
  int (*dump)(const char *);
  int (*extra)(const char *);
  int (*restart)(const char *);
  int (*grid)(void);
} gk_h5_api;

static const gk_h5_api *gk_h5(void) {
  static int tried = 0, ok = 0;
  static gk_h5_api api;
  if (!tried) {
    tried = 1;
    Dl_info info;
    if (dladdr((void *)&gk_h5, &info) && info.dli_fname) {
      char path[4096];
      snprintf(path, sizeof path, "%s", info.dli_fname);
      char *slash = strrchr(path, '/');
      if (slash)
        snprintf(slash + 1, sizeof path - (size_t)(slash + 1 - path), "%s", "libgamera_h5.so");
      else
        snprintf(path, sizeof path, "%s", "libgamera_h5.so");
      void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
      if (h) {
        api.dump = (int (*)(const char *))dlsym(h, "dump_data_hdf5");
        api.extra = (int (*)(const char *))dlsym(h, "dump_extra_data");
        api.restart = (int (*)(const char *))dlsym(h, "gk_read_restart_file");
        api.grid = (int (*)(void))dlsym(h, "gk_read_grid_file");
        ok = api.dump && api.extra && api.restart && api.grid;
      }
    }
    if (!ok) log_warn("HDF5 layer not available: no dumps, restarts or grid files");
  }
  return ok ? &api : NULL;
}

int set_domain_decomposition() {
  gk_default_decomposition();
  /* a grid file in the working directory overrides the uniform grid */
  if (gk_io_enabled && gk_h5()) {
    int st = gk_h5()->grid();
    if (st < 0) {
      log_error("Error reading grid.h5");
      return -1;
    }
  }
  return 0;
}

int initialize_solver(const char *base_name) {
  log_info("Allocating solver");
  if (allocate_solver() != 0) {
    log_error("host allocation failed");
    return -1;
  }
  read_restart = false;
  time_sim = 0.0;
  time_sim_start = 0.0;
  if (set_domain_decomposition() != 0) return -1;
  int st = 1;
  /* restart from this rank's last dump (src/solver_init.c:352-376) */
  if (gk_io_enabled && gk_h5()) st = gk_h5()->restart(base_name);
  if (st < 0) return st;
  read_restart = st == 0;
  if (read_restart) time_sim_start = time_sim;
  if (!read_restart && gk_problem_custom_grid) gk_problem_custom_grid();
  if (set_geometry_arrays() != 0) return -1;
  log_info("Solver initialized");
  return 0;
}

int finalize_solver() {
  if (g_dev) {
    gk_destroy(g_dev);
    g_dev = NULL;
  }
  gk_tables_free(&g_tab);
  double ***a3[12] = {x1, x2, x3, x1c, x2c, x3c, x1ctr, x2ctr, x3ctr, dx1, dx2, dx3};
  for (int n = 0; n < 12; n++) gk_free3(a3[n]);
  x1 = x2 = x3 = x1c = x2c = x3c = x1ctr = x2ctr = x3ctr = dx1 = dx2 = dx3 = NULL;
  gk_free4(geo, NF_geo);
  gk_free4(rec, NF_rec);
  gk_free4(gem, NF_gem);
  if (gas) {
    gk_free4(gas[0], NF_gas);
    free(gas);
  }
  geo = rec = gem = NULL;
  gas = NULL;
  g_dev_fresh = 1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* host versions of the derived-variable and CFL routines
   (src/solver_func.c:255-383), used on downloaded state */
/* The host passes below are elementwise or exact min/max reductions, so
   splitting them over OpenMP threads leaves every result unchanged. */
void get_derived_variables() {
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = isg; i <= ieg; i++)
    for (int j = jsg; j <= jeg; j++)
      for (int k = ksg; k <= keg; k++) {
        gem[mags_b1][i][j][k] = gem[mag_bi][i][j][k] * (x1[i + 1][j][k] - x1ctr[i][j][k]) / dx1[i][j][k] +
                                gem[mag_bi][i + 1][j][k] * (x1ctr[i][j][k] - x1[i][j][k]) / dx1[i][j][k];
        gem[mags_b2][i][j][k] = gem[mag_bj][i][j][k] * (x2[i][j + 1][k] - x2ctr[i][j][k]) / dx2[i][j][k] +
                                gem[mag_bj][i][j + 1][k] * (x2ctr[i][j][k] - x2[i][j][k]) / dx2[i][j][k];
        gem[mags_b3][i][j][k] = gem[mag_bk][i][j][k] * (x3[i][j][k + 1] - x3ctr[i][j][k]) / dx3[i][j][k] +
                                gem[mag_bk][i][j][k + 1] * (x3ctr[i][j][k] - x3[i][j][k]) / dx3[i][j][k];
        for (int m = 0; m < 3; m++)
          gem[magtot_b1 + m][i][j][k] = gem[mags_b1 + m][i][j][k] + gem[B0_b1 + m][i][j][k];
      }
  double dmax = 0.0;
#pragma omp parallel for collapse(2) schedule(static) reduction(max : dmax)
  for (int i = is; i <= ie; i++)
    for (int j = js; j <= je; j++)
      for (int k = ks; k <= ke; k++) {
        double d = ((gem[mag_bi][i + 1][j][k] + gem[B0AfaceI_b1][i + 1][j][k]) * geo[face_idir][i + 1][j][k] -
                    (gem[mag_bi][i][j][k] + gem[B0AfaceI_b1][i][j][k]) * geo[face_idir][i][j][k]) / geo[vol_center][i][j][k] +
                   ((gem[mag_bj][i][j + 1][k] + gem[B0AfaceJ_b2][i][j + 1][k]) * geo[face_jdir][i][j + 1][k] -
                    (gem[mag_bj][i][j][k] + gem[B0AfaceJ_b2][i][j][k]) * geo[face_jdir][i][j][k]) / geo[vol_center][i][j][k] +
                   ((gem[mag_bk][i][j][k + 1] + gem[B0AfaceK_b3][i][j][k + 1]) * geo[face_kdir][i][j][k + 1] -
                    (gem[mag_bk][i][j][k] + gem[B0AfaceK_b3][i][j][k]) * geo[face_kdir][i][j][k]) / geo[vol_center][i][j][k];
        if (fabs(d) > dmax) dmax = fabs(d);
      }
  divB_max = dmax;
}

double get_dt() {
  double m = INFINITY;
  for (int i = is; i <= ie; i++)
    for (int j = js; j <= je; j++)
      for (int k = ks; k <= ke; k++) {
        double b1 = gem[magtot_b1][i][j][k], b2 = gem[magtot_b2][i][j][k], b3 = gem[magtot_b3][i][j][k];
        double rho = gas[0][gas_rho][i][j][k];
        double Va = sqrt((b1 * b1 + b2 * b2 + b3 * b3) / rho);
        double Vae = Va * CA / (sqrt(Va * Va + CA * CA));
        double v1 = gas[0][gas_v1][i][j][k], v2 = gas[0][gas_v2][i][j][k], v3 = gas[0][gas_v3][i][j][k];
        double Vf = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
        double cs = sqrt(gamma_val * gas[0][gas_p][i][j][k] / rho);
        double V = Vf + sqrt(cs * cs + Vae * Vae);
        /* get_dt_ring (src/spherical/spherical_general.c:700-735) uses the
           ring-widened azimuthal width */
        const int qk = doRingAverage == 1 ? dxk_ring : dxk;
        double d = CFL / (V / geo[dxi][i][j][k] + V / geo[dxj][i][j][k] + V / geo[qk][i][j][k]);
        m = fmin(m, d);
      }
  return m;
}

void check_positivity() {
  for (int i = is; i <= ie; i++)
    for (int j = js; j <= je; j++)
      for (int k = ks; k <= ke; k++) {
        if (gas[0][gas_rho][i][j][k] < rho_floor) {
          gas[0][gas_rho][i][j][k] = rho_floor;
          gas[0][gas_p][i][j][k] = p_floor;
        }
        if (gas[0][gas_p][i][j][k] < p_floor) gas[0][gas_p][i][j][k] = p_floor;
      }
}

/* ------------------------------------------------------------------------ */
/* device plumbing */
static int record_bcs(void) {
  memset(gk_bc_kind, 0, sizeof gk_bc_kind);
  gk_bc_record_only = 1;
  boundary_conditions();
  gk_bc_record_only = 0;
  return 0;
}

int gkh_device_open(int precision) {
  if (g_dev) {
    gk_destroy(g_dev);
    g_dev = NULL;
  }
  gk_tables_free(&g_tab);
  record_bcs();
  if (gk_tables_build(&g_tab) != 0) return GK_ERR_ARG;
  gk_grid_desc d;
  int rc = gk_fill_desc(&d, &g_tab, precision);
  if (rc != GK_OK) {
    log_error("device solver: unsupported boundary kind or background field");
    return rc;
  }
  rc = gk_create(&d, &g_dev);
  if (rc != GK_OK) return rc;
  g_precision = precision;
  const double *prim[6], *bf[3], *pp[6], *bp[3];
  for (int m = 0; m < 6; m++) {
    prim[m] = &gas[0][m][0][0][0];
    pp[m] = &gas[0][gas_rho_p + m][0][0][0];
  }
  for (int m = 0; m < 3; m++) {
    bf[m] = &gem[m][0][0][0];
    bp[m] = &gem[mag_bi_p + m][0][0][0];
  }
  /* a restart carries both time levels and dt0 (src/solver_solve.c:55-92) */
  rc = read_restart ? gk_set_state(g_dev, prim, bf, pp, bp) : gk_set_state(g_dev, prim, bf, NULL, NULL);
  if (rc != GK_OK) return rc;
  rc = gk_prologue(g_dev, read_restart ? dt0 : 0.0);
  if (rc != GK_OK) return rc;
  rc = gk_set_time(g_dev, time_sim);
  g_dev_fresh = 0;
  return rc;
}

gk_solver *gkh_device(void) { return g_dev; }

/* ------------------------------------------------------------------------ */
/* Serialised grid description.  A process that set up a problem writes its
   descriptor (scalars plus every per-axis table) into one flat buffer; any
   process can rebuild a gk_grid_desc pointing into that buffer and create a
   solver from it with the initial state it was handed.  Used to prepare
   requests in worker processes (the host facade holds one problem per
   process). */
typedef struct {
  uint32_t magic, version;
  int ni, nj, nk, coord, bc[2][6], b0_mode, N[3], ncls[3][2];
  int ring, ring_pole[2], ring_rows, ring_nchunk[16];
  double CA, b0_cell[3], b0_face[3][8], b0_edge[3][3], b0_lorentz[3];
  double rho_floor, p_floor, gravity[3];
} gk_blob_head;
#define GK_BLOB_MAGIC 0x6b426c6fu

static size_t blob_layout(const gk_blob_head *h, size_t off[GK_G_COUNT + 3 + 1 + 1 + 1]) {
  size_t n = sizeof(gk_blob_head);
  n = (n + 7) / 8 * 8;
  size_t q = 0;
  for (int g = 0; g < GK_G_COUNT; g++) {
    off[q++] = n;
    n += sizeof(double) * (size_t)(h->N[0] + h->N[1] + h->N[2]);
  }
  for (int a = 0; a < 3; a++) {
    off[q++] = n;
    n += sizeof(double) * 3 * (size_t)h->N[a];
  }
  off[q++] = n; /* recw of the three axes */
  for (int a = 0; a < 3; a++)
    n += sizeof(double) * (size_t)h->N[a] * h->ncls[a][0] * h->ncls[a][1] * GK_RECW_STRIDE;
  off[q++] = n; /* class maps */
  for (int a = 0; a < 3; a++) {
    const int tb = a == 0 ? 1 : 0, tc = a == 2 ? 1 : 2;
    n += sizeof(int) * (size_t)(h->N[tb] + h->N[tc]);
  }
  n = (n + 7) / 8 * 8;
  off[q++] = n; /* per-point background field (b0_mode 2) */
  if (h->b0_mode == 2) n += sizeof(double) * GK_B0_FIELDS * (size_t)h->N[0] * h->N[1] * h->N[2];
  return n;
}

/* Write the current problem's descriptor into buf (cap bytes); returns the
   size needed (call with buf = NULL to ask), or a negative gk_status. */
long gkh_desc_blob(void *buf, size_t cap) {
  record_bcs();
  gk_tables t;
  if (gk_tables_build(&t) != 0) return GK_ERR_ARG;
  gk_grid_desc d;
  int rc = gk_fill_desc(&d, &t, 0);
  if (rc != GK_OK) {
    gk_tables_free(&t);
    return rc;
  }
  gk_blob_head h;
  memset(&h, 0, sizeof h);
  h.magic = GK_BLOB_MAGIC;
  h.version = 4;
  h.ring = d.ring;
  memcpy(h.ring_pole, d.ring_pole, sizeof h.ring_pole);
  h.ring_rows = d.ring_rows;
  memcpy(h.ring_nchunk, d.ring_nchunk, sizeof h.ring_nchunk);
  h.ni = d.ni;
  h.nj = d.nj;
  h.nk = d.nk;
  h.coord = d.coord;
  memcpy(h.bc, d.bc, sizeof h.bc);
  h.b0_mode = d.b0_mode;
  h.N[0] = config.NI;
  h.N[1] = config.NJ;
  h.N[2] = config.NK;
  memcpy(h.ncls, d.ncls, sizeof h.ncls);
  h.CA = d.CA;
  memcpy(h.b0_cell, d.b0_cell, sizeof h.b0_cell);
  memcpy(h.b0_face, d.b0_face, sizeof h.b0_face);
  memcpy(h.b0_edge, d.b0_edge, sizeof h.b0_edge);
  memcpy(h.b0_lorentz, d.b0_lorentz, sizeof h.b0_lorentz);
  h.rho_floor = d.rho_floor;
  h.p_floor = d.p_floor;
  memcpy(h.gravity, d.gravity, sizeof h.gravity);
  size_t off[GK_G_COUNT + 6];
  const size_t need = blob_layout(&h, off);
  if (buf && cap >= need) {
    char *b = (char *)buf;
    memset(b, 0, need);
    memcpy(b, &h, sizeof h);
    size_t q = 0;
    for (int g = 0; g < GK_G_COUNT; g++, q++) {
      double *dst = (double *)(b + off[q]);
      for (int a = 0; a < 3; a++) {
        memcpy(dst, d.geom[g][a], sizeof(double) * h.N[a]);
        dst += h.N[a];
      }
    }
    for (int a = 0; a < 3; a++, q++)
      for (int c = 0; c < 3; c++)
        memcpy(b + off[q] + sizeof(double) * (size_t)c * h.N[a], d.cell_coef[a][c],
               sizeof(double) * h.N[a]);
    char *w = b + off[q++];
    for (int a = 0; a < 3; a++) {
      const size_t nw = (size_t)h.N[a] * h.ncls[a][0] * h.ncls[a][1] * GK_RECW_STRIDE;
      memcpy(w, d.recw[a], sizeof(double) * nw);
      w += sizeof(double) * nw;
    }
    char *cm = b + off[q++];
    for (int a = 0; a < 3; a++) {
      const int tb = a == 0 ? 1 : 0, tc = a == 2 ? 1 : 2;
      memcpy(cm, d.cls[a][0], sizeof(int) * h.N[tb]);
      cm += sizeof(int) * h.N[tb];
      memcpy(cm, d.cls[a][1], sizeof(int) * h.N[tc]);
      cm += sizeof(int) * h.N[tc];
    }
    if (h.b0_mode == 2) {
      const size_t n3 = (size_t)h.N[0] * h.N[1] * h.N[2];
      for (int f = 0; f < GK_B0_FIELDS; f++)
        memcpy(b + off[q] + sizeof(double) * n3 * f, d.b0_field[f], sizeof(double) * n3);
    }
  }
  gk_tables_free(&t);
  return (long)need;
}

/* Solver from a descriptor blob and an initial state (reference layout,
   current time level only) on the given HIP stream (NULL: its own);
   prologue done, time set to 0. */
int gkh_solver_from_blob(const void *blob, size_t n, int precision, const double *const prim[6],
                         const double *const bface[3], void *stream, gk_solver **out) {
  if (!blob || !out || n < sizeof(gk_blob_head)) return GK_ERR_ARG;
  *out = NULL;
  gk_blob_head h;
  memcpy(&h, blob, sizeof h);
  if (h.magic != GK_BLOB_MAGIC || h.version != 4) return GK_ERR_ARG;
  size_t off[GK_G_COUNT + 6];
  if (blob_layout(&h, off) > n) return GK_ERR_ARG;
  const char *b = (const char *)blob;
  gk_grid_desc d;
  memset(&d, 0, sizeof d);
  d.ni = h.ni;
  d.nj = h.nj;
  d.nk = h.nk;
  d.coord = h.coord;
  memcpy(d.bc, h.bc, sizeof d.bc);
  d.precision = precision;
  d.CA = h.CA;
  size_t q = 0;
  for (int g = 0; g < GK_G_COUNT; g++, q++) {
    const double *src = (const double *)(b + off[q]);
    for (int a = 0; a < 3; a++) {
      d.geom[g][a] = src;
      src += h.N[a];
    }
  }
  for (int a = 0; a < 3; a++, q++)
    for (int c = 0; c < 3; c++)
      d.cell_coef[a][c] = (const double *)(b + off[q] + sizeof(double) * (size_t)c * h.N[a]);
  const char *w = b + off[q++];
  for (int a = 0; a < 3; a++) {
    d.recw[a] = (const double *)w;
    w += sizeof(double) * (size_t)h.N[a] * h.ncls[a][0] * h.ncls[a][1] * GK_RECW_STRIDE;
  }
  const char *cm = b + off[q++];
  for (int a = 0; a < 3; a++) {
    const int tb = a == 0 ? 1 : 0, tc = a == 2 ? 1 : 2;
    d.cls[a][0] = (const int *)cm;
    cm += sizeof(int) * h.N[tb];
    d.cls[a][1] = (const int *)cm;
    cm += sizeof(int) * h.N[tc];
  }
  memcpy(d.ncls, h.ncls, sizeof d.ncls);
  d.b0_mode = h.b0_mode;
  if (h.b0_mode == 2) {
    const size_t n3 = (size_t)h.N[0] * h.N[1] * h.N[2];
    for (int f = 0; f < GK_B0_FIELDS; f++)
      d.b0_field[f] = (const double *)(b + off[q] + sizeof(double) * n3 * f);
  }
  memcpy(d.b0_cell, h.b0_cell, sizeof d.b0_cell);
  memcpy(d.b0_face, h.b0_face, sizeof d.b0_face);
  memcpy(d.b0_edge, h.b0_edge, sizeof d.b0_edge);
  memcpy(d.b0_lorentz, h.b0_lorentz, sizeof d.b0_lorentz);
  d.rho_floor = h.rho_floor;
  d.p_floor = h.p_floor;
  memcpy(d.gravity, h.gravity, sizeof d.gravity);
  d.ring = h.ring;
  memcpy(d.ring_pole, h.ring_pole, sizeof d.ring_pole);
  d.ring_rows = h.ring_rows;
  memcpy(d.ring_nchunk, h.ring_nchunk, sizeof d.ring_nchunk);
  gk_solver *s = NULL;
  d.stream = stream; /* NULL: the solver's own stream */
  int rc = gk_create(&d, &s);
  if (rc != GK_OK) return rc;
  rc = gk_set_state(s, prim, bface, NULL, NULL);
  if (rc == GK_OK) rc = gk_prologue(s, 0.0);
  if (rc == GK_OK) rc = gk_set_time(s, 0.0);
  if (rc != GK_OK) {
    gk_destroy(s);
    return rc;
  }
  *out = s;
  return GK_OK;
}

/* A standalone device solver for the problem currently held by the host
   facade (geometry, weights, B0, initial state): the caller owns it and may
   set up the next problem right away (batched ensembles, serving). */
int gkh_device_create(int precision, gk_solver **out) {
  if (!out) return GK_ERR_ARG;
  *out = NULL;
  record_bcs();
  gk_tables t;
  if (gk_tables_build(&t) != 0) return GK_ERR_ARG;
  gk_grid_desc d;
  int rc = gk_fill_desc(&d, &t, precision);
  gk_solver *s = NULL;
  if (rc == GK_OK) rc = gk_create(&d, &s);
  gk_tables_free(&t);
  if (rc != GK_OK) return rc;
  const double *prim[6], *bf[3];
  for (int m = 0; m < 6; m++) prim[m] = &gas[0][m][0][0][0];
  for (int m = 0; m < 3; m++) bf[m] = &gem[m][0][0][0];
  rc = gk_set_state(s, prim, bf, NULL, NULL);
  if (rc == GK_OK) rc = gk_prologue(s, 0.0);
  if (rc == GK_OK) rc = gk_set_time(s, time_sim);
  if (rc != GK_OK) {
    gk_destroy(s);
    return rc;
  }
  *out = s;
  return GK_OK;
}

/* cell-centred B of the previous time level (what the reference's
   X_p := X copy leaves in mags_p), for dumps */
static void host_mags_prev(void) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = isg; i <= ieg; i++)
    for (int j = jsg; j <= jeg; j++)
      for (int k = ksg; k <= keg; k++) {
        gem[mags_b1_p][i][j][k] = gem[mag_bi_p][i][j][k] * (x1[i + 1][j][k] - x1ctr[i][j][k]) / dx1[i][j][k] +
                                  gem[mag_bi_p][i + 1][j][k] * (x1ctr[i][j][k] - x1[i][j][k]) / dx1[i][j][k];
        gem[mags_b2_p][i][j][k] = gem[mag_bj_p][i][j][k] * (x2[i][j + 1][k] - x2ctr[i][j][k]) / dx2[i][j][k] +
                                  gem[mag_bj_p][i][j + 1][k] * (x2ctr[i][j][k] - x2[i][j][k]) / dx2[i][j][k];
        gem[mags_b3_p][i][j][k] = gem[mag_bk_p][i][j][k] * (x3[i][j][k + 1] - x3ctr[i][j][k]) / dx3[i][j][k] +
                                  gem[mag_bk_p][i][j][k + 1] * (x3ctr[i][j][k] - x3[i][j][k]) / dx3[i][j][k];
      }
}

int gkh_device_fetch(void);

/* device state into the host arrays; without history only the current
   time level (the initial dump happens before X_p := X in the reference) */
static int device_fetch(int history) {
  if (!g_dev) return GK_ERR_STATE;
  double *prim[6], *bf[3];
  for (int m = 0; m < 6; m++) prim[m] = &gas[0][m][0][0][0];
  for (int m = 0; m < 3; m++) bf[m] = &gem[m][0][0][0];
  int rc;
  if (history) return gkh_device_fetch();
  rc = gk_get_state(g_dev, prim, bf, NULL, NULL);
  if (rc != GK_OK) return rc;
  get_derived_variables();
  return GK_OK;
}

int gkh_device_fetch(void) {
  if (!g_dev) return GK_ERR_STATE;
  double *prim[6], *bf[3], *pp[6], *bp[3];
  for (int m = 0; m < 6; m++) {
    prim[m] = &gas[0][m][0][0][0];
    pp[m] = &gas[0][gas_rho_p + m][0][0][0];
  }
  for (int m = 0; m < 3; m++) {
    bf[m] = &gem[m][0][0][0];
    bp[m] = &gem[mag_bi_p + m][0][0][0];
  }
  int rc = gk_get_state(g_dev, prim, bf, pp, bp);
  if (rc != GK_OK) return rc;
  double sc[6];
  rc = gk_get_scalars(g_dev, sc);
  if (rc != GK_OK) return rc;
  dt = sc[0];
  dt0 = sc[1];
  time_sim = sc[2];
  divB_max = sc[3];
  get_derived_variables();
  host_mags_prev();
  return GK_OK;
}

int solve(int nt) {
  int rc;
  if (g_dev_fresh || !g_dev) {
    rc = gkh_device_open(g_precision);
    if (rc != GK_OK) {
      log_error("Error: device solver setup failed (%d).", rc);
      return -1;
    }
  }
  time_sim_start = time_sim;
  const gk_h5_api *io = gk_io_enabled ? gk_h5() : NULL;
  const int dumps = io != NULL;
  double last_dump_time = time_sim;
  if (dumps && !read_restart) {
    /* initial state after ghost fill and BCs (src/solver_solve.c:44-53) */
    hdf_seq_num = -1;
    log_seq_num = 0;
    if ((rc = device_fetch(0)) != GK_OK) {
      log_error("Error: device state download failed (%d).", rc);
      return -1;
    }
    io->dump(BASE_DAT_NAME);
    io->extra(BASE_EXT_NAME);
    last_dump_time = 0;
  }
  /* The loop of src/solver_solve.c:102-529 in runs of steps: the device
     ends a run after the first step at which a dump or status line is due
     or time_stop is passed (gk_step_until), so the host meets the checks of
     every step that could trigger one, and no other step costs a round
     trip.  Only the last step of a run can meet them. */
  double sc[6];
  if ((rc = gk_get_scalars(g_dev, sc)) != GK_OK) {
    log_error("Error: device scalars failed (%d).", rc);
    return -1;
  }
  int n = 0;
  while (n < nt) {
    const double t_dump = dumps ? output_interval * (hdf_seq_num + 1) : INFINITY;
    const double t_log = 0.1 * output_interval * log_seq_num;
    const double t_event = fmin(t_dump, t_log);
    /* steps requested: all that are left, or about what reaching the next
       stop takes at the next step's dt (with a margin), so that a stop far
       short of nt does not leave many launched steps with nothing to do;
       a run that falls short just continues in the next pass */
    int want = nt - n;
    const double t_next = fmin(t_event, time_stop);
    const double dt_next = sc[4];
    if (t_next < INFINITY && dt_next > 0.0 && dt_next < INFINITY) {
      const double est = (t_next - sc[2]) / dt_next;
      if (est < (double)want) want = (int)fmax(1.0, ceil(1.25 * est) + 2.0);
      if (want > nt - n) want = nt - n;
    }
    int done = 0;
    rc = gk_step_until(g_dev, want, time_stop, t_event, &done);
    const int bad = rc == GK_ERR_NAN;
    if (rc != GK_OK && !bad) {
      log_error("Error: device step %d failed (%d).", n + 1, rc);
      return -1;
    }
    n += done;
    if ((rc = gk_get_scalars(g_dev, sc)) != GK_OK) {
      log_error("Error: device scalars failed (%d).", rc);
      return -1;
    }
    time_sim = sc[2];
    dt = sc[0];
    divB_max = sc[3];
    if (bad) {
      log_error("Error: non-finite state after step %d (time %.15f)", n, time_sim);
      return -1;
    }
    if (dumps && time_sim >= output_interval * (hdf_seq_num + 1)) {
      if ((rc = gkh_device_fetch()) != GK_OK) {
        log_error("Error: device state download failed (%d).", rc);
        return -1;
      }
      io->dump(BASE_DAT_NAME);
      log_info("Output data at time %.15f", time_sim);
      last_dump_time = time_sim;
    }
    if (time_sim >= 0.1 * output_interval * log_seq_num) {
      log_info("Status at step %d: rank 0, Time %.15f, dt %.15f, max_divB for this rank is %.4e", n,
               time_sim, dt, divB_max);
      log_seq_num++;
    }
    if (time_sim > time_stop) {
      log_info("Step %d, Time %f, dt %f", n, time_sim, dt);
      log_info("This is the end");
      break;
    }
    if (done == 0) break; /* nothing left to step */
  }
  rc = gkh_device_fetch();
  if (rc != GK_OK) {
    log_error("Error: device state download failed (%d).", rc);
    return -1;
  }
  if (dumps && last_dump_time != time_sim) io->dump(BASE_DAT_NAME);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Single-GPU stand-ins for the reference's communicator set-up
   (src/setup_mpi.h:7,26-29): one rank owns the whole grid, so the periodic
   images come from the device ghost fill.  rank and size are the globals a
   reference main() fills from MPI_COMM_WORLD (src/main.c:27-28); a caller
   without MPI leaves them at 0 and 1. */
int rank = 0, size = 1;

int setup_comm_and_types(int rank_) {
  if (config.proc_dims[0] * config.proc_dims[1] * config.proc_dims[2] != 1) {
    log_error("the device solver runs one rank per grid (proc_dims must be 1 1 1)");
    return -1;
  }
  if (rank_ != 0 || size != 1) {
    log_error("the device solver runs one process per grid (rank %d of %d)", rank_, size);
    return -1;
  }
  config.proc_coords[0] = config.proc_coords[1] = config.proc_coords[2] = 0;
  return 0;
}

void free_mpi_types(void) {}

/* The reference main() sequence for one rank (src/main.c:30-75), without
   MPI: config file, logging, solver set-up or restart, weights, problem
   initial state (fresh start only), background field, then solve(nt) with
   the reference's dump cadence.  Arrays stay allocated for inspection until
   finalize_solver(). */
int gkh_main(const char *config_file, int nt) {
  if (gas) finalize_solver();
  gk_io_enabled = 1;
  gk_problem_kind = -1; /* taken from the environment (GK_PROBLEM, GK_SEED) */
  gk_problem_seed_set = 0;
  B10 = B20 = B30 = Field_Default;
  Source1_B0 = Source2_B0 = Source3_B0 = Field_Default;
  A1 = A2 = A3 = Field_Default;
  set_problem_config();
  /* coordinate system of the seeded problem: spherical for the dipole
     shells (kinds 2, 11), the full sphere (kind 4) and the polar cap
     (kind 5), Cartesian for the others */
  gk_coord_type = (gk_problem_kind == 2 || gk_problem_kind == 4 || gk_problem_kind == 5 ||
                   gk_problem_kind == 11) ? 2 : 0;
  if (initialize_config(0, 1, config_file) != 0) return -1;
  if (setup_comm_and_types(0) != 0) return -1;
  if (initialize_logging(BASE_LOG_NAME, 0) != 0) return -1;
  if (initialize_solver(BASE_DAT_NAME) != 0) return -2;
  getweights();
  if (!read_restart) problem_init();
  set_background_field();
  if (doRingAverage == 1) { /* src/main.c:66-70 */
    set_RingAverage_geo_data();
    set_RingAverage_config();
  }
  g_dev_fresh = 1;
  if (nt > 0 && solve(nt) != 0) return -3;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* programmatic set-up used by the Python bindings, tests and the benchmark:
   the reference main() sequence (src/main.c:30-63) without a YAML file */

static gk_tables g_probe;
void gkh_table_reset(void);
int gkh_setup(int kind, uint64_t seed, int ni, int nj, int nk, int coord, int full_b0) {
  if (gas) finalize_solver();
  gk_io_enabled = 0;
  gkh_table_reset();
  gk_problem_kind = kind;
  gk_problem_seed = seed;
  gk_problem_seed_set = 1;
  gk_coord_type = coord;
  B10 = B20 = B30 = Field_Default;
  Source1_B0 = Source2_B0 = Source3_B0 = Field_Default;
  A1 = A2 = A3 = Field_Default;
  time_sim = 0.0;
  set_problem_config();
  const int dims[3] = {1, 1, 1};
  if (gk_config_set_sizes(ni, nj, nk, dims) != 0) return -1;
  config.proc_coords[0] = config.proc_coords[1] = config.proc_coords[2] = 0;
  gk_config_finish();
  Nt = 0;
  time_stop = 1e300;
  output_interval = 1e300;
  if (initialize_solver(BASE_DAT_NAME) != 0) return -2;
  getweights();
  problem_init();
  (void)full_b0; /* the background-field arrays are always computed */
  set_background_field();
  if (gk_b0.mode < 0) return -3;
  if (doRingAverage == 1) {
    set_RingAverage_geo_data();
    set_RingAverage_config();
  }
  record_bcs();
  g_dev_fresh = 1;
  return 0;
}

/* Zero n bytes at p: the partial pages at either end are cleared, and the
   whole pages between are handed back to the kernel (MADV_DONTNEED on the
   private anonymous memory the field arrays live in), which maps zero pages
   in again when they are next touched.  Clearing them one by one instead
   made every page of every kept field fault in (0.84 s for a 128^3 grid's
   fields, where a fresh set-up pays only for the fields it writes). */
static void zero_fill(void *p, size_t n) {
  const uintptr_t pg = (uintptr_t)sysconf(_SC_PAGESIZE);
  const uintptr_t a = (uintptr_t)p, e = a + n;
  const uintptr_t h = (a + pg - 1) / pg * pg, t = e / pg * pg;
  if (t <= h || madvise((void *)h, t - h, MADV_DONTNEED) != 0) {
    memset(p, 0, n);
    return;
  }
  memset(p, 0, h - a);
  memset((void *)t, 0, e - t);
}

/* host array access for the bindings: 0 gas f, 1 gem f, 2 geo f, 3 rec f,
   4.. coordinates x1 x2 x3 x1c x2c x3c x1ctr x2ctr x3ctr dx1 dx2 dx3 */
/* New initial state for the problem already set up, with another seed: the
   grid, geometry and reconstruction weights do not depend on the seed, so
   only the problem's parameters, initial state and background field are
   redone.  Equal, array for array, to gkh_setup with the new seed. */
int gkh_reseed(uint64_t seed) {
  if (!gas || !gem) return -1;
  gk_problem_seed = seed;
  gk_problem_seed_set = 1;
  B10 = B20 = B30 = Field_Default;
  Source1_B0 = Source2_B0 = Source3_B0 = Field_Default;
  A1 = A2 = A3 = Field_Default;
  time_sim = 0.0;
  set_problem_config();
  const size_t n3 = (size_t)config.NI * config.NJ * config.NK;
  for (int f = 0; f < NF_gas; f++)
    if (gas_keep[f] && gas[0][f]) zero_fill(&gas[0][f][0][0][0], n3 * sizeof(double));
  for (int f = 0; f < NF_gem; f++)
    if (gem_keep[f] && gem[f]) zero_fill(&gem[f][0][0][0], n3 * sizeof(double));
  problem_init();
  set_background_field();
  if (gk_b0.mode < 0) return -3;
  record_bcs();
  g_dev_fresh = 1;
  return 0;
}

double *gkh_ptr(int which, int f) {
  double ***c3[12] = {x1, x2, x3, x1c, x2c, x3c, x1ctr, x2ctr, x3ctr, dx1, dx2, dx3};
  switch (which) {
    case 0: return (f >= 0 && f < NF_gas && gas && gas[0][f]) ? &gas[0][f][0][0][0] : NULL;
    case 1: return (f >= 0 && f < NF_gem && gem && gem[f]) ? &gem[f][0][0][0] : NULL;
    case 2: return (f >= 0 && f < NF_geo && geo) ? &geo[f][0][0][0] : NULL;
    case 3: return (f >= 0 && f < NF_rec && rec) ? &rec[f][0][0][0] : NULL;
    default:
      if (which >= 4 && which < 16 && c3[which - 4]) return &c3[which - 4][0][0][0];
      return NULL;
  }
}

void gkh_sizes(int out[6]) {
  out[0] = config.ni;
  out[1] = config.nj;
  out[2] = config.nk;
  out[3] = config.NI;
  out[4] = config.NJ;
  out[5] = config.NK;
}

void gkh_bc_table(int out[2][6]) { memcpy(out, gk_bc_kind, sizeof gk_bc_kind); }

double gkh_CA(void) { return CA; }
int gkh_b0_mode(void) { return gk_b0.mode; }

/* per-axis device tables, for checking them against the 3-D host geometry:
   kind 0 geometry factor (q, axis), 1 cell-centre coefficient (axis, q),
   2 reconstruction table (axis) */
const double *gkh_table(int kind, int a, int b) {
  if (!g_probe.recw[0]) gk_tables_build(&g_probe);
  switch (kind) {
    case 0: return g_probe.geom[a][b];
    case 1: return g_probe.cc[a][b];
    default: return g_probe.recw[a];
  }
}
void gkh_table_reset(void) { gk_tables_free(&g_probe); }
const int *gkh_table_cls(int a, int q, int *ncls) {
  if (!g_probe.recw[0]) gk_tables_build(&g_probe);
  *ncls = g_probe.ncls[a][q];
  return g_probe.cls[a][q];
}
