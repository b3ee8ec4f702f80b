/*
 * Per-rank HDF5 dumps, restart and grid files of the host facade, in the
 * reference's on-disk layout (reference src/hdf5_io.c:11-926,
 * src/solver_init.c:100-349):
 *
 *   <base>_<pi>-<pj>-<pk>_<seq>.h5      (config.h: out_file_pattern)
 *   /time_sim /log_seq_num /dt0           1-element datasets
 *   /config (int) /problem_config (int) /norm_config (double) /proc_coords
 *   /x1min_global ... /x3max_global
 *   /x1 /x2 /x3                           [NI][NJ][NK]
 *   /gem                                  [18][NI][NJ][NK]  (fields 0..17)
 *   /gas                                  [1][18][NI][NJ][NK]
 *   first file also: /geo [28][..] /rec [42][..] /B0_i /B0_j /B0_k /B0_1..3
 *                    /x1c.. /x1ctr.. /dx1..
 *
 * Every dataset is deflate-6 compressed with chunks of at most 64 along each
 * axis.  /log_seq_num is stored as one H5T_NATIVE_DOUBLE whose bytes are the
 * 4-byte int counter in its low half (the reference writes the int through a
 * double-typed dataset); readers take the low four bytes.
 *
 * The device holds the state between dumps; solve() fetches it before each
 * dump.  The half-step fields (gas 6..11, gem 6..8 and 12..14) are transient
 * inside a step and are written as the host holds them (zero); a restart
 * reads only the current and previous time levels, as the reference's does.
 */
#include <dirent.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <hdf5.h>

#include "common.h"
#include "config.h"
#include "curvilinear.h"
#include "gk_host.h"
#include "log.h"

static hid_t make_dcpl(int rank, const hsize_t *dims) {
  hid_t p = H5Pcreate(H5P_DATASET_CREATE);
  if (p < 0) return -1;
  hsize_t ch[5];
  for (int a = 0; a < rank; a++) ch[a] = dims[a] < 64 ? dims[a] : 64;
  if (H5Pset_chunk(p, rank, ch) < 0 || H5Pset_deflate(p, 6) < 0) {
    H5Pclose(p);
    return -1;
  }
  return p;
}

/* one dataset of `rank` dimensions from a contiguous buffer */
static int put(hid_t f, const char *name, hid_t type, int rank, const hsize_t *dims,
               const void *data) {
  hid_t sp = H5Screate_simple(rank, dims, NULL);
  hid_t pl = make_dcpl(rank, dims);
  hid_t ds = (sp >= 0 && pl >= 0)
                 ? H5Dcreate2(f, name, type, sp, H5P_DEFAULT, pl, H5P_DEFAULT)
                 : -1;
  int rc = -1;
  if (ds >= 0 && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0) rc = 0;
  if (ds >= 0) H5Dclose(ds);
  if (pl >= 0) H5Pclose(pl);
  if (sp >= 0) H5Sclose(sp);
  if (rc) log_error("Error: Unable to write dataset %s", name);
  return rc;
}

static int put1(hid_t f, const char *name, hid_t type, const void *data, hsize_t n) {
  return put(f, name, type, 1, &n, data);
}

static int put3(hid_t f, const char *name, double ***a) {
  const hsize_t d[3] = {(hsize_t)config.NI, (hsize_t)config.NJ, (hsize_t)config.NK};
  return put(f, name, H5T_NATIVE_DOUBLE, 3, d, &a[0][0][0]);
}

/* whole dataset into a buffer of the expected number of elements */
static int get(hid_t f, const char *name, hid_t type, void *data, hsize_t expect) {
  hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
  if (ds < 0) {
    log_error("Error: dataset %s missing", name);
    return -1;
  }
  hid_t sp = H5Dget_space(ds);
  hssize_t n = sp >= 0 ? H5Sget_simple_extent_npoints(sp) : -1;
  int rc = -1;
  if (n == (hssize_t)expect && H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0) rc = 0;
  if (rc) log_error("Error: unable to read dataset %s", name);
  if (sp >= 0) H5Sclose(sp);
  H5Dclose(ds);
  return rc;
}

void get_out_filename(const char *base, char *filename, size_t size, int seq, const char *ext) {
  snprintf(filename, size, out_file_pattern, base, proc_coords[0], proc_coords[1], proc_coords[2],
           seq, ext);
}

static double seq_as_double(int v) {
  double d = 0.0;
  memcpy(&d, &v, sizeof v); /* low four bytes; the rest stay zero */
  return d;
}

int dump_data_hdf5(const char *base) {
  char fn[256];
  hdf_seq_num++;
  get_out_filename(base, fn, sizeof fn, hdf_seq_num, "h5");
  log_info("Dumping hdf5 file: %s", fn);
  hid_t f = H5Fcreate(fn, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) {
    log_error("Error: Unable to create HDF5 file %s", fn);
    return -1;
  }
  const double lsn = seq_as_double(log_seq_num);
  const hsize_t NI = config.NI, NJ = config.NJ, NK = config.NK;
  const hsize_t dgem[4] = {(hsize_t)NF_gem_store, NI, NJ, NK};
  const hsize_t dgas[5] = {1, (hsize_t)NF_gas_store, NI, NJ, NK};
  int rc = 0;
  rc |= put1(f, "/time_sim", H5T_NATIVE_DOUBLE, &time_sim, 1);
  rc |= put1(f, "/log_seq_num", H5T_NATIVE_DOUBLE, &lsn, 1);
  rc |= put1(f, "/dt0", H5T_NATIVE_DOUBLE, &dt0, 1);
  rc |= put1(f, "/config", H5T_NATIVE_INT, &config, sizeof(config_t) / sizeof(int));
  rc |= put1(f, "/problem_config", H5T_NATIVE_INT, &problem_config,
             sizeof(Problem_config_t) / sizeof(Boundary_condition_t));
  rc |= put1(f, "/norm_config", H5T_NATIVE_DOUBLE, &norm_config, sizeof(Norm_t) / sizeof(double));
  rc |= put1(f, "/proc_coords", H5T_NATIVE_INT, proc_coords, 3);
  rc |= put1(f, "/x1min_global", H5T_NATIVE_DOUBLE, &x1min_global, 1);
  rc |= put1(f, "/x1max_global", H5T_NATIVE_DOUBLE, &x1max_global, 1);
  rc |= put1(f, "/x2min_global", H5T_NATIVE_DOUBLE, &x2min_global, 1);
  rc |= put1(f, "/x2max_global", H5T_NATIVE_DOUBLE, &x2max_global, 1);
  rc |= put1(f, "/x3min_global", H5T_NATIVE_DOUBLE, &x3min_global, 1);
  rc |= put1(f, "/x3max_global", H5T_NATIVE_DOUBLE, &x3max_global, 1);
  rc |= put3(f, "/x1", x1);
  rc |= put3(f, "/x2", x2);
  rc |= put3(f, "/x3", x3);
  /* fields 0..17 of each array are one contiguous block */
  rc |= put(f, "/gem", H5T_NATIVE_DOUBLE, 4, dgem, &gem[0][0][0][0]);
  rc |= put(f, "/gas", H5T_NATIVE_DOUBLE, 5, dgas, &gas[0][0][0][0][0]);
  H5Fclose(f);
  return rc ? -1 : 0;
}

int dump_extra_data(const char *base) {
  char fn[256];
  get_out_filename(base, fn, sizeof fn, hdf_seq_num, "h5");
  hid_t f = H5Fopen(fn, H5F_ACC_RDWR, H5P_DEFAULT);
  if (f < 0) {
    log_error("Error: Unable to open HDF5 file %s", fn);
    return -1;
  }
  const hsize_t NI = config.NI, NJ = config.NJ, NK = config.NK;
  const hsize_t dgeo[4] = {NF_geo, NI, NJ, NK}, drec[4] = {NF_rec, NI, NJ, NK};
  int rc = 0;
  rc |= put(f, "/geo", H5T_NATIVE_DOUBLE, 4, dgeo, &geo[0][0][0][0]);
  if (doRingAverage == 1) {
    /* the ring copy of the face areas (src/hdf5_io.c:258-265): on one rank
       the ring arrays span the whole box, so they are geo's face fields */
    const hsize_t dring[4] = {3, NI, NJ, NK};
    rc |= put(f, "/geo_ring", H5T_NATIVE_DOUBLE, 4, dring, &geo[face_idir][0][0][0]);
  }
  rc |= put(f, "/rec", H5T_NATIVE_DOUBLE, 4, drec, &rec[0][0][0][0]);
  rc |= put3(f, "/B0_i", gem[B0AfaceI_b1]);
  rc |= put3(f, "/B0_j", gem[B0AfaceJ_b2]);
  rc |= put3(f, "/B0_k", gem[B0AfaceK_b3]);
  rc |= put3(f, "/B0_1", gem[B0_b1]);
  rc |= put3(f, "/B0_2", gem[B0_b2]);
  rc |= put3(f, "/B0_3", gem[B0_b3]);
  rc |= put3(f, "/x1c", x1c);
  rc |= put3(f, "/x2c", x2c);
  rc |= put3(f, "/x3c", x3c);
  rc |= put3(f, "/x1ctr", x1ctr);
  rc |= put3(f, "/x2ctr", x2ctr);
  rc |= put3(f, "/x3ctr", x3ctr);
  rc |= put3(f, "/dx1", dx1);
  rc |= put3(f, "/dx2", dx2);
  rc |= put3(f, "/dx3", dx3);
  H5Fclose(f);
  return rc ? -1 : 0;
}

int read_data_hdf5(const char *filename) {
  hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) {
    log_error("Error: Unable to open HDF5 file %s", filename);
    return -1;
  }
  const hsize_t n3 = (hsize_t)config.NI * config.NJ * config.NK;
  double lsn = 0.0;
  config_t cfg;
  Norm_t nrm;
  Problem_config_t pc;
  int rc = 0;
  rc |= get(f, "/time_sim", H5T_NATIVE_DOUBLE, &time_sim, 1);
  rc |= get(f, "/log_seq_num", H5T_NATIVE_DOUBLE, &lsn, 1);
  rc |= get(f, "/dt0", H5T_NATIVE_DOUBLE, &dt0, 1);
  rc |= get(f, "/config", H5T_NATIVE_INT, &cfg, sizeof(config_t) / sizeof(int));
  if (!rc && memcmp(&cfg, &config, sizeof cfg) != 0) {
    log_error("Config is not the same as the one in the file.");
    rc = -1;
  }
  if (!rc) rc |= get(f, "/problem_config", H5T_NATIVE_INT, &pc, sizeof pc / sizeof(int));
  if (!rc) problem_config = pc;
  if (!rc) rc |= get(f, "/norm_config", H5T_NATIVE_DOUBLE, &nrm, sizeof nrm / sizeof(double));
  if (!rc && memcmp(&nrm, &norm_config, sizeof nrm) != 0) {
    log_error("Norm config is not the same as the one in the file.");
    rc = -1;
  }
  if (!rc) {
    rc |= get(f, "/x1min_global", H5T_NATIVE_DOUBLE, &x1min_global, 1);
    rc |= get(f, "/x1max_global", H5T_NATIVE_DOUBLE, &x1max_global, 1);
    rc |= get(f, "/x2min_global", H5T_NATIVE_DOUBLE, &x2min_global, 1);
    rc |= get(f, "/x2max_global", H5T_NATIVE_DOUBLE, &x2max_global, 1);
    rc |= get(f, "/x3min_global", H5T_NATIVE_DOUBLE, &x3min_global, 1);
    rc |= get(f, "/x3max_global", H5T_NATIVE_DOUBLE, &x3max_global, 1);
    rc |= get(f, "/x1", H5T_NATIVE_DOUBLE, &x1[0][0][0], n3);
    rc |= get(f, "/x2", H5T_NATIVE_DOUBLE, &x2[0][0][0], n3);
    rc |= get(f, "/x3", H5T_NATIVE_DOUBLE, &x3[0][0][0], n3);
    rc |= get(f, "/gem", H5T_NATIVE_DOUBLE, &gem[0][0][0][0], (hsize_t)NF_gem_store * n3);
    rc |= get(f, "/gas", H5T_NATIVE_DOUBLE, &gas[0][0][0][0][0], (hsize_t)NF_gas_store * n3);
  }
  if (!rc) memcpy(&log_seq_num, &lsn, sizeof log_seq_num);
  H5Fclose(f);
  return rc ? -1 : 0;
}

/* lexicographically last "<base>_<pi>-<pj>-<pk>_*.h5" in the working
   directory; 1 when there is none (reference src/solver_init.c:290-330) */
static int last_dump_name(const char *base, char *out, size_t size) {
  char prefix[128];
  snprintf(prefix, sizeof prefix, out_prefix_pattern, base, proc_coords[0], proc_coords[1],
           proc_coords[2]);
  DIR *d = opendir(".");
  if (!d) return -1;
  char best[256] = {0};
  struct dirent *e;
  while ((e = readdir(d)) != NULL) {
    size_t L = strlen(e->d_name);
    if (strncmp(e->d_name, prefix, strlen(prefix)) == 0 && L > 3 &&
        strcmp(e->d_name + L - 3, ".h5") == 0 && strcmp(e->d_name, best) > 0)
      snprintf(best, sizeof best, "%s", e->d_name);
  }
  closedir(d);
  if (!best[0]) return 1;
  snprintf(out, size, "%s", best);
  return 0;
}

/* restart from the last dump of this rank: 0 read, 1 none found, <0 error */
int gk_read_restart_file(const char *base) {
  char fn[256];
  int st = last_dump_name(base, fn, sizeof fn);
  if (st != 0) {
    if (st > 0) log_warn("No restart file found for %s", base);
    return st;
  }
  log_info("Reading restart file: %s", fn);
  if (read_data_hdf5(fn) < 0) return -1;
  if (sscanf(fn, "%*[^_]_%*[^_]_%d.h5", &hdf_seq_num) != 1) return -1;
  return 0;
}

/* grid.h5: /x1grid /x2grid /x3grid of shape [nk+9][nj+9][ni+9] (i fastest)
   over the global grid; 1 when the file is absent (src/solver_init.c:100-197) */
int gk_read_grid_file(void) {
  if (access("grid.h5", F_OK) != 0) return 1;
  hid_t f = H5Fopen("grid.h5", H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  const int GI = config.ni_global + 2 * NG + 1, GJ = config.nj_global + 2 * NG + 1,
            GK = config.nk_global + 2 * NG + 1;
  const size_t n = (size_t)GI * GJ * GK;
  double *t[3];
  int rc = 0;
  const char *names[3] = {"/x1grid", "/x2grid", "/x3grid"};
  for (int a = 0; a < 3; a++) {
    t[a] = (double *)malloc(n * sizeof(double));
    if (!t[a] || get(f, names[a], H5T_NATIVE_DOUBLE, t[a], n) != 0) rc = -1;
  }
  H5Fclose(f);
  if (!rc) {
    const int oi = proc_coords[0] * config.ni, oj = proc_coords[1] * config.nj,
              ok = proc_coords[2] * config.nk;
#define G(a, k, j, i) t[a][((size_t)(k)*GJ + (j)) * GI + (i)]
    for (int i = isg; i <= ieg + 1; i++)
      for (int j = jsg; j <= jeg + 1; j++)
        for (int k = ksg; k <= keg + 1; k++) {
          x1[i][j][k] = G(0, k + ok, j + oj, i + oi);
          x2[i][j][k] = G(1, k + ok, j + oj, i + oi);
          x3[i][j][k] = G(2, k + ok, j + oj, i + oi);
        }
    x1min_global = G(0, 0, 0, NG);
    x1max_global = G(0, 0, 0, config.ni_global + NG);
    x2min_global = G(1, 0, NG, 0);
    x2max_global = G(1, 0, config.nj_global + NG, 0);
    x3min_global = G(2, NG, 0, 0);
    x3max_global = G(2, config.nk_global + NG, 0, 0);
#undef G
    x1min = x1[is][js][ks];
    x1max = x1[ie + 1][je][ke];
    x2min = x2[is][js][ks];
    x2max = x2[ie][je + 1][ke];
    x3min = x3[is][js][ks];
    x3max = x3[ie][je][ke + 1];
  }
  for (int a = 0; a < 3; a++) free(t[a]);
  return rc;
}
