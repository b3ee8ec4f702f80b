/*
 * Globals, config.yaml reader and logging of the host facade.
 *
 * Behaviour follows the reference: the flat "key: value" config.yaml schema
 * and derived sizes (src/config.c:134-194), the onface stagger tables
 * (src/config.c:196-260), grid indices (src/config.c:267-280) and physical
 * constants (src/config.c:82-106).  Unknown YAML keys are ignored, as there.
 */
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "config.h"
#include "log.h"

FILE *log_file;
int Nt;
double time_stop;
double output_interval;
config_t config;
int *proc_coords = config.proc_coords;
int pdm_index_gas = gas_rho_p - gas_rho_h;
int pdm_index_gem = mag_bi_p - mag_bi_h;
Problem_config_t problem_config;
Norm_t norm_config;
int is, ie, isg, ieg;
int js, je, jsg, jeg;
int ks, ke, ksg, keg;
double dt, dt0, dxx, dyy, dzz, dtout = 0.03, time_sim = 0., time_sim_start = 0.;
double ***x1, ***x2, ***x3;
double ***x1c, ***x2c, ***x3c;
double ***x1ctr, ***x2ctr, ***x3ctr;
double ***dx1, ***dx2, ***dx3;
double ****geo;
double ****rec;
double ****gem;
double *****gas;
int hdf_seq_num = -1;
int log_seq_num = 0;
bool read_restart = true;
int NF_gem_prim = mags_b3 + 1;
int NF_gas_prim = gas_p_S + 1;
int NF_gem_store = mags_b3_p + 1;
int NF_gas_store = gas_p_S_p + 1;
double x1min_global = 0.0, x1max_global = 1.0;
double x2min_global = 0.0, x2max_global = 1.0;
double x3min_global = 0.0, x3max_global = 1.0;
double x1min, x1max, x2min, x2max, x3min, x3max;
int geo_onface_i[NF_geo], geo_onface_j[NF_geo], geo_onface_k[NF_geo];
int gem_onface_i[NF_gem], gem_onface_j[NF_gem], gem_onface_k[NF_gem];
int gas_onface_i[NF_gas], gas_onface_j[NF_gas], gas_onface_k[NF_gas];
int doRingAverage;
int doRingAverage_thisrank_positivepole = 0;
int doRingAverage_thisrank_negativepole = 0;
int Nchunk[32];
int NAVR;

const double gamma_val = 5.0 / 3.0;
const double CFL = 0.3;
const double PDMB = 1.0;
double CA = 10.0;
const double nlf_cs = 0;
const double nlf_va = 0;
double rho_floor = 0.00316227;
double p_floor = 0.000316227;
double G01 = 0.0, G02 = 0.0, G03 = 0.0;
double divB_max = 0.0;
/* 8th-order face-average and 7-point PDM interpolation coefficients */
const double fac8th[8] = {-3.0 / 840, 29.0 / 840, -139.0 / 840, 533.0 / 840,
                          533.0 / 840, -139.0 / 840, 29.0 / 840, -3.0 / 840};
const double facPDMU7[7] = {-1.0 / 140, 5.0 / 84, -101.0 / 420, 319.0 / 420,
                            107.0 / 210, -19.0 / 210, 1.0 / 105};

/* ------------------------------------------------------------------------ */
/* logging */
static int g_log_level = LOG_TRACE;
static bool g_log_quiet = false;
static FILE *g_log_fp = NULL;
static int g_log_fp_level = LOG_TRACE;
static const char *g_level_names[] = {"TRACE", "DEBUG", "INFO", "WARN", "ERROR", "FATAL"};

void log_set_level(int level) { g_log_level = level; }
void log_set_quiet(bool enable) { g_log_quiet = enable; }
int log_add_fp(FILE *fp, int level) {
  g_log_fp = fp;
  g_log_fp_level = level;
  return 0;
}

void log_log(int level, const char *file, int line, const char *fmt, ...) {
  if (level < 0 || level > LOG_FATAL) level = LOG_FATAL;
  time_t t = time(NULL);
  struct tm tmv;
  localtime_r(&t, &tmv);
  char ts[32];
  strftime(ts, sizeof ts, "%Y-%m-%d %H:%M:%S", &tmv);
  if (!g_log_quiet && level >= g_log_level && level >= LOG_WARN) {
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "%s %-5s %s:%d: ", ts, g_level_names[level], file, line);
    vfprintf(stderr, fmt, ap);
    fputc('\n', stderr);
    va_end(ap);
  }
  if (g_log_fp && level >= g_log_fp_level) {
    va_list ap;
    va_start(ap, fmt);
    fprintf(g_log_fp, "%s %-5s %s:%d: ", ts, g_level_names[level], file, line);
    vfprintf(g_log_fp, fmt, ap);
    fputc('\n', g_log_fp);
    fflush(g_log_fp);
    va_end(ap);
  }
}

/* ------------------------------------------------------------------------ */
/* flat YAML: "key: value" per line, '#' comments, 256-character lines */
typedef struct {
  char key[64];
  char val[192];
} kv_t;

static int yaml_load(const char *path, kv_t **out, int *n) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  int cap = 32, cnt = 0;
  kv_t *kv = (kv_t *)malloc(sizeof(kv_t) * cap);
  char line[256];
  while (fgets(line, sizeof line, f)) {
    char *h = strchr(line, '#');
    if (h) *h = 0;
    char *c = strchr(line, ':');
    if (!c) continue;
    *c = 0;
    char *k = line, *v = c + 1;
    while (isspace((unsigned char)*k)) k++;
    char *ke_ = k + strlen(k);
    while (ke_ > k && isspace((unsigned char)ke_[-1])) *--ke_ = 0;
    while (isspace((unsigned char)*v)) v++;
    char *ve = v + strlen(v);
    while (ve > v && isspace((unsigned char)ve[-1])) *--ve = 0;
    if (!*k) continue;
    if (cnt == cap) {
      cap *= 2;
      kv = (kv_t *)realloc(kv, sizeof(kv_t) * cap);
    }
    snprintf(kv[cnt].key, sizeof kv[cnt].key, "%s", k);
    snprintf(kv[cnt].val, sizeof kv[cnt].val, "%s", v);
    cnt++;
  }
  fclose(f);
  *out = kv;
  *n = cnt;
  return 0;
}

static const char *yaml_find(kv_t *kv, int n, const char *key) {
  for (int q = 0; q < n; q++)
    if (strcmp(kv[q].key, key) == 0) return kv[q].val;
  return NULL;
}

static void yaml_int(kv_t *kv, int n, const char *key, int *out) {
  const char *v = yaml_find(kv, n, key);
  if (v) *out = atoi(v);
}

static void yaml_double(kv_t *kv, int n, const char *key, double *out) {
  const char *v = yaml_find(kv, n, key);
  if (v) *out = strtod(v, NULL);
}

static int verify_config(void) {
  if (config.ni_global <= 0 || config.nj_global <= 0 || config.nk_global <= 0) {
    printError("ni_global, nj_global, and nk_global must be positive.\n");
    return -1;
  }
  for (int d = 0; d < 3; d++)
    if (config.proc_dims[d] <= 0) {
      printError("proc_dims must be positive.\n");
      return -1;
    }
  if (config.ni_global % config.proc_dims[0] || config.nj_global % config.proc_dims[1] ||
      config.nk_global % config.proc_dims[2]) {
    printError("global sizes must be divisible by proc_dims.\n");
    return -1;
  }
  return 0;
}

/* derived normalisation units (src/config.c:165-171); the mu0 member is left
   at zero as the reference leaves it (restart files compare the struct) */
static void derive_norms(void) {
  double mu0 = 4.0 * PI * 1e-7;
  norm_config.Time_Norm = norm_config.x_Norm / norm_config.u_Norm;
  norm_config.p_Norm = norm_config.rho_Norm * norm_config.u_Norm * norm_config.u_Norm;
  norm_config.B_Norm = sqrt(mu0 * norm_config.rho_Norm) * norm_config.u_Norm;
  norm_config.Magnetic_moment_Norm =
      norm_config.B_Norm * pow(norm_config.x_Norm, 3) / (mu0);
  norm_config.Omega_Norm = 1.0 / norm_config.Time_Norm;
  norm_config.Sigma_Norm = 1.0 / (mu0 * norm_config.u_Norm);
}

/* sizes from global counts (src/config.c:181-186) */
int gk_config_set_sizes(int ni_global, int nj_global, int nk_global, const int dims[3]) {
  config.ni_global = ni_global;
  config.nj_global = nj_global;
  config.nk_global = nk_global;
  for (int d = 0; d < 3; d++) config.proc_dims[d] = dims[d];
  if (verify_config() != 0) return -1;
  config.ni = config.ni_global / config.proc_dims[0];
  config.nj = config.nj_global / config.proc_dims[1];
  config.nk = config.nk_global / config.proc_dims[2];
  config.NI = config.ni + 2 * NG + 1;
  config.NJ = config.nj + 2 * NG + 1;
  config.NK = config.nk + 2 * NG + 1;
  return 0;
}

int read_config(const char *filename) {
  kv_t *kv = NULL;
  int n = 0;
  if (yaml_load(filename, &kv, &n) != 0) {
    printError("Failed to parse YAML file\n");
    return 1;
  }
  int gi = 0, gj = 0, gk = 0, dims[3] = {0, 0, 0};
  yaml_int(kv, n, "ni_global", &gi);
  yaml_int(kv, n, "nj_global", &gj);
  yaml_int(kv, n, "nk_global", &gk);
  yaml_int(kv, n, "Nt", &Nt);
  yaml_double(kv, n, "time_stop", &time_stop);
  yaml_int(kv, n, "proc_dims_i", &dims[0]);
  yaml_int(kv, n, "proc_dims_j", &dims[1]);
  yaml_int(kv, n, "proc_dims_k", &dims[2]);
  yaml_double(kv, n, "output_interval", &output_interval);
  yaml_double(kv, n, "x_Norm", &norm_config.x_Norm);
  yaml_double(kv, n, "u_Norm", &norm_config.u_Norm);
  yaml_double(kv, n, "rho_Norm", &norm_config.rho_Norm);
  free(kv);
  derive_norms();
  if (gk_config_set_sizes(gi, gj, gk, dims) != 0) {
    printError("Invalid configuration.\n");
    return -1;
  }
  return 0;
}

static void set_onface_tables(void) {
  memset(geo_onface_i, 0, sizeof geo_onface_i);
  memset(geo_onface_j, 0, sizeof geo_onface_j);
  memset(geo_onface_k, 0, sizeof geo_onface_k);
  /* staggering of the geometry fields (src/config.c:196-248) */
  static const int gi[] = {edge_jdir, edge_kdir, face_idir, vol_idir, IfaceAedgeJ,
                           IfaceAedgeK, JfaceAedgeI, KfaceAedgeI};
  static const int gj[] = {edge_idir, edge_kdir, face_jdir, vol_jdir, IfaceAedgeJ,
                           JfaceAedgeK, JfaceAedgeI, KfaceAedgeJ};
  static const int gk_[] = {edge_idir, edge_jdir, face_kdir, vol_kdir, IfaceAedgeK,
                            JfaceAedgeK, KfaceAedgeI, KfaceAedgeJ};
  for (size_t q = 0; q < sizeof gi / sizeof gi[0]; q++) geo_onface_i[gi[q]] = 1;
  for (size_t q = 0; q < sizeof gj / sizeof gj[0]; q++) geo_onface_j[gj[q]] = 1;
  for (size_t q = 0; q < sizeof gk_ / sizeof gk_[0]; q++) geo_onface_k[gk_[q]] = 1;
  memset(gem_onface_i, 0, sizeof gem_onface_i);
  memset(gem_onface_j, 0, sizeof gem_onface_j);
  memset(gem_onface_k, 0, sizeof gem_onface_k);
  gem_onface_i[mag_bi] = gem_onface_i[mag0_bi] = 1;
  gem_onface_j[mag_bj] = gem_onface_j[mag0_bj] = 1;
  gem_onface_k[mag_bk] = gem_onface_k[mag0_bk] = 1;
  memset(gas_onface_i, 0, sizeof gas_onface_i);
  memset(gas_onface_j, 0, sizeof gas_onface_j);
  memset(gas_onface_k, 0, sizeof gas_onface_k);
}

void initialize_grid_indices() {
  isg = 0;
  is = isg + NO2;
  ie = is + config.ni - 1;
  ieg = ie + NO2;
  jsg = 0;
  js = jsg + NO2;
  je = js + config.nj - 1;
  jeg = je + NO2;
  ksg = 0;
  ks = ksg + NO2;
  ke = ks + config.nk - 1;
  keg = ke + NO2;
}

/* finish configuration once sizes are known (shared by the YAML and the
   programmatic paths) */
int gk_config_finish(void) {
  set_onface_tables();
  initialize_grid_indices();
  return 0;
}

int initialize_config(const int rank, const int size, const char *filename) {
  log_info("Reading config file %s for rank %d of %d processes", filename, rank, size);
  if (read_config(filename) != 0) return -1;
  if (size != config.proc_dims[0] * config.proc_dims[1] * config.proc_dims[2]) {
    printf("The number of processes %d must be equal to the product of the dimensions %d*%d*%d.\n",
           size, config.proc_dims[0], config.proc_dims[1], config.proc_dims[2]);
    return -1;
  }
  return gk_config_finish();
}

/* per-rank log file "<base>_XX-YY-ZZ_NNNNNN.log" (src/config.c:307-371) */
int initialize_logging(const char *base_filename, int rank) {
  char name[256];
  int seq = 0;
  for (;; seq++) {
    snprintf(name, sizeof name, out_file_pattern, base_filename, proc_coords[0], proc_coords[1],
             proc_coords[2], seq, "log");
    FILE *t = fopen(name, "r");
    if (!t) break;
    fclose(t);
  }
  log_file = fopen(name, "a");
  if (!log_file) {
    printError("Failed to open log file.\n");
    return -1;
  }
  log_add_fp(log_file, LOG_TRACE);
  log_set_quiet(true);
  log_info("Log initialized for rank %d", rank);
  return 0;
}
