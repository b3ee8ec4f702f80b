/*
 * Seeded problem generators for the GAMERA-OP per-step MHD benchmarks.
 *
 * This file is a problem plugin written against the reference's problem.h
 * template (src/problem.h:5-11): it defines set_problem_config(),
 * problem_init() and boundary_conditions() and touches only the globals that
 * template exposes (gas, gem, x1c, x1ctr, is..ke, proc_coords, CA, ...).  It
 * therefore compiles unchanged against the reference headers (to produce
 * golden fixtures from the reference CPU solver) and against this project's
 * host headers (the drop-in facade), and both builds generate bit-identical
 * initial states.
 *
 * Generators (selected by gk_problem_kind, or env GK_PROBLEM / GK_SEED):
 *   GK_SINUSOID        Cartesian periodic [0,1]^3; v=(-A1 sin(2 pi y+p1),
 *                      A2 sin(2 pi x+p2), 0); rho=25/9 (1+0.05 xi), p=5/3;
 *                      B from the edge potential
 *                      Az = B0 (cos(4 pi x+p3)/4pi + cos(2 pi y+p4)/2pi).
 *   GK_SINUSOID_BLAST  GK_SINUSOID + uniform background |B0|=1 in a seeded
 *                      isotropic direction (through the B0 path) + an
 *                      over-pressure sphere (R=0.1, p_in/p_out=100, rho=1).
 *   GK_DIPOLE_SHELL    Spherical shell, log-spaced r in [1,10],
 *                      theta in [pi/8, 7pi/8], phi periodic; B from the edge
 *                      potential A_phi = m sin(theta)/r^2; rho=1, p=0.1, plus an
 *                      over-pressure blob (R=0.5, p ratio 50).
 *   GK_SINUSOID_SPHERE GK_SINUSOID + the over-pressure sphere, no B0.
 *   GK_POLE_SPHERE     Full sphere: log-spaced r in [1,10], theta in [0, pi]
 *                      (both poles), phi periodic, pole ring averaging on
 *                      (doRingAverage = 1, as src/problem_FieldLoop_spherical.c
 *                      sets it) with the pole boundary on both theta sides;
 *                      dipole B as GK_DIPOLE_SHELL, rho = 1 + 0.1 xi, a uniform
 *                      Cartesian flow of speed 0.3 in a seeded direction (it
 *                      crosses the poles, so every pole ring carries azimuthal
 *                      structure), p = 0.1 plus an over-pressure blob (ratio 50)
 *                      centred near the north pole.
 *   GK_POLE_CAP        GK_POLE_SPHERE's state on the northern hemisphere,
 *                      theta in [0, pi/2], with the boundary layout of
 *                      src/problem_FieldLoop_spherical.c: the pole boundary at
 *                      theta = 0 only, every other ghost fixed at its initial
 *                      value (the state is set on the whole box).
 *   GK_WALL_BOX        GK_SINUSOID_SPHERE in the unit box with reflective
 *                      walls in x (face B there keeps its initial ghosts: the
 *                      reference's gem reflective routines are empty),
 *                      extrapolated (zero-gradient) boundaries in y, periodic
 *                      in z.
 *   GK_INFLOW_BOX      GK_SINUSOID_SPHERE in the unit box with a fixed
 *                      boundary at x = 0 (rho = 1, v = 0, p = 1, face B = 0
 *                      in the ghosts), extrapolated at x = 1, symmetric in y,
 *                      periodic in z.
 *   GK_ENTROPY_BOX     GK_SINUSOID_SPHERE with the entropy channel carried:
 *                      p_S = p at the start (every other seeded problem has
 *                      p_S = 0).
 *   GK_FLOOR_BOX       Cartesian periodic box whose state crosses the
 *                      positivity floors: rho = 1 + 0.05 xi, p = 0.5 outside
 *                      a seeded sphere (R = 0.2) that holds rho = 1.5 x the
 *                      density floor and p = 1.02 x the pressure floor and
 *                      expands uniformly (v = 2.5 d/R radially, d the distance
 *                      from its centre), so pressure and then density fall
 *                      below the floors within a few steps (at 24 x 20 x 12
 *                      cells both branches of the positivity check are
 *                      taken, the pressure-only one first); the sinusoid's magnetic
 *                      field.  The plugin sets non-default floors
 *                      (rho_floor = 0.01, p_floor = 0.002) and a uniform
 *                      gravity of 0.4 in a seeded direction (G01..G03), the
 *                      globals the reference leaves to problem files
 *                      (src/config.c:89-97).
 *   GK_VARYING_B0_BOX  GK_SINUSOID_SPHERE with a background field that
 *                      varies in space: B0 = b + 0.3 (sin(2 pi y + q1),
 *                      sin(2 pi z + q2), sin(2 pi x + q3)), b the seeded unit
 *                      vector of GK_SINUSOID_BLAST (divergence-free and
 *                      periodic; its stress divergence, the reference's
 *                      dmagsf_G, is not zero).
 *   GK_DIPOLE_B0_SHELL GK_DIPOLE_SHELL's grid, boundaries and state with the
 *                      dipole as the background field B0 (spherical
 *                      components B_r = 2 m cos(theta)/r^3,
 *                      B_theta = m sin(theta)/r^3) and a tenth of it as the
 *                      evolved field (edge potential A_phi = 0.1 m
 *                      sin(theta)/r^2).
 *
 * Every random draw comes from a counter-based hash of (seed, stream, global
 * i, j, k), so the state does not depend on the domain decomposition.
 * Spheres and blobs are smoothed with s = (1 - tanh((d - R)/dx))/2, dx the
 * local cell width along x1, i.e. a two-cell ramp.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "log.h"

#include "common.h"
#include "config.h"
#include "problem.h"
#include "solver.h"

enum {
  GK_SINUSOID = 0,
  GK_SINUSOID_BLAST = 1,
  GK_DIPOLE_SHELL = 2,
  GK_SINUSOID_SPHERE = 3,
  GK_POLE_SPHERE = 4,
  GK_POLE_CAP = 5,
  GK_WALL_BOX = 6,
  GK_INFLOW_BOX = 7,
  GK_ENTROPY_BOX = 8,
  GK_FLOOR_BOX = 9,
  GK_VARYING_B0_BOX = 10,
  GK_DIPOLE_B0_SHELL = 11,
};


/* Pole boundary routines exist only in spherical builds of the reference
   (src/spherical/spherical_boundary.c); weak references keep this plugin
   linkable into a Cartesian build, where GK_POLE_SPHERE is not available. */
void gas_bc_pole_j_low(void) __attribute__((weak));
void gas_bc_pole_j_high(void) __attribute__((weak));
void gem_bc_pole_j_low(void) __attribute__((weak));
void gem_bc_pole_j_high(void) __attribute__((weak));

int gk_problem_kind = -1;
uint64_t gk_problem_seed = 0;
int gk_problem_seed_set = 0;

static int gk_polar(void) {
  return gk_problem_kind == GK_POLE_SPHERE || gk_problem_kind == GK_POLE_CAP;
}

/* splitmix64 finaliser; the counter hash chains it over the key words. */
static uint64_t gk_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double gk_hash_u01(uint64_t seed, uint64_t stream, int64_t i, int64_t j,
                   int64_t k) {
  uint64_t h = gk_mix64(seed ^ gk_mix64(stream * 0xD1B54A32D192ED03ULL + 1));
  h = gk_mix64(h ^ (uint64_t)i);
  h = gk_mix64(h ^ (uint64_t)j);
  h = gk_mix64(h ^ (uint64_t)k);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

static double gk_scalar(uint64_t stream) {
  return gk_hash_u01(gk_problem_seed, stream, -1, -1, -1);
}

/* seeded parameters */
static double p_A1, p_A2, p_Bamp, p_ph[4];
static double p_b0[3];
static double p_ctr[3];
static double p_m;
static double p_q[3];  /* phases of the varying background field */

/* the spherical-shell problems (log-spaced radii, symmetric r and theta
   boundaries) */
static int gk_shell(void) {
  return gk_problem_kind == GK_DIPOLE_SHELL || gk_problem_kind == GK_DIPOLE_B0_SHELL;
}

static void gk_resolve_selection(void) {
  if (gk_problem_kind < 0) {
    const char *s = getenv("GK_PROBLEM");
    gk_problem_kind = s ? atoi(s) : GK_SINUSOID;
  }
  if (!gk_problem_seed_set) {
    const char *s = getenv("GK_SEED");
    gk_problem_seed = s ? (uint64_t)strtoull(s, NULL, 10) : 1;
    gk_problem_seed_set = 1;
  }
}

static void gk_draw_parameters(void) {
  const double twopi = 2.0 * PI;
  p_A1 = 0.8 + 0.4 * gk_scalar(1);
  p_A2 = 0.8 + 0.4 * gk_scalar(2);
  p_Bamp = 0.2 + 0.1 * gk_scalar(3);
  for (int n = 0; n < 4; n++) p_ph[n] = twopi * gk_scalar(4 + n);
  /* isotropic unit vector for the background field */
  double cz = 2.0 * gk_scalar(8) - 1.0;
  double az = twopi * gk_scalar(9);
  double sz = sqrt(fmax(0.0, 1.0 - cz * cz));
  p_b0[0] = sz * cos(az);
  p_b0[1] = sz * sin(az);
  p_b0[2] = cz;
  for (int n = 0; n < 3; n++) p_q[n] = twopi * gk_scalar(22 + n);
  if (gk_shell() || gk_polar()) {
    p_m = 0.5 + 1.5 * gk_scalar(13);
    p_ctr[0] = 3.0 + 4.0 * gk_scalar(14);               /* r */
    if (gk_polar())
      p_ctr[1] = PI * (0.05 + 0.1 * gk_scalar(15));     /* theta, near the pole */
    else
      p_ctr[1] = PI * (0.375 + 0.25 * gk_scalar(15));   /* theta */
    p_ctr[2] = twopi * gk_scalar(16);                   /* phi */
  } else {
    for (int n = 0; n < 3; n++) p_ctr[n] = gk_scalar(10 + n);
  }
}

static double gk_Az_sinusoid(double x, double y, double z) {
  (void)z;
  return p_Bamp * (cos(4.0 * PI * x + p_ph[2]) / (4.0 * PI) +
                   cos(2.0 * PI * y + p_ph[3]) / (2.0 * PI));
}

static double gk_Aphi_dipole(double r, double th, double ph) {
  (void)ph;
  return p_m * sin(th) / (r * r);
}

static double gk_Aphi_dipole_tenth(double r, double th, double ph) {
  (void)ph;
  return 0.1 * p_m * sin(th) / (r * r);
}

/* background fields that vary in space */
static double gk_B0x_var(double x, double y, double z) {
  (void)x; (void)z;
  return p_b0[0] + 0.3 * sin(2.0 * PI * y + p_q[0]);
}
static double gk_B0y_var(double x, double y, double z) {
  (void)x; (void)y;
  return p_b0[1] + 0.3 * sin(2.0 * PI * z + p_q[1]);
}
static double gk_B0z_var(double x, double y, double z) {
  (void)y; (void)z;
  return p_b0[2] + 0.3 * sin(2.0 * PI * x + p_q[2]);
}
static double gk_B0r_dipole(double r, double th, double ph) {
  (void)ph;
  return 2.0 * p_m * cos(th) / (r * r * r);
}
static double gk_B0th_dipole(double r, double th, double ph) {
  (void)ph;
  return p_m * sin(th) / (r * r * r);
}

static double gk_B0x(double x, double y, double z) { (void)x; (void)y; (void)z; return p_b0[0]; }
static double gk_B0y(double x, double y, double z) { (void)x; (void)y; (void)z; return p_b0[1]; }
static double gk_B0z(double x, double y, double z) { (void)x; (void)y; (void)z; return p_b0[2]; }

/* minimum-image distance on the unit periodic box */
static double gk_periodic_dist(double x, double y, double z) {
  double d[3] = {x - p_ctr[0], y - p_ctr[1], z - p_ctr[2]};
  double s = 0.0;
  for (int n = 0; n < 3; n++) {
    d[n] -= floor(d[n] + 0.5);
    s += d[n] * d[n];
  }
  return sqrt(s);
}

static double gk_ramp(double d, double R, double w) {
  return 0.5 * (1.0 - tanh((d - R) / w));
}

void set_problem_config() {
  gk_resolve_selection();
  gk_draw_parameters();
  CA = 10.0;
  doRingAverage = gk_polar() ? 1 : 0;
  /* floors and gravity: the reference's initial values unless the floor box
     sets its own (every kind sets them, so one process can set up several
     problems in turn) */
  rho_floor = 0.00316227;
  p_floor = 0.000316227;
  G01 = G02 = G03 = 0.0;
  if (gk_problem_kind == GK_FLOOR_BOX) {
    const double cz = 2.0 * gk_scalar(20) - 1.0, az = 2.0 * PI * gk_scalar(21);
    const double sz = sqrt(fmax(0.0, 1.0 - cz * cz));
    rho_floor = 0.01;
    p_floor = 0.002;
    G01 = 0.4 * sz * cos(az);
    G02 = 0.4 * sz * sin(az);
    G03 = 0.4 * cz;
  }
  if (gk_shell() || gk_polar()) {
    x1min_global = 1.0;
    x1max_global = 10.0;
    x2min_global = gk_polar() ? 0.0 : PI / 8.0;
    x2max_global = gk_problem_kind == GK_POLE_SPHERE ? PI
                   : gk_problem_kind == GK_POLE_CAP ? PI / 2.0 : 7.0 * PI / 8.0;
    x3min_global = 0.0;
    x3max_global = 2.0 * PI;
    problem_config.Boundary_i = BC_NON_PERIODIC;
    problem_config.Boundary_j = BC_NON_PERIODIC;
    problem_config.Boundary_k = BC_PERIODIC;
  } else {
    x1min_global = 0.0;
    x1max_global = 1.0;
    x2min_global = 0.0;
    x2max_global = 1.0;
    x3min_global = 0.0;
    x3max_global = 1.0;
    const int walls = gk_problem_kind == GK_WALL_BOX || gk_problem_kind == GK_INFLOW_BOX;
    problem_config.Boundary_i = walls ? BC_NON_PERIODIC : BC_PERIODIC;
    problem_config.Boundary_j = walls ? BC_NON_PERIODIC : BC_PERIODIC;
    problem_config.Boundary_k = BC_PERIODIC;
  }
  /* the background field is a property of the problem, not of the initial
     state: set it here so a restarted run (which skips problem_init) keeps it */
  if (gk_problem_kind == GK_SINUSOID_BLAST) {
    B10 = gk_B0x;
    B20 = gk_B0y;
    B30 = gk_B0z;
  }
  if (gk_problem_kind == GK_VARYING_B0_BOX) {
    B10 = gk_B0x_var;
    B20 = gk_B0y_var;
    B30 = gk_B0z_var;
  }
  if (gk_problem_kind == GK_DIPOLE_B0_SHELL) {
    B10 = gk_B0r_dipole;
    B20 = gk_B0th_dipole;
    B30 = Field_Default;
  }
}

/*
 * Optional grid hook: the shell problem uses log-spaced radii,
 * r = 10^(gi/ni_global) at global face index gi (ghost faces continue the
 * spacing).  Returns 1 when it rewrote x1, in which case the caller recomputes
 * the geometry arrays.
 */
int gk_problem_custom_grid(void) {
  if (!gk_shell() && !gk_polar()) return 0;
  for (int i = isg; i <= ieg + 1; i++) {
    int gi = i - is + proc_coords[0] * config.ni;
    double r = pow(10.0, (double)gi / (double)config.ni_global);
    for (int j = jsg; j <= jeg + 1; j++)
      for (int k = ksg; k <= keg + 1; k++) x1[i][j][k] = r;
  }
  return 1;
}

void boundary_conditions() {
  if (gk_shell()) {
    gas_bc_symmetric_i_low();
    gas_bc_symmetric_i_high();
    gas_bc_symmetric_j_low();
    gas_bc_symmetric_j_high();
    gem_bc_symmetric_i_low();
    gem_bc_symmetric_i_high();
    gem_bc_symmetric_j_low();
    gem_bc_symmetric_j_high();
  } else if (gk_problem_kind == GK_POLE_SPHERE && gas_bc_pole_j_low) {
    gas_bc_symmetric_i_low();
    gas_bc_symmetric_i_high();
    gas_bc_pole_j_low();
    gas_bc_pole_j_high();
    gem_bc_symmetric_i_low();
    gem_bc_symmetric_i_high();
    gem_bc_pole_j_low();
    gem_bc_pole_j_high();
  } else if (gk_problem_kind == GK_POLE_CAP && gas_bc_pole_j_low) {
    gas_bc_pole_j_low();
    gem_bc_pole_j_low();
  } else if (gk_problem_kind == GK_INFLOW_BOX) {
    gas_bc_fixed_i_low();
    gas_bc_extrapolated_i_high();
    gas_bc_symmetric_j_low();
    gas_bc_symmetric_j_high();
    gem_bc_fixed_i_low();
    gem_bc_extrapolated_i_high();
    gem_bc_symmetric_j_low();
    gem_bc_symmetric_j_high();
  } else if (gk_problem_kind == GK_WALL_BOX) {
    gas_bc_reflective_i_low();
    gas_bc_reflective_i_high();
    gas_bc_extrapolated_j_low();
    gas_bc_extrapolated_j_high();
    gem_bc_reflective_i_low();
    gem_bc_reflective_i_high();
    gem_bc_extrapolated_j_low();
    gem_bc_extrapolated_j_high();
  }
}

static void gk_init_sinusoid(int sphere) {
  const double twopi = 2.0 * PI;
  for (int s = 0; s < NS1; s++) {
    for (int i = is; i <= ie; i++) {
      for (int j = js; j <= je; j++) {
        for (int k = ks; k <= ke; k++) {
          int64_t gi = i - is + proc_coords[0] * config.ni;
          int64_t gj = j - js + proc_coords[1] * config.nj;
          int64_t gk = k - ks + proc_coords[2] * config.nk;
          double xi = 2.0 * gk_hash_u01(gk_problem_seed, 100, gi, gj, gk) - 1.0;
          double x = x1c[i][j][k], y = x2c[i][j][k], z = x3c[i][j][k];
          double rho = 25.0 / 9.0 * (1.0 + 0.05 * xi);
          double p = 5.0 / 3.0;
          if (sphere) {
            double sm = gk_ramp(gk_periodic_dist(x, y, z), 0.1, dx1[i][j][k]);
            rho = rho + (1.0 - rho) * sm;
            p = p * (1.0 + 99.0 * sm);
          }
          gas[s][gas_rho][i][j][k] = rho;
          gas[s][gas_v1][i][j][k] = -p_A1 * sin(twopi * y + p_ph[0]);
          gas[s][gas_v2][i][j][k] = p_A2 * sin(twopi * x + p_ph[1]);
          gas[s][gas_v3][i][j][k] = 0.0;
          gas[s][gas_p][i][j][k] = p;
        }
      }
    }
  }
  A3 = gk_Az_sinusoid;
  AinitB(gem[mag_bi], gem[mag_bj], gem[mag_bk], A1, A2, A3, Field_Default,
         Field_Default, Field_Default);
}

static void gk_init_dipole_shell(int b0) {
  double xc = p_ctr[0] * sin(p_ctr[1]) * cos(p_ctr[2]);
  double yc = p_ctr[0] * sin(p_ctr[1]) * sin(p_ctr[2]);
  double zc = p_ctr[0] * cos(p_ctr[1]);
  for (int s = 0; s < NS1; s++) {
    for (int i = is; i <= ie; i++) {
      for (int j = js; j <= je; j++) {
        for (int k = ks; k <= ke; k++) {
          double r = x1ctr[i][j][k], th = x2ctr[i][j][k], ph = x3ctr[i][j][k];
          double x = r * sin(th) * cos(ph) - xc;
          double y = r * sin(th) * sin(ph) - yc;
          double z = r * cos(th) - zc;
          double sm = gk_ramp(sqrt(x * x + y * y + z * z), 0.5, dx1[i][j][k]);
          gas[s][gas_rho][i][j][k] = 1.0;
          gas[s][gas_v1][i][j][k] = 0.0;
          gas[s][gas_v2][i][j][k] = 0.0;
          gas[s][gas_v3][i][j][k] = 0.0;
          gas[s][gas_p][i][j][k] = 0.1 * (1.0 + 49.0 * sm);
        }
      }
    }
  }
  A3 = b0 ? gk_Aphi_dipole_tenth : gk_Aphi_dipole;
  AinitB(gem[mag_bi], gem[mag_bj], gem[mag_bk], A1, A2, A3, Field_Default,
         Field_Default, Field_Default);
}

static void gk_init_pole_sphere(int whole_box) {
  const double twopi = 2.0 * PI;
  double xc = p_ctr[0] * sin(p_ctr[1]) * cos(p_ctr[2]);
  double yc = p_ctr[0] * sin(p_ctr[1]) * sin(p_ctr[2]);
  double zc = p_ctr[0] * cos(p_ctr[1]);
  /* flow direction: isotropic unit vector of the seeded draws 8, 9 */
  double cz = 2.0 * gk_scalar(8) - 1.0;
  double az = twopi * gk_scalar(9);
  double sz = sqrt(fmax(0.0, 1.0 - cz * cz));
  double u[3] = {0.3 * sz * cos(az), 0.3 * sz * sin(az), 0.3 * cz};
  const int i0 = whole_box ? isg : is, i1 = whole_box ? ieg : ie;
  const int j0 = whole_box ? jsg : js, j1 = whole_box ? jeg : je;
  const int k0 = whole_box ? ksg : ks, k1 = whole_box ? keg : ke;
  for (int s = 0; s < NS1; s++) {
    for (int i = i0; i <= i1; i++) {
      for (int j = j0; j <= j1; j++) {
        for (int k = k0; k <= k1; k++) {
          int64_t gi = i - is + proc_coords[0] * config.ni;
          int64_t gj = j - js + proc_coords[1] * config.nj;
          int64_t gk = k - ks + proc_coords[2] * config.nk;
          double xi = 2.0 * gk_hash_u01(gk_problem_seed, 100, gi, gj, gk) - 1.0;
          double r = x1ctr[i][j][k], th = x2ctr[i][j][k], ph = x3ctr[i][j][k];
          double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
          double x = r * st * cp - xc;
          double y = r * st * sp - yc;
          double z = r * ct - zc;
          double sm = gk_ramp(sqrt(x * x + y * y + z * z), 0.5, dx1[i][j][k]);
          gas[s][gas_rho][i][j][k] = 1.0 + 0.1 * xi;
          gas[s][gas_v1][i][j][k] = u[0] * st * cp + u[1] * st * sp + u[2] * ct;
          gas[s][gas_v2][i][j][k] = u[0] * ct * cp + u[1] * ct * sp - u[2] * st;
          gas[s][gas_v3][i][j][k] = -u[0] * sp + u[1] * cp;
          gas[s][gas_p][i][j][k] = 0.1 * (1.0 + 49.0 * sm);
        }
      }
    }
  }
  A3 = gk_Aphi_dipole;
  AinitB(gem[mag_bi], gem[mag_bj], gem[mag_bk], A1, A2, A3, Field_Default,
         Field_Default, Field_Default);
}

/* the floor box: an expanding low-density, low-pressure sphere */
static void gk_init_floor_box(void) {
  for (int s = 0; s < NS1; s++) {
    for (int i = is; i <= ie; i++) {
      for (int j = js; j <= je; j++) {
        for (int k = ks; k <= ke; k++) {
          int64_t gi = i - is + proc_coords[0] * config.ni;
          int64_t gj = j - js + proc_coords[1] * config.nj;
          int64_t gk = k - ks + proc_coords[2] * config.nk;
          double xi = 2.0 * gk_hash_u01(gk_problem_seed, 100, gi, gj, gk) - 1.0;
          double x = x1c[i][j][k], y = x2c[i][j][k], z = x3c[i][j][k];
          double d[3] = {x - p_ctr[0], y - p_ctr[1], z - p_ctr[2]};
          for (int n = 0; n < 3; n++) d[n] -= floor(d[n] + 0.5);
          double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          double sm = gk_ramp(r, 0.2, dx1[i][j][k]);
          double rho = 1.0 + 0.05 * xi;
          gas[s][gas_rho][i][j][k] = rho + (1.5 * rho_floor - rho) * sm;
          gas[s][gas_p][i][j][k] = 0.5 + (1.02 * p_floor - 0.5) * sm;
          for (int n = 0; n < 3; n++)
            gas[s][gas_v1 + n][i][j][k] = 2.5 * d[n] / 0.2 * sm;
        }
      }
    }
  }
  A3 = gk_Az_sinusoid;
  AinitB(gem[mag_bi], gem[mag_bj], gem[mag_bk], A1, A2, A3, Field_Default,
         Field_Default, Field_Default);
}

void problem_init() {
  log_info("Initializing seeded problem kind %d seed %llu", gk_problem_kind,
           (unsigned long long)gk_problem_seed);
  switch (gk_problem_kind) {
    case GK_SINUSOID:
      gk_init_sinusoid(0);
      break;
    case GK_SINUSOID_BLAST:
      gk_init_sinusoid(1);
      break;
    case GK_SINUSOID_SPHERE:
    case GK_WALL_BOX:
    case GK_INFLOW_BOX:
      gk_init_sinusoid(1);
      break;
    case GK_ENTROPY_BOX:
      gk_init_sinusoid(1);
      for (int s = 0; s < NS1; s++)
        for (int i = is; i <= ie; i++)
          for (int j = js; j <= je; j++)
            for (int k = ks; k <= ke; k++)
              gas[s][gas_p_S][i][j][k] = gas[s][gas_p][i][j][k];
      break;
    case GK_FLOOR_BOX:
      gk_init_floor_box();
      break;
    case GK_DIPOLE_SHELL:
      gk_init_dipole_shell(0);
      break;
    case GK_DIPOLE_B0_SHELL:
      gk_init_dipole_shell(1);
      break;
    case GK_VARYING_B0_BOX:
      gk_init_sinusoid(1);
      break;
    case GK_POLE_SPHERE:
      gk_init_pole_sphere(0);
      break;
    case GK_POLE_CAP:
      gk_init_pole_sphere(1);
      break;
    default:
      log_error("unknown seeded problem kind %d", gk_problem_kind);
      break;
  }
}
