/*
 * CPU oracle for the GAMERA-OP per-step MHD update -- TEST INFRASTRUCTURE.
 *
 * A C restatement of the reference solver's time step (src/solver_solve.c and
 * src/solver_func.c), used only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker and the CPU baseline.  The
 * product path never links or calls it.
 *
 * The oracle keeps the reference's data layout: one contiguous block per field,
 * [f][i][j][k] with k fastest, NI = ni + 2*4 + 1 points per axis.  Field
 * indices follow the reference enums (src/config.h:161-499).  The caller fills
 * the geometry arrays (geo, rec, coordinates, background-field arrays) and the
 * initial primitive state, then calls go_prologue() once and go_step().
 */
#ifndef GAMERA_ORACLE_H
#define GAMERA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct go_solver go_solver;

/* array selectors for go_array() */
enum {
  GO_GAS = 0,   /* NF_gas = 82 fields */
  GO_GEM = 1,   /* NF_gem = 97 fields */
  GO_GEO = 2,   /* NF_geo = 28 fields */
  GO_REC = 3,   /* NF_rec = 42 fields */
  GO_X1 = 4, GO_X2, GO_X3,
  GO_X1C, GO_X2C, GO_X3C,
  GO_X1CTR, GO_X2CTR, GO_X3CTR,
  GO_DX1, GO_DX2, GO_DX3,
};

/* boundary kinds per (dimension, side); GO_BC_POLE only on the j sides of
   a spherical grid whose theta range ends at a pole; GO_BC_REFLECTIVE acts
   on gas only (the reference's gem reflective routines are empty) */
enum {
  GO_BC_PERIODIC = 0, GO_BC_SYMMETRIC = 1, GO_BC_NONE = 2, GO_BC_POLE = 3,
  GO_BC_REFLECTIVE = 4, GO_BC_EXTRAPOLATED = 5, GO_BC_FIXED = 6
};

/* coord: 0 Cartesian, 2 spherical (selects the geometric source terms). */
go_solver *go_create(int ni, int nj, int nk, int coord, const int bc[6],
                     double CA);
void go_destroy(go_solver *s);
double *go_array(go_solver *s, int which);
int64_t go_npoints(const go_solver *s);

/* Pole ring averaging (src/spherical/spherical_RingAverage.c): on = 1 turns
   it on (doRingAverage); pole_lo / pole_hi: theta = 0 / theta = pi is a pole
   of the grid.  Call after the geometry is loaded and before go_prologue():
   it sets geo[dxk_ring] as set_RingAverage_config does.  Returns the number
   of averaged rows per pole (NAVR), or -1 if nk is not 32, 64, 128 or 256. */
int go_set_ring(go_solver *s, int on, int pole_lo, int pole_hi);

/* solve() prologue (solver_solve.c:25-99): ghost fill, BCs, derived
   variables, X_p := X, dt0 := get_dt(). */
void go_prologue(go_solver *s);
/* n AB2 steps of the solve() loop (solver_solve.c:102-529). */
void go_step(go_solver *s, int n);

/* enable and return per-direction copies of the face fluxes of the last
   step: 27 fields [dir*9 + m], m = rho, rho v1..3, energy, entropy, magnetic
   stress 1..3; then 18 fields [27 + dir*6 + m] of the reconstructed cell B,
   left b1..b3 then right b1..b3 */
double *go_debug_fluxes(go_solver *s);

/* unit-test hooks: one PDMU7 evaluation (src/solver_func.c:215-253) and
   the Rusanov flux + magnetic stress of one face
   (src/solver_func.c:934-1515) */
double go_pdmu7(const double fh0[7], const double fp0[7], const double w[7], double faceim1,
                double facei, double vol, double dx);
void go_face_flux(int dir, const double L[6], const double R[6], const double bl[3],
                  const double br[3], const double b0[8], double Bn, double CA, double out[9]);

/* unit-test hooks of the ring average: FilterBaseMode (Q -= its m = 0, 1
   Fourier part Q0; src/solver_func.c:2187-2212) and RingBlockAvg
   (src/solver_func.c:2214-2257), N <= 256, nch <= 128 */
void go_ring_filter(double *Q, int N, double *Q0);
void go_ring_block_avg(double *Q, int N, int nch);

/* scalars: 0 dt, 1 dt0, 2 time_sim, 3 divB_max, 4 rho_floor, 5 p_floor,
   6..8 gravity G01..G03 (src/config.c:89-97; settable like the reference's
   globals) */
double go_scalar(const go_solver *s, int which);
void go_set_scalar(go_solver *s, int which, double v);

#ifdef __cplusplus
}
#endif
#endif
