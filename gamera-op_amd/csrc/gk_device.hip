// MI355X (gfx950) device solver for the GAMERA-OP per-step ideal-MHD update.
//
// One AB2 step (reference src/solver_solve.c:102-508) runs as a chain of
// stencil kernels on one HIP stream; dt, dt0 and the simulated time stay in
// device memory, so a run of steps needs no host round trip:
//
//   k_begin      dt from the fused CFL reduction, AB2 factor, dt0 := dt
//   k_sweep<D>   per D-face: PDM reconstruction of the 6 primitives and the
//                3 cell-centred B components (half-step values limited
//                against X^n), Rusanov hydro flux and magnetic stress with the
//                Boris diffusion term; also the face-averaged velocities that
//                feed the edge electric fields
//   k_emf<D>     per D-edge: second (self-limited) velocity reconstruction,
//                face-B reconstruction to the edge, EMF with the fast-speed
//                diffusion term
//   k_ct         constrained-transport face-B update (keeps div B at roundoff)
//   k_cell       flux divergence, geometric sources, conservative update,
//                positivity floors, Boris-corrected momentum; fused CFL min /
//                |div B| max / NaN reduction for the next step
//   k_ghost      periodic / symmetric ghost fill of the new state (replaces
//                the MPI halo exchange and the boundary-condition library)
//   k_mags       cell-centred B of the new state over the ghosted box
//
// Arithmetic: the FP64 path keeps the reference's floating-point operation
// order and is built without FMA contraction.  The PDM limiter branches on
// the sign of differences of neighbouring values; near symmetry planes those
// differences are pure rounding noise, so matching the reference there
// requires the same roundings, not just the same formulas.
//
// Memory: every field is one [i][j][k] block with k fastest; the k pitch is
// padded to whole 128-byte lines and every field block is placed so that the
// first active cell of each row starts a line.  Geometry
// lives in per-axis tables (orthogonal grids are separable), so per-cell HBM
// traffic is state only.  The AB2 history is kept by rotating buffers
// (X^n, X^{n-1}; three face-B and cell-B buffers) instead of copying.  The
// stencil path is memory- and FP64-VALU-bound, so there is no MFMA here.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>
#include <algorithm>
#include <mutex>

#include "gamera_op.h"

namespace {

constexpr int NGH = 4;
constexpr double GAMMA = 5.0 / 3.0;
constexpr double CFL = 0.3;
// initial values of the reference's positivity floors (src/config.c:89-90);
// a problem may change them, so the cell update takes them as arguments
constexpr double RHO_FLOOR = 0.00316227;
constexpr double P_FLOOR = 0.000316227;

struct Grid {
  int N[3];   // points per axis (n + 9)
  int s[3];   // first active index (4)
  int e[3];   // last active index
  long st[3]; // strides (i, j, k)
  long fs;    // field stride
  double CA;
};

struct Geo {
  const double *g[GK_G_COUNT][3];
  const double *cc[3][3];   // cell-centre B coefficients per axis
  const double *recw[3];
  const float *recwf[3];    // FP32 copies of recw (FP32-bulk form)
  const int *cls[3][2];
  int ncls[3][2];
  // element strides of the reconstruction tables of axis a: tfs[a] from one
  // face to the next, tps[a] from one class pair (class_b * ncls[a][1] +
  // class_c) to the next.  Axes i and j are face-major (a wave's lanes, which
  // run along k, share the face and differ in the k class); axis k is
  // class-major (the lanes differ in the face and share the classes), so a
  // wave reads contiguous rows either way.
  long tfs[3], tps[3];
};

struct B0c {
  double cell[3];
  double face[3][8];
  double edge[3][3];
  double lor[3];
};

// Background field given per point (b0_mode 2): B0A_COUNT FP64 fields of the
// solver's field stride, in the device layout.  Faces: for each direction D
// the face averages b1 b2 b3 and the face normal bn (the reference's
// B0Aface*_b1..bn, src/solver_func.c:1883-1934); edges: for each direction D
// the two components the edge field reads, (D+1)%3 then (D+2)%3
// (B0Aedge*, :1935-1980); the cell-centre field B0_b1..3; the Lorentz force
// dmagsf_G (:1981-2077).  Null when B0 is zero or the same at every point
// (b0_mode 0 / 1: the B0c constants).  Kernels that read B0 take a template
// flag B0V (per-point arrays) so the constant form keeps its registers.
enum { B0A_FACE = 0, B0A_EDGE = 12, B0A_CELL = 18, B0A_LOR = 21, B0A_COUNT = 24 };

// B0 at a face of direction D (b1 b2 b3 bn)
template <int D, typename R, bool B0V>
__device__ __forceinline__ void face_b0(const B0c &b0c, const double *b0a, long fs, long c,
                                        R b0[4]) {
  if constexpr (B0V) {
#pragma unroll
    for (int m = 0; m < 4; m++) b0[m] = (R)b0a[(B0A_FACE + 4 * D + m) * fs + c];
  } else {
#pragma unroll
    for (int m = 0; m < 4; m++) b0[m] = (R)b0c.face[D][m];
  }
}

// B0 at an edge of direction D (the three components; the one along the
// edge is not read)
template <int D, bool B0V>
__device__ __forceinline__ void edge_b0(const B0c &b0c, const double *b0a, long fs, long c,
                                        double e0[3]) {
  if constexpr (B0V) {
    e0[D] = 0.0;
    e0[(D + 1) % 3] = b0a[(B0A_EDGE + 2 * D) * fs + c];
    e0[(D + 2) % 3] = b0a[(B0A_EDGE + 2 * D + 1) * fs + c];
  } else {
#pragma unroll
    for (int m = 0; m < 3; m++) e0[m] = b0c.edge[D][m];
  }
}

// Device-resident scalars.  The fused reductions (next CFL dt, max |div B|)
// land in NSLOT slots chosen by block coordinates, so that tens of thousands
// of blocks do not all serialise on one atomic word; k_begin folds the slots.
constexpr int NSLOT = 64;
constexpr unsigned long long DBL_INF_BITS = 0x7FF0000000000000ULL;

// Run control: a call arms the grid (k_arm) with a step budget and two stop
// times; k_begin stops the grid after the step that spends the budget or
// whose time exceeds t_stop (the reference's time_stop test,
// src/solver_solve.c:523-527) or reaches t_event (a dump or status line is
// due), and every kernel of a stopped grid's later steps returns at once, so
// a run of steps needs no host round trip to end where the reference's loop
// would pause.
struct Scal {
  double dt, dt0, time, c;
  unsigned long long smin[NSLOT];  // bit images of partial minima of the next CFL dt (>= 0)
  unsigned long long smax[NSLOT];  // bit images of partial maxima of |div B|
  unsigned int nanflag;
  unsigned int halt;     // this step's kernels do nothing
  unsigned int stop;     // the next step halts
  unsigned int pad;
  long long budget;      // steps the grid may still take (< 0: no limit)
  long long steps;       // steps taken since the grid was armed
  double t_stop, t_event;
};

// every stage kernel of a halted step returns at once (null: not a step)
__device__ __forceinline__ bool halted(const Scal *sc) { return sc && sc->halt; }

__device__ __forceinline__ double geo3(const Geo &g, int q, int i, int j, int k) {
  return g.g[q][0][i] * g.g[q][1][j] * g.g[q][2][k];
}

__device__ __forceinline__ double entropy_times(double s, double rho) {
  if (s == 0.0 && rho >= 0.0 && rho < INFINITY) return s;
  return s * pow(rho, GAMMA - 1.0);
}

__device__ __forceinline__ double cfl_dt(double rho, double v1, double v2, double v3, double p,
                                         double m1, double m2, double m3, double dxi, double dxj,
                                         double dxk, double CA) {
  double bsq = m1 * m1 + m2 * m2 + m3 * m3;
  double Va = sqrt(bsq / rho);
  double Vae = Va * CA / (sqrt(Va * Va + CA * CA));
  double Vf = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
  double cs = sqrt(GAMMA * p / rho);
  double V = Vf + sqrt(cs * cs + Vae * Vae);
  return CFL / (V / dxi + V / dxj + V / dxk);
}

// cell-centred B component m from the face field (src/solver_func.c:310-322)
__device__ __forceinline__ double cell_b_of(double b_lo, double b_hi, const Geo &G, int n, int m) {
  const double *d0 = G.cc[m][0], *d1 = G.cc[m][1], *dx = G.cc[m][2];
  return b_lo * d0[n] / dx[n] + b_hi * d1[n] / dx[n];
}
__device__ __forceinline__ double cell_b(const double *Bm, const Geo &G, long c, long sm, int n,
                                         int m) {
  return cell_b_of(Bm[c], Bm[c + sm], G, n, m);
}

// block-wide min / max of non-negative doubles through their bit images
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

__device__ __forceinline__ void block_reduce(double dtc, double divb, Scal *sc, int slot) {
  __shared__ double smin[8], smax[8];
  dtc = wave_min(dtc);
  divb = wave_max(divb);
  const int lane = threadIdx.x + threadIdx.y * blockDim.x;
  const int w = lane >> 6;
  if ((lane & 63) == 0) {
    smin[w] = dtc;
    smax[w] = divb;
  }
  __syncthreads();
  if (lane == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    double mn = smin[0], mx = smax[0];
    for (int n = 1; n < nw; n++) {
      mn = fmin(mn, smin[n]);
      mx = fmax(mx, smax[n]);
    }
    atomicMin(&sc->smin[slot & (NSLOT - 1)], (unsigned long long)__double_as_longlong(mn));
    atomicMax(&sc->smax[slot & (NSLOT - 1)], (unsigned long long)__double_as_longlong(mx));
  }
}

// ---------------------------------------------------------------------------
// Start of a step (one wave per grid): fold the reduction slots into dt, set
// the AB2 factor and time, dt0 := dt, and reset the slots.
__device__ __forceinline__ void begin_step(Scal *sc) {
  const int q = threadIdx.x;  // NSLOT == 64 lanes
  if (sc->halt) return;
  if (sc->stop) {  // the previous step ended the run: this step and later ones halt
    if (q == 0) sc->halt = 1;
    return;
  }
  const double mn = wave_min(__longlong_as_double((long long)sc->smin[q]));
  sc->smin[q] = DBL_INF_BITS;
  sc->smax[q] = 0ULL;
  if (q == 0) {
    const double dt = mn;
    sc->dt = dt;
    sc->time += dt;
    sc->c = 0.5 * dt / sc->dt0;
    sc->dt0 = dt;
    sc->steps += 1;
    if (sc->budget > 0) sc->budget -= 1;
    if (sc->budget == 0 || sc->time > sc->t_stop || sc->time >= sc->t_event) sc->stop = 1;
  }
}

__global__ __launch_bounds__(64) void k_begin(Scal *sc) { begin_step(sc); }

__global__ __launch_bounds__(64) void k_begin_batch(Scal *const *scs, int n) {
  if ((int)blockIdx.x < n) begin_step(scs[blockIdx.x]);
}

// arm a grid for a run: step budget (< 0: none) and stop times
__global__ __launch_bounds__(64) void k_arm(Scal *sc, long long budget, double t_stop,
                                            double t_event) {
  if (threadIdx.x == 0) {
    sc->budget = budget;
    sc->steps = 0;
    sc->t_stop = t_stop;
    sc->t_event = t_event;
    sc->halt = 0;
    sc->stop = budget == 0 ? 1u : 0u;
  }
}

// non-finite flags of the grids of a batch (one word each)
__global__ __launch_bounds__(64) void k_flags(Scal *const *scs, int n, unsigned int *out) {
  for (int q = threadIdx.x; q < n; q += 64) out[q] = scs[q]->nanflag;
}

// dt0 := CFL dt of the initial state (prologue of a fresh start)
__global__ __launch_bounds__(64) void k_set_dt0(Scal *sc) {
  const double mn = wave_min(__longlong_as_double((long long)sc->smin[threadIdx.x]));
  if (threadIdx.x == 0) sc->dt0 = mn;
}

// Arithmetic of the reconstruction, flux and EMF kernels in real type R.
// R = double is the reference arithmetic (plain IEEE division and square
// root, reference operation order).  R = float is the FP32-bulk form: the
// same formulas with single-precision hardware reciprocal and square root
// (1 ulp each); state, geometry, CT and the dt reduction stay FP64.
__device__ __forceinline__ double qdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double qsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float qsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double qpow(double x, double y) { return pow(x, y); }
// num == 0 and qdiv(G * num, den) is a zero (so that its square root squared
// is +0): den not NaN and not 0 in FP64; in FP32, whose quotient multiplies
// by a reciprocal, den also not subnormal (its reciprocal would overflow)
__device__ __forceinline__ bool zero_ratio(double num, double den) {
  return num == 0.0 && den == den && den != 0.0;
}
__device__ __forceinline__ bool zero_ratio(float num, float den) {
  return num == 0.0f && fabsf(den) >= 1.17549435e-38f;
}
__device__ __forceinline__ float qpow(float x, float y) { return powf(x, y); }
// a * b + c: in FP64 a rounded product, then a rounded sum (the reference's
// operations; c + a * b gives the same bits); in FP32-bulk one fused
// multiply-add (one rounding)
__device__ __forceinline__ double madd(double a, double b, double c) { return a * b + c; }
__device__ __forceinline__ float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
// store of a value no later part of the same kernel reads (streaming)
template <typename T>
__device__ __forceinline__ void st_nt(T *p, T v) { __builtin_nontemporal_store(v, p); }
// load of a value no other thread of the kernel reads (streaming)
template <typename T>
__device__ __forceinline__ T ld_nt(const T *p) { return __builtin_nontemporal_load(p); }
// half-step value x + c (x - x_prev) of the AB2 predictor
template <typename R>
__device__ __forceinline__ R half_step(R x, R xp, R c) { return madd(c, x - xp, x); }

// reconstruction table of sweep axis A at face f for a point with
// coordinates idx (the two transverse coordinates select the class)
template <int A, typename R>
__device__ __forceinline__ const R *rtab(const Geo &g, int f, const int idx[3]) {
  constexpr int B = A == 0 ? 1 : 0;
  constexpr int C = A == 2 ? 1 : 2;
  const int cb = g.cls[A][0][idx[B]], ccl = g.cls[A][1][idx[C]];
  const long off = (long)f * g.tfs[A] + ((long)cb * g.ncls[A][1] + ccl) * g.tps[A];
  if constexpr (sizeof(R) == 8)
    return (const R *)(g.recw[A] + off);
  else
    return (const R *)(g.recwf[A] + off);
}

// The same table when the grid has one class on each transverse axis of A
// (Cartesian grids whose cell widths are bitwise equal along each transverse
// axis, and every spherical grid): then it depends on the face alone, and a
// march position's face is uniform across a wave.  The 18 values are read
// through the constant address space (the tables are read-only while a step
// runs), so a wave fetches them once with scalar loads into scalar registers
// instead of into 36 (FP64) vector registers of every lane.
template <int A, typename R>
__device__ __forceinline__ void utab(const Geo &g, int f, R T[GK_RECW_STRIDE]) {
  typedef const __attribute__((address_space(4))) R CR;
  const R *p = (sizeof(R) == 8 ? (const R *)g.recw[A] : (const R *)g.recwf[A]) +
               (long)f * GK_RECW_STRIDE;
  CR *c = (CR *)(const void *)p;
#pragma unroll
  for (int n = 0; n < GK_RECW_STRIDE; n++) T[n] = c[n];
}

template <typename R>
__device__ __forceinline__ R sgn(R x) {
  return (x > R(0)) ? R(1) : ((x < R(0)) ? R(-1) : R(0));
}

// PDMU7 limiter (reference src/solver_func.c:215-253) given the unlimited
// interpolant and the three limiter cells (i-1, i, i+1 in the upwind frame);
// the geometry ratio terms (q - e A_i)/(e A_i) and (e A_im1 - q)/(e A_i) are
// precomputed per face as c1, c2 (the reference forms them first as well).
template <typename R>
__device__ __forceinline__ R pdm_limit(R f_itp, R fm1, R f0, R fp1, R c1, R c2) {
  R maxf = fmax(f0, fp1), minf = fmin(f0, fp1);
  R f = fmax(minf, fmin(f_itp, maxf));
  R df0 = f0 - fm1, df1 = fp1 - f0;
  R s0 = sgn(df0), s1 = sgn(df1);
  df0 = fabs(df0);
  R q0 = fabs(s0 + s1);
  R dfl = f - f0;
  R df0c = fabs(madd(c2, fm1, c1 * f0)) / R(2);
  df0c = madd(R(0.5), df0, R(0.5) * df0c);
  // f - s1 max(0, |dfl| - q0 df0c), with x - y z written x + (-y) z
  return madd(-s1, fmax(R(0), madd(-q0, df0c, fabs(dfl))), f);
}

// Left/right states at a face from 8 stencil cells (f-4 .. f+3): xh are the
// half-step values that are interpolated, xn the values that limit.
template <typename R>
__device__ __forceinline__ void pdm_lr(const R xh[8], const R xn[8], const R *T, R &L, R &Rr) {
  R fl = R(0), fr = R(0);
#pragma unroll
  for (int n = 0; n < 7; n++) {
    fl = madd(xh[n], T[n], fl);
    fr = madd(xh[7 - n], T[7 + n], fr);
  }
  L = pdm_limit(fl, xn[2], xn[3], xn[4], T[14], T[15]);
  Rr = pdm_limit(fr, xn[5], xn[4], xn[3], T[16], T[17]);
}

// Reconstruct a stored field p (history q; self-limited when q == nullptr)
// at a face along stride sd from offset c.
template <typename R, typename S>
__device__ __forceinline__ void recon(const S *__restrict__ p, const S *__restrict__ q, long c,
                                      long sd, R cc, const R *T, R &L, R &Rr) {
  R xh[8], xn[8];
#pragma unroll
  for (int n = 0; n < 8; n++) {
    long o = c + (n - 4) * sd;
    R x = (R)p[o];
    xn[n] = x;
    xh[n] = q ? half_step(x, (R)q[o], cc) : x;
  }
  pdm_lr(xh, xn, T, L, Rr);
}

template <typename R>
__device__ __forceinline__ R entropy_over(R ps, R rho) {
  // p_S / rho^(gamma-1); p_S is identically zero unless a problem sets it, so
  // skip the pow in that case (0/x = 0 for finite positive x, as in IEEE).
  if (ps == R(0) && rho > R(0) && rho < R(INFINITY)) return ps;
  return ps / qpow(rho, R(GAMMA - 1.0));
}

// Face flux from the reconstructed states: Rusanov hydro flux (reference
// src/solver_func.c:934-1251) and magnetic stress with B0 cross terms and
// Boris diffusion (src/solver_func.c:1268-1515).  L/R: rho, v1..3, p, p_S;
// bl/br: cell-centred B (without B0); Bn: face-normal B at the face.
// Fo: mass, momentum 1..3, energy, entropy, magnetic stress 1..3.
template <int D, typename R>
__device__ __forceinline__ void face_flux(const R L[6], const R Rs[6], const R bl[3],
                                          const R br[3], const R Bn, const R b0[4],
                                          const double CAd, R Fo[9]) {
  const R CA = (R)CAd;
  const R G = (R)GAMMA;
  const R rl = L[0], rr = Rs[0], pl = L[4], pr = Rs[4], psl = L[5], psr = Rs[5];
  const R vl[3] = {L[1], L[2], L[3]}, vr[3] = {Rs[1], Rs[2], Rs[3]};
  R csl = qsqrt(qdiv(G * pl, rl)), csr = qsqrt(qdiv(G * pr, rr));
  // entropy sound speeds, used only squared: with p_S = 0 on both sides (every
  // problem that does not set it) each square is +0 exactly, and a wave whose
  // lanes all have it skips the divisions and square roots
  R cel2 = R(0), cer2 = R(0);
  if (!(zero_ratio(psl, rl) && zero_ratio(psr, rr))) {
    const R cel = qsqrt(qdiv(G * psl, rl)), cer = qsqrt(qdiv(G * psr, rr));
    cel2 = cel * cel;
    cer2 = cer * cer;
  }
  // sums of products are written with madd in the reference's order:
  // a + b * c as madd(b, c, a), a - b * c as madd(-b, c, a)
  const R tl[3] = {bl[0] + b0[0], bl[1] + b0[1], bl[2] + b0[2]};
  const R tr[3] = {br[0] + b0[0], br[1] + b0[1], br[2] + b0[2]};
  R Btl = madd(tl[2], tl[2], madd(tl[1], tl[1], tl[0] * tl[0]));
  R Btr = madd(tr[2], tr[2], madd(tr[1], tr[1], tr[0] * tr[0]));
  R Val = qsqrt(qdiv(Btl, rl)), Var = qsqrt(qdiv(Btr, rr));
  R Vel = qdiv(Val * CA, qsqrt(madd(Val, Val, CA * CA)));
  R Ver = qdiv(Var * CA, qsqrt(madd(Var, Var, CA * CA)));
  R vnl = vl[D], vnr = vr[D];
  R VF = fmax(fabs(vnl) + qsqrt(madd(Vel, Vel, csl * csl)),
              fabs(vnr) + qsqrt(madd(Ver, Ver, csr * csr)));
  R VFe = fmax(fabs(vnl) + qsqrt(madd(Vel, Vel, cel2)), fabs(vnr) + qsqrt(madd(Ver, Ver, cer2)));
  R rvl[3], rvr[3];
#pragma unroll
  for (int m = 0; m < 3; m++) {
    rvl[m] = rl * vl[m];
    rvr[m] = rr * vr[m];
  }
  const R h = R(0.5);
  const R v2l = madd(vl[2], vl[2], madd(vl[1], vl[1], vl[0] * vl[0]));
  const R v2r = madd(vr[2], vr[2], madd(vr[1], vr[1], vr[0] * vr[0]));
  R engl = madd(h * rl, v2l, qdiv(pl, R(GAMMA - 1.0)));
  R engr = madd(h * rr, v2r, qdiv(pr, R(GAMMA - 1.0)));
  R Sl = entropy_over(psl, rl), Sr = entropy_over(psr, rr);
  Fo[0] = h * madd(-VF, rr - rl, rvl[D] + rvr[D]);
#pragma unroll
  for (int m = 0; m < 3; m++) {
    if (m == D)
      Fo[1 + m] = h * madd(-VF, rvr[m] - rvl[m], madd(rvr[m], vnr, madd(rvl[m], vnl, pl)) + pr);
    else
      Fo[1 + m] = h * madd(-VF, rvr[m] - rvl[m], madd(rvr[m], vnr, rvl[m] * vnl));
  }
  Fo[4] = h * madd(-VF, engr - engl, madd(pr, vnr, madd(engr, vnr, madd(pl, vnl, engl * vnl))));
  Fo[5] = h * madd(-VFe, Sr - Sl, madd(Sr, vnr, Sl * vnl));

  const R b0n = b0[3];
  R bsql = madd(R(2), madd(bl[2], b0[2], madd(bl[1], b0[1], bl[0] * b0[0])),
                madd(bl[2], bl[2], madd(bl[1], bl[1], bl[0] * bl[0])));
  R bsqr = madd(R(2), madd(br[2], b0[2], madd(br[1], b0[1], br[0] * b0[0])),
                madd(br[2], br[2], madd(br[1], br[1], br[0] * br[0])));
  R Bsq = qdiv(qdiv(Btl + Btr, R(4)), CA);
#pragma unroll
  for (int m = 0; m < 3; m++) {
    R sp, sn;
    if (m == D) {
      sp = h * madd(-bl[m], b0n, madd(-tl[m], Bn, h * bsql));
      sn = h * madd(-br[m], b0n, madd(-tr[m], Bn, h * bsqr));
    } else {
      sp = h * madd(-bl[m], b0n, -tl[m] * Bn);
      sn = h * madd(-br[m], b0n, -tr[m] * Bn);
    }
    R ms = sp + sn;
    Fo[6 + m] = madd(-Bsq, vr[m] - vl[m], ms);
  }
}

// XCD-aware block order.  Workgroups of a launch are dealt to the 8 XCDs in
// turn and every XCD has its own L2; a 1-D launch is therefore decoded so that
// each XCD receives a contiguous range of the logical blocks (logical order
// z fastest when ZF, x fastest otherwise).  Neighbouring blocks, which share
// stencil halos, chunk prologues and faces, then meet in the same L2.
constexpr int NXCD = 8;

template <bool ZF>
__device__ __forceinline__ bool xcd_block(int gx, int gy, int gz, int &bx, int &by, int &bz) {
  const int total = gx * gy * gz;
  const int per = (total + NXCD - 1) / NXCD;
  const int b = blockIdx.x;
  const int L = (b % NXCD) * per + b / NXCD;
  if (L >= total) return false;
  if (ZF) {
    bz = L % gz;
    const int t = L / gz;
    bx = t % gx;
    by = t / gx;
  } else {
    bx = L % gx;
    const int t = L / gx;
    by = t % gy;
    bz = t / gy;
  }
  return true;
}

// xcd_block for block b of a launch part whose first block is a multiple of
// NXCD (the hardware deals block b to XCD b % NXCD as it would block
// blockIdx.x = b)
template <bool ZF>
__device__ __forceinline__ bool xcd_block_at(int b, int gx, int gy, int gz, int &bx, int &by,
                                             int &bz) {
  const int total = gx * gy * gz;
  const int per = (total + NXCD - 1) / NXCD;
  const int L = (b % NXCD) * per + b / NXCD;
  if (L >= total) return false;
  if (ZF) {
    bz = L % gz;
    const int t = L / gz;
    bx = t % gx;
    by = t / gx;
  } else {
    bx = L % gx;
    const int t = L / gx;
    by = t % gy;
    bz = t / gy;
  }
  return true;
}

// logical block of a launch of `total` blocks in the order of xcd_block, for
// launches that split their blocks between several parts
__device__ __forceinline__ bool xcd_linear(int total, int &L) {
  const int per = (total + NXCD - 1) / NXCD;
  const int b = blockIdx.x;
  L = (b % NXCD) * per + b / NXCD;
  return L < total;
}

// XCD-aware order in 8 x 8 patches of (x, z) for launches with gy == 1
// whose blocks read faces of both neighbours, x + 1 (the next row tile) and
// z + 1 (the next plane): inside a patch the two neighbours of a block run
// within 8 blocks of it on the same XCD, so the shared faces are still in
// its L2.  The launch covers the grid rounded up to whole patches
// (patch_blocks); blocks outside the grid return at once.
constexpr int PATCH = 8;

__device__ __forceinline__ bool xcd_patch(int gx, int gz, int &bx, int &bz) {
  const int npz = (gz + PATCH - 1) / PATCH;
  const int total = (gx + PATCH - 1) / PATCH * PATCH * npz * PATCH;
  const int per = (total + NXCD - 1) / NXCD;
  const int b = blockIdx.x;
  const int L = (b % NXCD) * per + b / NXCD;
  if (L >= total) return false;
  const int pp = L / (PATCH * PATCH), w = L % (PATCH * PATCH);
  bz = pp % npz * PATCH + w % PATCH;
  bx = pp / npz * PATCH + w / PATCH;
  return bx < gx && bz < gz;
}

// work-list entry of a batched launch: argument set and block coordinates.
// A batched launch has one block per entry; consecutive entries (neighbouring
// blocks of one grid) go to the same XCD, as in xcd_block.
__device__ __forceinline__ bool xcd_item(int nwork, int &w) {
  const int per = (nwork + NXCD - 1) / NXCD;
  const int b = blockIdx.x;
  w = (b % NXCD) * per + b / NXCD;
  return w < nwork;
}

struct WorkItem {
  int arg, bx, by, bz;
};

// A batched launch reads each grid's argument set from global memory, so the
// compiler cannot tell that the pointers in it address global memory and
// would access them with flat instructions (slower to issue, and counted
// against the LDS counter too, so waits on scalar and LDS loads also wait for
// them).  A batched kernel copies its set and marks every pointer in it as a
// global-memory pointer (an integer round trip, which the optimiser does not
// fold away), as it does by itself for a kernel argument's pointers.
template <typename T>
__device__ __forceinline__ void gptr(T *&p) {
  p = (T *)(__attribute__((address_space(1))) T *)(unsigned long)p;
}

__device__ __forceinline__ void globalize(Geo &g) {
#pragma unroll
  for (int q = 0; q < GK_G_COUNT; q++)
#pragma unroll
    for (int a = 0; a < 3; a++) gptr(g.g[q][a]);
#pragma unroll
  for (int a = 0; a < 3; a++) {
#pragma unroll
    for (int b = 0; b < 3; b++) gptr(g.cc[a][b]);
    gptr(g.recw[a]);
    gptr(g.recwf[a]);
    gptr(g.cls[a][0]);
    gptr(g.cls[a][1]);
  }
}

template <typename T>
__device__ __forceinline__ T batch_args(const T *args, int arg) {
  T a = args[arg];
  globalize(a);
  return a;
}

// Marching kernels split the march range [lo, hi] into nch near-equal chunks
// (lengths differ by at most one), one per thread; chunk bz is [f0, f1).
// Equal shares avoid a nearly empty last chunk (257 faces at 32 per thread
// had left a one-face chunk in every column).
__device__ __forceinline__ void march_range(int lo, int hi, int nch, int bz, int &f0, int &f1) {
  const long nf = (long)hi - lo + 1;
  f0 = lo + (int)(bz * nf / nch);
  f1 = lo + (int)((bz + 1) * nf / nch);
}

// chunks for a march range of nf points at about `per` points per thread
static inline int march_chunks(int nf, int per) {
  return std::max(1, std::min(nf, (nf + per / 2) / per));
}

// ---------------------------------------------------------------------------
struct SweepArgs {
  Grid g;
  Geo geo;
  B0c b0;
  const double *b0a;     // per-point B0 (b0_mode 2), else null
  // state read by the stencil: of type R (FP64 state, or its FP32 copy in
  // the FP32-bulk form)
  const void *X, *Xp;    // 6 fields
  const void *M, *Mp;    // 3 cell-B fields at n and n-1
  const void *B;         // face B at n (face-normal component for the stress)
  void *F;               // 9 fields of R: 6 hydro + 3 magnetic stress
  void *V;               // 2 fields of R: face-averaged velocity components
  const Scal *sc;
  int lo[3], hi[3];
  int flo[3], fhi[3];    // range where the full flux is needed
  int nch;               // marching sweeps: near-equal chunks of the march range
  int gap;               // velocity strips: strip-axis indices from NGH on are
                         // shifted by gap (both ghost strips in one box)
};

__device__ __forceinline__ void globalize(SweepArgs &a) {
  globalize(a.geo);
  gptr(a.b0a);
  gptr(a.X);
  gptr(a.Xp);
  gptr(a.M);
  gptr(a.Mp);
  gptr(a.B);
  gptr(a.F);
  gptr(a.V);
  gptr(a.sc);
}

// Direct sweep for the velocity-only strips of the j sweep (k in the four
// ghost layers on either side).  Each thread reads its eight stencil cells;
// the points of the launch box are walked k fastest, then j, then i, 256 per
// block.
template <int D, typename R, bool B0V>
__device__ __forceinline__ void k_sweep_body(const SweepArgs &a, const int bx, const int by,
                                        const int bz) {
  if (halted(a.sc)) return;
  (void)by;
  (void)bz;
  const int W = a.hi[2] - a.lo[2] + 1, H = a.hi[1] - a.lo[1] + 1;
  const int p = bx * 256 + threadIdx.y * 64 + threadIdx.x;
  const int t = p / W;
  int k = a.lo[2] + (p - t * W);
  int j = a.lo[1] + t % H;
  const int i = a.lo[0] + t / H;
  if (i > a.hi[0]) return;
  if (D == 0 && j >= NGH) j += a.gap;
  if (D == 1 && k >= NGH) k += a.gap;
  const Grid &g = a.g;
  const long c = i * g.st[0] + j * g.st[1] + k;
  const long sd = g.st[D];
  const int idx[3] = {i, j, k};
  const int f = idx[D];
  const bool full = i >= a.flo[0] && i <= a.fhi[0] && j >= a.flo[1] && j <= a.fhi[1] &&
                    k >= a.flo[2] && k <= a.fhi[2];
  const R cc = (R)a.sc->c;
  const R *T = rtab<D, R>(a.geo, f, idx);
  R *F = (R *)a.F, *V = (R *)a.V;

  // velocity components that feed the edge fields: D=0 (v1,v2), 1 (v2,v3), 2 (v3,v1)
  constexpr int va = D == 0 ? 1 : (D == 1 ? 2 : 3);
  constexpr int vb = D == 0 ? 2 : (D == 1 ? 3 : 1);
  R L[6], Rs[6];
#pragma unroll
  for (int m = 0; m < 6; m++) {
    if (!full && m != va && m != vb) continue;
    recon((const R *)a.X + m * g.fs, (const R *)a.Xp + m * g.fs, c, sd, cc, T, L[m], Rs[m]);
  }
  V[c] = R(0.5) * (L[va] + Rs[va]);
  V[g.fs + c] = R(0.5) * (L[vb] + Rs[vb]);
  if (!full) return;

  R bl[3], br[3];
#pragma unroll
  for (int m = 0; m < 3; m++)
    recon((const R *)a.M + m * g.fs, (const R *)a.Mp + m * g.fs, c, sd, cc, T, bl[m], br[m]);

  R Fo[9], b0[4];
  face_b0<D, R, B0V>(a.b0, a.b0a, g.fs, c, b0);
  face_flux<D, R>(L, Rs, bl, br, ((const R *)a.B)[D * g.fs + c], b0, g.CA, Fo);
#pragma unroll
  for (int m = 0; m < 9; m++) F[m * g.fs + c] = Fo[m];
}

template <int D, typename R, bool B0V>
__global__ __launch_bounds__(256) void k_sweep(SweepArgs a, int gx, int gy, int gz) {
  int bx, by, bz;
  if (!xcd_block<false>(gx, gy, gz, bx, by, bz)) return;
  k_sweep_body<D, R, B0V>(a, bx, by, bz);
}

// one launch for many grids: block w runs work-list entry w (argument set, block)
template <int D, typename R, bool B0V>
__global__ __launch_bounds__(256) void k_sweep_batch(const SweepArgs *__restrict__ args,
                                                    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  // uniform per block: keep the argument set in scalar registers
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  const int bx = __builtin_amdgcn_readfirstlane(work[w].bx);
  const int by = __builtin_amdgcn_readfirstlane(work[w].by);
  const int bz = __builtin_amdgcn_readfirstlane(work[w].bz);
  k_sweep_body<D, R, B0V>(batch_args(args, arg), bx, by, bz);
}

// Marching sweep along D = 0 (i) or 1 (j).  Lanes run along k (coalesced),
// block rows along the other transverse axis RA, and each thread walks a
// chunk of faces along D, keeping for every reconstructed field a register
// window of the 8 half-step values and the 6 current values the stencil and
// limiter need: each face loads one new cell instead of eight.  Rows outside
// [flo, fhi] of RA only form the two face-averaged velocities.  The per-face
// arithmetic is that of k_sweep.
// k faces of a few k columns (those left over after the 64-face tiles of
// k_sweepk): a block is one wave on one column and 64 rows (j fastest, then
// i).  A lane's stencil lies along k inside its own row, so each field is
// loaded with 8 lanes per row (8 rows x 64 contiguous bytes per load) and
// transposed through LDS, instead of every lane loading its own 8 cells (64
// rows per load).  Rows of planes outside [flo0, fhi0] only store the face
// velocities.  The arithmetic is that of k_sweepk.
constexpr int KCOL_LDS = 4 * 64 * 9;  // elements of k_kcol's LDS
template <typename R, bool B0V>
__device__ __forceinline__ void k_kcol_body(const SweepArgs &a, const int bx, const int by,
                                       const int bz, R *lds) {
  if (halted(a.sc)) return;
  (void)by;
  (void)bz;
  R(*sx)[9] = (R(*)[9])lds;  // [row][cell], padded against bank conflicts
  R(*sh)[9] = (R(*)[9])(lds + 64 * 9);
  const Grid &g = a.g;
  const int W = a.hi[2] - a.lo[2] + 1, H = a.hi[1] - a.lo[1] + 1;
  const int nrow = H * (a.hi[0] - a.lo[0] + 1);
  const int col = bx % W, grp = bx / W;
  const int k = a.lo[2] + col;
  const int tx = threadIdx.x;
  const long fs = g.fs;
  const R cc = (R)a.sc->c;
  auto rowbase = [&](int q) {
    q = min(q, nrow - 1);
    return (long)(a.lo[0] + q / H) * g.st[0] + (long)(a.lo[1] + q % H) * g.st[1];
  };
  // loads: lane tx brings cell k-4+(tx%8) of row 8s + tx/8, s = 0..7
  long lo_off[8];
#pragma unroll
  for (int s2 = 0; s2 < 8; s2++) lo_off[s2] = rowbase(grp * 64 + 8 * s2 + tx / 8) + k - 4 + tx % 8;
  const R *X = (const R *)a.X, *Xp = (const R *)a.Xp;
  const R *M = (const R *)a.M, *Mp = (const R *)a.Mp;
  const int q = grp * 64 + tx;
  const bool valid = q < nrow;
  const int i = a.lo[0] + min(q, nrow - 1) / H, j = a.lo[1] + min(q, nrow - 1) % H;
  int idx[3] = {i, j, k};
  const R *T = rtab<2, R>(a.geo, k, idx);
  const long c = (long)i * g.st[0] + (long)j * g.st[1] + k;
  const bool full = i >= a.flo[0] && i <= a.fhi[0];
  // one field at a time (a loop, not unrolled, so the loads of later fields
  // are not hoisted into registers); left/right states wait in LDS
  R(*sl)[64] = (R(*)[64])(lds + 2 * 64 * 9);
  R(*sr)[64] = (R(*)[64])(lds + 3 * 64 * 9);
#pragma unroll 1
  for (int m = 0; m < 9; m++) {
    const R *Pm = m < 6 ? X + m * fs : M + (m - 6) * fs;
    const R *Qm = m < 6 ? Xp + m * fs : Mp + (m - 6) * fs;
    R xv[8], qv[8];
#pragma unroll
    for (int s2 = 0; s2 < 8; s2++) {
      xv[s2] = Pm[lo_off[s2]];
      qv[s2] = Qm[lo_off[s2]];
    }
#pragma unroll
    for (int s2 = 0; s2 < 8; s2++) {
      sx[8 * s2 + tx / 8][tx % 8] = xv[s2];
      sh[8 * s2 + tx / 8][tx % 8] = half_step(xv[s2], qv[s2], cc);
    }
    __syncthreads();
    R xn[8], xh[8];
#pragma unroll
    for (int n = 0; n < 8; n++) {
      xn[n] = sx[tx][n];
      xh[n] = sh[tx][n];
    }
    R l, r;
    pdm_lr(xh, xn, T, l, r);
    sl[m][tx] = l;
    sr[m][tx] = r;
    __syncthreads();  // the next field is staged over this one
  }
  R L[6], Rs[6], bl[3], br[3];
#pragma unroll
  for (int m = 0; m < 6; m++) {
    L[m] = sl[m][tx];
    Rs[m] = sr[m][tx];
  }
#pragma unroll
  for (int m = 0; m < 3; m++) {
    bl[m] = sl[6 + m][tx];
    br[m] = sr[6 + m][tx];
  }
  if (!valid) return;
  R *F = (R *)a.F, *V = (R *)a.V;
  st_nt(V + c, R(0.5) * (L[3] + Rs[3]));
  st_nt(V + fs + c, R(0.5) * (L[1] + Rs[1]));
  if (!full) return;
  R Fo[9], b0[4];
  face_b0<2, R, B0V>(a.b0, a.b0a, fs, c, b0);
  face_flux<2, R>(L, Rs, bl, br, ((const R *)a.B)[2 * fs + c], b0, g.CA, Fo);
#pragma unroll
  for (int m = 0; m < 9; m++) st_nt(F + m * fs + c, Fo[m]);
}

template <typename R, bool B0V>
__global__ __launch_bounds__(64) void k_kcol(SweepArgs a, int gx, int gy, int gz) {
  int bx, by, bz;
  if (!xcd_block<false>(gx, gy, gz, bx, by, bz)) return;
  __shared__ R lds[KCOL_LDS];
  k_kcol_body<R, B0V>(a, bx, by, bz, lds);
}

template <typename R, bool B0V>
__global__ __launch_bounds__(64) void k_kcol_batch(const SweepArgs *__restrict__ args,
                                                  const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  const int bx = __builtin_amdgcn_readfirstlane(work[w].bx);
  const int by = __builtin_amdgcn_readfirstlane(work[w].by);
  const int bz = __builtin_amdgcn_readfirstlane(work[w].bz);
  __shared__ R lds[KCOL_LDS];
  k_kcol_body<R, B0V>(batch_args(args, arg), bx, by, bz, lds);
}

// Points of a marching launch: the (row, k) plane of the launch box is walked
// row by row, k fastest, 256 points per block (bx), so a k extent that is not
// a multiple of 64 (257 faces or edges, small grids) does not leave waves
// mostly empty; the lanes of a wave may span two rows.
__device__ __forceinline__ bool row_point(int klo, int khi, int rlo, int rhi, int bx, int &k,
                                          int &r) {
  const int wk = khi - klo + 1;
  const int p = bx * 256 + threadIdx.y * 64 + threadIdx.x;
  const int q = p / wk;
  r = rlo + q;
  k = klo + (p - q * wk);
  return r <= rhi;
}

template <int NF, typename R>
struct Window {
  R h[NF][8];  // half-step values at cells f-4 .. f+3
  R x[NF][6];  // current values at cells f-2 .. f+3
};

template <int NF, typename R, typename S>
__device__ __forceinline__ void march_load(Window<NF, R> &w, const S *const P[NF],
                                           const S *const Q[NF], long o, R cc, int slot) {
#pragma unroll
  for (int m = 0; m < NF; m++) {
    const R x = (R)P[m][o];
    w.h[m][slot] = half_step(x, (R)Q[m][o], cc);
    if (slot >= 2) w.x[m][slot - 2] = x;
  }
}

template <int NF, typename R>
__device__ __forceinline__ void march_shift(Window<NF, R> &w) {
#pragma unroll
  for (int m = 0; m < NF; m++) {
#pragma unroll
    for (int n = 0; n < 7; n++) w.h[m][n] = w.h[m][n + 1];
#pragma unroll
    for (int n = 0; n < 5; n++) w.x[m][n] = w.x[m][n + 1];
  }
}

template <int NF, typename R>
__device__ __forceinline__ void march_lr(const Window<NF, R> &w, int m, const R *T, R &L, R &Rr) {
  R fl = R(0), fr = R(0);
#pragma unroll
  for (int n = 0; n < 7; n++) {
    fl = madd(w.h[m][n], T[n], fl);
    fr = madd(w.h[m][7 - n], T[7 + n], fr);
  }
  // limiter cells: x[0] = f-2, x[1] = f-1, x[2] = f, x[3] = f+1
  L = pdm_limit(fl, w.x[m][0], w.x[m][1], w.x[m][2], T[14], T[15]);
  Rr = pdm_limit(fr, w.x[m][3], w.x[m][2], w.x[m][1], T[16], T[17]);
}

template <int D, typename R, bool B0V, bool UT>
__device__ __forceinline__ void k_march_body(const SweepArgs &a, const int bx, const int by,
                                        const int bz) {
  if (halted(a.sc)) return;
  constexpr int RA = D == 0 ? 1 : 0;  // row axis
  const Grid &g = a.g;
  (void)by;
  int k, r;
  if (!row_point(a.lo[2], a.hi[2], a.lo[RA], a.hi[RA], bx, k, r)) return;
  int f0, f1;
  march_range(a.lo[D], a.hi[D], a.nch, bz, f0, f1);
  int idx[3];
  idx[2] = k;
  idx[RA] = r;
  idx[D] = f0;
  const long sd = g.st[D];
  long c = idx[0] * g.st[0] + idx[1] * g.st[1] + k;
  const R cc = (R)a.sc->c;
  const R *T = rtab<D, R>(a.geo, f0, idx);
  const long tstep = a.geo.tfs[D];
  constexpr int va = D == 0 ? 1 : 2;
  constexpr int vb = D == 0 ? 2 : 3;
  const long fs = g.fs;
  R *F = (R *)a.F, *V = (R *)a.V;
  const R *X = (const R *)a.X, *Xp = (const R *)a.Xp;
  if (r >= a.flo[RA] && r <= a.fhi[RA]) {
    const R *P[9], *Q[9];
#pragma unroll
    for (int m = 0; m < 6; m++) {
      P[m] = X + m * fs;
      Q[m] = Xp + m * fs;
    }
#pragma unroll
    for (int m = 0; m < 3; m++) {
      P[6 + m] = (const R *)a.M + m * fs;
      Q[6 + m] = (const R *)a.Mp + m * fs;
    }
    Window<9, R> w;
#pragma unroll
    for (int p = 0; p < 7; p++) march_load<9, R, R>(w, P, Q, c + (p - 4) * sd, cc, p);
    // FP64: the raw values of the cell entering the window and the
    // face-normal B are loaded one face ahead, so their HBM latency overlaps
    // the previous face's arithmetic (one wave per SIMD leaves no other wave
    // to cover it).  FP32 keeps the plain order, which measured faster.
    constexpr bool AHEAD = sizeof(R) == 8;
    const R *Bd = (const R *)a.B + D * fs;
    R nx[9], nq[9], nb = R(0);
    if constexpr (AHEAD) {
#pragma unroll
      for (int m = 0; m < 9; m++) {
        nx[m] = P[m][c + 3 * sd];
        nq[m] = Q[m][c + 3 * sd];
      }
      nb = Bd[c];
    }
    for (int f = f0; f < f1; f++) {
      R bn;
      if constexpr (AHEAD) {
#pragma unroll
        for (int m = 0; m < 9; m++) {
          w.h[m][7] = half_step(nx[m], nq[m], cc);
          w.x[m][5] = nx[m];
        }
        bn = nb;
        if (f + 1 < f1) {
#pragma unroll
          for (int m = 0; m < 9; m++) {
            nx[m] = P[m][c + 4 * sd];
            nq[m] = Q[m][c + 4 * sd];
          }
          nb = Bd[c + sd];
        }
      } else {
        march_load<9, R, R>(w, P, Q, c + 3 * sd, cc, 7);
        bn = Bd[c];
      }
      R Tu[UT ? GK_RECW_STRIDE : 1];
      const R *Tf = T;
      if constexpr (UT) {
        utab<D, R>(a.geo, f, Tu);
        Tf = Tu;
      }
      R L[6], Rs[6], bl[3], br[3];
#pragma unroll
      for (int m = 0; m < 6; m++) march_lr(w, m, Tf, L[m], Rs[m]);
#pragma unroll
      for (int m = 0; m < 3; m++) march_lr(w, 6 + m, Tf, bl[m], br[m]);
      st_nt(V + c, R(0.5) * (L[va] + Rs[va]));
      st_nt(V + fs + c, R(0.5) * (L[vb] + Rs[vb]));
      R Fo[9], b0[4];
      face_b0<D, R, B0V>(a.b0, a.b0a, fs, c, b0);
      face_flux<D, R>(L, Rs, bl, br, bn, b0, g.CA, Fo);
#pragma unroll
      for (int m = 0; m < 9; m++) st_nt(F + m * fs + c, Fo[m]);
      march_shift(w);
      c += sd;
      T += tstep;
    }
  } else {
    const R *P[2] = {X + va * fs, X + vb * fs};
    const R *Q[2] = {Xp + va * fs, Xp + vb * fs};
    Window<2, R> w;
#pragma unroll
    for (int p = 0; p < 7; p++) march_load<2, R, R>(w, P, Q, c + (p - 4) * sd, cc, p);
    for (int f = f0; f < f1; f++) {
      march_load<2, R, R>(w, P, Q, c + 3 * sd, cc, 7);
      R Tu[UT ? GK_RECW_STRIDE : 1];
      const R *Tf = T;
      if constexpr (UT) {
        utab<D, R>(a.geo, f, Tu);
        Tf = Tu;
      }
      R l, rr;
      march_lr(w, 0, Tf, l, rr);
      V[c] = R(0.5) * (l + rr);
      march_lr(w, 1, Tf, l, rr);
      V[fs + c] = R(0.5) * (l + rr);
      march_shift(w);
      c += sd;
      T += tstep;
    }
  }
}

template <int D, typename R, bool B0V, bool UT>
__global__ __launch_bounds__(256) void k_march(SweepArgs a, int gx, int gy, int gz) {
  int bx, by, bz;
  if (!xcd_block<true>(gx, gy, gz, bx, by, bz)) return;
  k_march_body<D, R, B0V, UT>(a, bx, by, bz);
}

// one launch for many grids: block w runs work-list entry w (argument set, block)
template <int D, typename R, bool B0V, bool UT>
__global__ __launch_bounds__(256) void k_march_batch(const SweepArgs *__restrict__ args,
                                                    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  // uniform per block: keep the argument set in scalar registers
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  const int bx = __builtin_amdgcn_readfirstlane(work[w].bx);
  const int by = __builtin_amdgcn_readfirstlane(work[w].by);
  const int bz = __builtin_amdgcn_readfirstlane(work[w].bz);
  k_march_body<D, R, B0V, UT>(batch_args(args, arg), bx, by, bz);
}

// k sweep (D = 2) with the stencil staged in LDS.  A block is one wave
// covering 64 faces along k of one row (i, j); each reconstructed field is
// staged once (72 cells: 64 faces + 8 stencil halo) as half-step and current
// values.  All of the row's global loads are issued before the first LDS
// write, so the wave waits for memory once, not once per field.  Planes i
// outside [flo0, fhi0] only form the face-averaged velocities (v3, v1).
constexpr int KSPAN = 72;
constexpr int SWEEPK_LDS = 2 * 9 * KSPAN;  // elements of k_sweepk's LDS
constexpr int CELL_CHUNK = 1;  // cells each thread of the cell update walks along i

template <typename R, bool B0V>
__device__ __forceinline__ void k_sweepk_body(const SweepArgs &a, const int bx, const int by,
                                         const int bz, R *lds) {
  if (halted(a.sc)) return;
  R(*sh)[KSPAN] = (R(*)[KSPAN])lds;
  R(*sx)[KSPAN] = (R(*)[KSPAN])(lds + 9 * KSPAN);
  const Grid &g = a.g;
  const int k0 = a.lo[2] + bx * 64;
  const int j = a.lo[1] + by;
  const int i = a.lo[0] + bz;
  const long fs = g.fs;
  const bool full = i >= a.flo[0] && i <= a.fhi[0];
  const R cc = (R)a.sc->c;
  R *F = (R *)a.F, *V = (R *)a.V;
  // lane tx stages point k0-4+tx and, for tx < 8, point k0+60+tx of the row
  // (clamped to the box)
  const int tx = threadIdx.x;
  const bool halo = tx < KSPAN - 64;
  const long rowb = i * g.st[0] + j * g.st[1];
  const long oa = rowb + min(k0 - 4 + tx, g.N[2] - 1);
  const long ob = rowb + min(k0 + 60 + tx, g.N[2] - 1);
  // staged fields: all nine on full planes, else v3 (slot 3) and v1 (slot 1)
  const R *X = (const R *)a.X, *Xp = (const R *)a.Xp;
  const R *M = (const R *)a.M, *Mp = (const R *)a.Mp;
  const R *P[9], *Q[9];
#pragma unroll
  for (int m = 0; m < 6; m++) {
    P[m] = X + m * fs;
    Q[m] = Xp + m * fs;
  }
#pragma unroll
  for (int m = 0; m < 3; m++) {
    P[6 + m] = M + m * fs;
    Q[6 + m] = Mp + m * fs;
  }
  R xa[9], pa[9], xb[9], pb[9];
  auto fetch = [&](int m) {
    xa[m] = P[m][oa];
    pa[m] = Q[m][oa];
    if (halo) {
      xb[m] = P[m][ob];
      pb[m] = Q[m][ob];
    }
  };
  auto stage = [&](int m) {
    sh[m][tx] = half_step(xa[m], pa[m], cc);
    sx[m][tx] = xa[m];
    if (halo) {
      sh[m][64 + tx] = half_step(xb[m], pb[m], cc);
      sx[m][64 + tx] = xb[m];
    }
  };
  if (full) {
#pragma unroll
    for (int m = 0; m < 9; m++) fetch(m);
#pragma unroll
    for (int m = 0; m < 9; m++) stage(m);
  } else {
    fetch(3);
    fetch(1);
    stage(3);
    stage(1);
  }
  __syncthreads();
  const int k = k0 + tx;
  if (k > a.hi[2]) return;
  int idx[3] = {i, j, k};
  const R *T = rtab<2, R>(a.geo, k, idx);
  const long c = rowb + k;
  const int b = tx;  // stencil cells b .. b+7 in LDS
  auto lr = [&](int m, R &L, R &Rr) {
    R fl = R(0), fr = R(0);
#pragma unroll
    for (int n = 0; n < 7; n++) {
      fl = madd(sh[m][b + n], T[n], fl);
      fr = madd(sh[m][b + 7 - n], T[7 + n], fr);
    }
    L = pdm_limit(fl, sx[m][b + 2], sx[m][b + 3], sx[m][b + 4], T[14], T[15]);
    Rr = pdm_limit(fr, sx[m][b + 5], sx[m][b + 4], sx[m][b + 3], T[16], T[17]);
  };
  if (!full) {
    R l, rr;
    lr(3, l, rr);
    V[c] = R(0.5) * (l + rr);
    lr(1, l, rr);
    V[fs + c] = R(0.5) * (l + rr);
    return;
  }
  R L[6], Rs[6], bl[3], br[3];
#pragma unroll
  for (int m = 0; m < 6; m++) lr(m, L[m], Rs[m]);
#pragma unroll
  for (int m = 0; m < 3; m++) lr(6 + m, bl[m], br[m]);
  st_nt(V + c, R(0.5) * (L[3] + Rs[3]));
  st_nt(V + fs + c, R(0.5) * (L[1] + Rs[1]));
  R Fo[9], b0[4];
  face_b0<2, R, B0V>(a.b0, a.b0a, fs, c, b0);
  face_flux<2, R>(L, Rs, bl, br, ((const R *)a.B)[2 * fs + c], b0, g.CA, Fo);
#pragma unroll
  for (int m = 0; m < 9; m++) st_nt(F + m * fs + c, Fo[m]);
}

template <typename R, bool B0V>
__global__ __launch_bounds__(256) void k_sweepk(SweepArgs a, int gx, int gy, int gz) {
  int bx, by, bz;
  if (!xcd_block<false>(gx, gy, gz, bx, by, bz)) return;
  __shared__ R lds[SWEEPK_LDS];
  k_sweepk_body<R, B0V>(a, bx, by, bz, lds);
}

// The k sweep's 64-face tiles and the k_kcol columns left over after them
// in one launch: blocks [0, nrem) run the columns (dispatched first, beside
// the tiles), blocks from nrem8 = nrem rounded up to a multiple of NXCD the
// tiles, in their own XCD-aware order.  The two parts share one LDS array.
template <typename R, bool B0V>
__global__ __launch_bounds__(64) void k_sweepk_rem(SweepArgs a, SweepArgs rem, int gx, int gy,
                                                   int gz, int nrem, int nrem8) {
  __shared__ R lds[KCOL_LDS > SWEEPK_LDS ? KCOL_LDS : SWEEPK_LDS];
  const int b = blockIdx.x;
  if (b < nrem8) {
    if (b < nrem) k_kcol_body<R, B0V>(rem, b, 0, 0, lds);
    return;
  }
  int bx, by, bz;
  if (!xcd_block_at<false>(b - nrem8, gx, gy, gz, bx, by, bz)) return;
  k_sweepk_body<R, B0V>(a, bx, by, bz, lds);
}

// one launch for many grids: block w runs work-list entry w (argument set, block)
template <typename R, bool B0V>
__global__ __launch_bounds__(256) void k_sweepk_batch(const SweepArgs *__restrict__ args,
                                                    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  // uniform per block: keep the argument set in scalar registers
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  const int bx = __builtin_amdgcn_readfirstlane(work[w].bx);
  const int by = __builtin_amdgcn_readfirstlane(work[w].by);
  const int bz = __builtin_amdgcn_readfirstlane(work[w].bz);
  __shared__ R lds[SWEEPK_LDS];
  k_sweepk_body<R, B0V>(batch_args(args, arg), bx, by, bz, lds);
}

// Edge EMF from the edge velocities (va, vb), the two reconstructed face-B
// components (b1 = field F1 along R1, b2 = field F2 along R2) and the sum of
// the four half-step densities around the edge (reference
// src/solver_func.c:779-932; E_i: va = v2, vb = v3, b1 = B_j, b2 = B_k;
// E_j: va = v3, vb = v1, b1 = B_k, b2 = B_i; E_k: va = v1, vb = v2,
// b1 = B_i, b2 = B_j).
template <int D, typename R>
__device__ __forceinline__ R emf_formula(R va, R vb, R b1l, R b1r, R b2l, R b2r, R rsum,
                                         const double *e0d, double CAd) {
  const R CA = (R)CAd, h = R(0.5);
  const R e0[3] = {(R)e0d[0], (R)e0d[1], (R)e0d[2]};
  R e, dv;
  // a + b * c as madd(b, c, a), a - b * c as madd(-b, c, a): the reference's
  // operations and order
  if (D == 0) {
    R bja = madd(h, b1l + b1r, e0[1]);
    R bka = madd(h, b2l + b2r, e0[2]);
    e = -madd(-vb, bja, va * bka);
    dv = qsqrt(qdiv(madd(bka, bka, bja * bja), R(0.25) * rsum));
    dv = qdiv(dv * CA, qsqrt(madd(dv, dv, CA * CA))) + qsqrt(madd(vb, vb, va * va));
    e = madd(h * dv, b1l + b2r - b1r - b2l, e);
  } else if (D == 1) {
    R bka = madd(h, b1l + b1r, e0[2]);
    R bia = madd(h, b2l + b2r, e0[0]);
    e = -madd(-vb, bka, va * bia);
    dv = qsqrt(qdiv(madd(bia, bia, bka * bka), R(0.25) * rsum));
    dv = qdiv(dv * CA, qsqrt(madd(dv, dv, CA * CA))) + qsqrt(madd(va, va, vb * vb));
    e = madd(h * dv, b1l + b2r - b1r - b2l, e);
  } else {
    R bia = madd(h, b1l + b1r, e0[0]);
    R bja = madd(h, b2l + b2r, e0[1]);
    e = -madd(-vb, bia, va * bja);
    dv = qsqrt(qdiv(madd(bja, bja, bia * bia), R(0.25) * rsum));
    dv = qdiv(dv * CA, qsqrt(madd(dv, dv, CA * CA))) + qsqrt(madd(va, va, vb * vb));
    e = madd(h * dv, b1l + b2r - b1r - b2l, e);
  }
  return e;
}

// ---------------------------------------------------------------------------
struct EmfArgs {
  Grid g;
  Geo geo;
  B0c b0;
  const double *b0a;   // per-point B0 (b0_mode 2), else null
  const void *X, *Xp;  // of type R (FP64 state or its FP32 copy)
  const void *B, *Bp;
  const void *V;  // 2 fields of R: face-averaged velocity (from the feeding sweep)
  void *E;        // 1 field of R
  const Scal *sc;
  int lo[3], hi[3];
  int nch;        // near-equal chunks of the march range (one per thread)
};

__device__ __forceinline__ void globalize(EmfArgs &a) {
  globalize(a.geo);
  gptr(a.b0a);
  gptr(a.X);
  gptr(a.Xp);
  gptr(a.B);
  gptr(a.Bp);
  gptr(a.V);
  gptr(a.E);
  gptr(a.sc);
}

// Marching EMF along the axis MA that two of the three reconstructions share
// (E_i: j, with B_k windowed and the k-reconstructions of v and B_j read
// directly; E_j: i, v and B_k windowed; E_k: j, v and B_i windowed).  Lanes
// run along k; the four half-step densities around the edge are carried
// from one march position to the next.  E_i on edges (i, j-1/2, k-1/2) etc.,
// edge ranges [s, e+1] across and [s, e] along the edge.
template <typename R>
struct Window1 {  // self-limited field: half-step = current value
  R h[8];
};

template <typename R>
__device__ __forceinline__ void lr_self(const Window1<R> &w, const R *T, R &L, R &Rr) {
  R fl = R(0), fr = R(0);
#pragma unroll
  for (int n = 0; n < 7; n++) {
    fl = madd(w.h[n], T[n], fl);
    fr = madd(w.h[7 - n], T[7 + n], fr);
  }
  L = pdm_limit(fl, w.h[2], w.h[3], w.h[4], T[14], T[15]);
  Rr = pdm_limit(fr, w.h[5], w.h[4], w.h[3], T[16], T[17]);
}

template <typename R>
__device__ __forceinline__ void shift1(Window1<R> &w) {
#pragma unroll
  for (int n = 0; n < 7; n++) w.h[n] = w.h[n + 1];
}

// Reconstruction table of axis A at a fixed face for a point that marches
// along MA, one of A's transverse axes: only the class of the march
// coordinate changes from one position to the next.  That class is looked up
// one position ahead, so the table loads of a position are issued together
// with its field loads instead of waiting for the lookup.
template <int A, int MA, typename R>
struct TableWalk {
  const R *base;
  long mul;
  const int *cls;
  int next;
  __device__ __forceinline__ void init(const Geo &g, int face, const int idx[3], int f0) {
    constexpr int B = A == 0 ? 1 : 0;
    constexpr int C = A == 2 ? 1 : 2;
    static_assert(MA == B || MA == C, "march axis must be transverse to the table axis");
    const R *t;
    if constexpr (sizeof(R) == 8)
      t = (const R *)g.recw[A];
    else
      t = (const R *)g.recwf[A];
    const long n1 = g.ncls[A][1];
    if (MA == B) {  // fixed class on C
      base = t + (long)face * g.tfs[A] + g.cls[A][1][idx[C]] * g.tps[A];
      mul = n1 * g.tps[A];
      cls = g.cls[A][0];
    } else {        // fixed class on B
      base = t + (long)face * g.tfs[A] + g.cls[A][0][idx[B]] * n1 * g.tps[A];
      mul = g.tps[A];
      cls = g.cls[A][1];
    }
    next = cls[f0];
  }
  // table at march position f (positions are visited f0, f0 + 1, ... < f1)
  __device__ __forceinline__ const R *at(int f, int f1) {
    const R *t = base + next * mul;
    if (f + 1 < f1) next = cls[f + 1];
    return t;
  }
};

// Occupancy bounds of the edge-field kernels.  The FP64 E_j and E_k marches
// keep three windows and reconstruct a fourth field directly; at two waves
// per SIMD they have the registers to issue a position's table and field
// loads together (at three the compiler loads the table piecewise), and at
// least two keep the compiler from spending more registers than two waves
// leave (with uniform tables it would otherwise drop to one).
template <int D, typename R, bool UT>
struct EmfWaves {
  static constexpr bool two = sizeof(R) == 8 && D != 0;
  static constexpr int lo = two ? 2 : 1;
  static constexpr int hi = two ? 2 : 8;
};

// STAGED (E_i and E_j, whose direct reconstructions run along k; the plan
// uses it for FP64 E_i): a block is
// one wave on 64 consecutive edges of one row; at each march position the
// row's 72 cells of every directly reconstructed field are loaded once and
// staged in LDS (current and half-step values), and each lane reads its
// 8-cell stencil from there instead of loading it.  Whole 64-edge tiles only;
// the unstaged form covers the rest.
template <int D, typename R, bool STAGED, bool B0V, bool UT>
__device__ __forceinline__ void k_emf_march_body(const EmfArgs &a, const int bx, const int by,
                                            const int bz) {
  if (halted(a.sc)) return;
  constexpr int MA = D == 1 ? 0 : 1;  // march axis
  constexpr int RA = D == 1 ? 1 : 0;  // row axis
  constexpr int A2 = D == 0 ? 2 : (D == 1 ? 0 : 1);  // velocity (and F1) axis
  constexpr int F1 = D == 0 ? 1 : (D == 1 ? 2 : 0);
  constexpr int F2 = D == 0 ? 2 : (D == 1 ? 0 : 1);
  constexpr int R2 = D == 0 ? 1 : (D == 1 ? 2 : 0);
  constexpr int P = D == 0 ? 1 : 0;  // density cells: c - st[P] - st[Q] ... c
  constexpr int Q = D == 2 ? 1 : 2;
  constexpr int O = P == MA ? Q : P;  // density axis other than the march axis
  constexpr bool VW = A2 == MA;       // v and F1 windowed
  const Grid &g = a.g;
  int k, r;
  if (STAGED) {
    static_assert(!STAGED || D != 2, "E_k reconstructs across i, not k");
    k = a.lo[2] + bx * 64 + (int)threadIdx.x;
    r = a.lo[RA] + by;
  } else {
    (void)by;
    if (!row_point(a.lo[2], a.hi[2], a.lo[RA], a.hi[RA], bx, k, r)) return;
  }
  int f0, f1;
  march_range(a.lo[MA], a.hi[MA], a.nch, bz, f0, f1);
  int idx[3];
  idx[2] = k;
  idx[RA] = r;
  idx[MA] = f0;
  const long sm = g.st[MA], so = g.st[O];
  const long fs = g.fs;
  long c = idx[0] * g.st[0] + idx[1] * g.st[1] + k;
  const R cc = (R)a.sc->c;
  const R *V0 = (const R *)a.V, *V1 = (const R *)a.V + fs;
  R *E = (R *)a.E;
  const R *B1 = (const R *)a.B + F1 * fs, *B1p = (const R *)a.Bp + F1 * fs;
  const R *B2 = (const R *)a.B + F2 * fs, *B2p = (const R *)a.Bp + F2 * fs;
  const R *X = (const R *)a.X, *Xp = (const R *)a.Xp;
  auto rh = [&](long o) {
    R x = (R)X[o];
    return half_step(x, (R)Xp[o], cc);
  };
  const R *Tm = rtab<MA, R>(a.geo, f0, idx);  // windowed reconstructions
  const long tstep = a.geo.tfs[MA];
  Window1<R> wv0, wv1;
  Window<1, R> wb;
  const R *Bw = VW ? B1 : B2, *Bwp = VW ? B1p : B2p;
  const R *Pw[1] = {Bw}, *Qw[1] = {Bwp};
#pragma unroll
  for (int p = 0; p < 7; p++) {
    const long o = c + (p - 4) * sm;
    if (VW) {
      wv0.h[p] = V0[o];
      wv1.h[p] = V1[o];
    }
    march_load<1, R, R>(wb, Pw, Qw, o, cc, p);
  }
  // direct reconstructions: v and F1 along A2 (E_i), else F2 along R2
  constexpr int DA = VW ? R2 : A2;
  TableWalk<DA, MA, R> tw;
  // uniform tables: the direct reconstructions' face does not change along
  // the march, so neither does their table.  E_i loads it once; E_j and E_k
  // (which keep three windows) load it per position, as holding it across
  // the march leaves them too few registers at two waves per SIMD.
  constexpr bool HOIST = UT && (!VW || STAGED);
  R Tdu[HOIST ? GK_RECW_STRIDE : 1];
  if constexpr (HOIST) {
    const R *t = (sizeof(R) == 8 ? (const R *)a.geo.recw[DA] : (const R *)a.geo.recwf[DA]) +
                 (long)idx[DA] * GK_RECW_STRIDE;
#pragma unroll
    for (int n = 0; n < GK_RECW_STRIDE; n++) Tdu[n] = t[n];
  } else {
    tw.init(a.geo, idx[DA], idx, f0);
  }
  // staged form: LDS rows of the direct fields (E_i: v_a, v_b, then F1
  // current and half-step; E_j: F2 current and half-step)
  constexpr int NS = STAGED ? (VW ? 2 : 4) : 1;
  __shared__ R stg[NS][KSPAN];
  const int tx = threadIdx.x;
  const long ka = (long)min(k - 4, g.N[2] - 1) - k;   // lane's first staged cell, offset from k
  const long kb = (long)min(k + 60, g.N[2] - 1) - k;  // its halo cell (lanes < 8)
  // densities at march position f0 - 1 (offsets -1 and 0 along O)
  R u0 = rh(c - sm - so), u1 = rh(c - sm);
  for (int f = f0; f < f1; f++) {
    const long o = c + 3 * sm;
    if (VW) {
      wv0.h[7] = V0[o];
      wv1.h[7] = V1[o];
    }
    march_load<1, R, R>(wb, Pw, Qw, o, cc, 7);
    const R *Td;
    if constexpr (HOIST)
      Td = Tdu;
    else
      Td = tw.at(f, f1);
    R Tu[UT ? GK_RECW_STRIDE : 1];
    const R *Tw = Tm;
    if constexpr (UT) {
      utab<MA, R>(a.geo, f, Tu);
      Tw = Tu;
    }
    R va, vb, b1l, b1r, b2l, b2r, l, rr;
    if constexpr (STAGED) {
      // stage the row: every global load first, then the LDS writes
      const R *S0 = VW ? B2 : V0, *S1 = VW ? B2p : V1;
      R ya[4], yb[4];
      ya[0] = S0[c + ka];
      ya[1] = S1[c + ka];
      if (!VW) {
        ya[2] = B1[c + ka];
        ya[3] = B1p[c + ka];
      }
      if (tx < KSPAN - 64) {
        yb[0] = S0[c + kb];
        yb[1] = S1[c + kb];
        if (!VW) {
          yb[2] = B1[c + kb];
          yb[3] = B1p[c + kb];
        }
      }
      auto put = [&](const R v[4], int at) {
        if (VW) {  // F2: current, half-step
          stg[0][at] = v[0];
          stg[1][at] = half_step(v[0], v[1], cc);
        } else {   // v_a, v_b, F1 current, F1 half-step
          stg[0][at] = v[0];
          stg[1][at] = v[1];
          stg[2][at] = v[2];
          stg[3][at] = half_step(v[2], v[3], cc);
        }
      };
      put(ya, tx);
      if (tx < KSPAN - 64) put(yb, 64 + tx);
      __syncthreads();
      auto lds_lr = [&](int sx, int sh, R &L, R &Rr) {
        R xh[8], xn[8];
#pragma unroll
        for (int n = 0; n < 8; n++) {
          xn[n] = stg[sx][tx + n];
          xh[n] = stg[sh][tx + n];
        }
        pdm_lr(xh, xn, Td, L, Rr);
      };
      if (VW) {
        lr_self(wv0, Tw, l, rr);
        va = R(0.5) * (l + rr);
        lr_self(wv1, Tw, l, rr);
        vb = R(0.5) * (l + rr);
        march_lr(wb, 0, Tw, b1l, b1r);
        lds_lr(0, 1, b2l, b2r);
      } else {
        lds_lr(0, 0, l, rr);
        va = R(0.5) * (l + rr);
        lds_lr(1, 1, l, rr);
        vb = R(0.5) * (l + rr);
        lds_lr(2, 3, b1l, b1r);
        march_lr(wb, 0, Tw, b2l, b2r);
      }
      __syncthreads();  // the next position is staged over this one
    } else if (VW) {
      lr_self(wv0, Tw, l, rr);
      va = R(0.5) * (l + rr);
      lr_self(wv1, Tw, l, rr);
      vb = R(0.5) * (l + rr);
      march_lr(wb, 0, Tw, b1l, b1r);
      recon(B2, B2p, c, g.st[R2], cc, Td, b2l, b2r);
    } else {
      const R *T2 = Td;
      recon(V0, (const R *)nullptr, c, g.st[A2], cc, T2, l, rr);
      va = R(0.5) * (l + rr);
      recon(V1, (const R *)nullptr, c, g.st[A2], cc, T2, l, rr);
      vb = R(0.5) * (l + rr);
      recon(B1, B1p, c, g.st[A2], cc, T2, b1l, b1r);
      march_lr(wb, 0, Tw, b2l, b2r);
    }
    const R w0 = rh(c - so), w1 = rh(c);
    // reference order: rho(c-sP-sQ) + rho(c-sP) + rho(c-sQ) + rho(c)
    R rsum;
    if (Q == MA)
      rsum = u0 + w0 + u1 + w1;
    else
      rsum = u0 + u1 + w0 + w1;
    double e0[3];
    edge_b0<D, B0V>(a.b0, a.b0a, fs, c, e0);
    st_nt(E + c, emf_formula<D, R>(va, vb, b1l, b1r, b2l, b2r, rsum, e0, g.CA));
    u0 = w0;
    u1 = w1;
    if (VW) {
      shift1(wv0);
      shift1(wv1);
    }
    march_shift(wb);
    c += sm;
    Tm += tstep;
  }
}

template <int D, typename R, bool B0V, bool UT>
__global__ __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(EmfWaves<D, R, UT>::lo, EmfWaves<D, R, UT>::hi))) void k_emf_march(EmfArgs a, int gx, int gy, int gz) {
  int bx, by, bz;
  // E_k reads its direct reconstruction across eight i rows: rows fastest,
  // so that the blocks an XCD runs together share those rows in its L2 (the
  // others run a row's march chunks together, which share window halos)
  if (!xcd_block<D != 2>(gx, gy, gz, bx, by, bz)) return;
  k_emf_march_body<D, R, false, B0V, UT>(a, bx, by, bz);
}

// The k columns of E_i and E_j left over after their whole 64-edge tiles, in
// one launch (one edge per thread; blocks [0, g0) are E_i's, decoded as its
// own launch would be, the rest E_j's): each alone is a latency-bound launch
// of about a thousand waves.
template <typename R, bool B0V, bool UT>
__global__ __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(EmfWaves<1, R, UT>::lo, EmfWaves<1, R, UT>::hi))) void k_emf_cols(EmfArgs a0, EmfArgs a1, dim3 g0, dim3 g1) {
  const int t0 = (int)(g0.x * g0.y * g0.z), t1 = (int)(g1.x * g1.y * g1.z);
  int L;
  if (!xcd_linear(t0 + t1, L)) return;
  const bool first = L < t0;
  const dim3 gr = first ? g0 : g1;
  if (!first) L -= t0;
  const int bz = L % (int)gr.z, t = L / (int)gr.z;  // z fastest, as xcd_block<true>
  const int bx = t % (int)gr.x, by = t / (int)gr.x;
  if (first)
    k_emf_march_body<0, R, false, B0V, UT>(a0, bx, by, bz);
  else
    k_emf_march_body<1, R, false, B0V, UT>(a1, bx, by, bz);
}

// staged form over whole 64-edge tiles along k (E_i, E_j): one wave per block
template <int D, typename R, bool B0V, bool UT>
__global__ __launch_bounds__(64) void k_emf_stage(EmfArgs a, int gx, int gy, int gz) {
  int bx, by, bz;
  if (!xcd_block<true>(gx, gy, gz, bx, by, bz)) return;
  k_emf_march_body<D, R, true, B0V, UT>(a, bx, by, bz);
}

template <int D, typename R, bool B0V, bool UT>
__global__ __launch_bounds__(64) void k_emf_stage_batch(const EmfArgs *__restrict__ args,
                                                       const WorkItem *__restrict__ work,
                                                       int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  const int bx = __builtin_amdgcn_readfirstlane(work[w].bx);
  const int by = __builtin_amdgcn_readfirstlane(work[w].by);
  const int bz = __builtin_amdgcn_readfirstlane(work[w].bz);
  k_emf_march_body<D, R, true, B0V, UT>(batch_args(args, arg), bx, by, bz);
}

// one launch for many grids: block w runs work-list entry w (argument set, block)
template <int D, typename R, bool B0V, bool UT>
__global__ __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(EmfWaves<D, R, UT>::lo, EmfWaves<D, R, UT>::hi))) void k_emf_march_batch(const EmfArgs *__restrict__ args,
                                                    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  // uniform per block: keep the argument set in scalar registers
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  const int bx = __builtin_amdgcn_readfirstlane(work[w].bx);
  const int by = __builtin_amdgcn_readfirstlane(work[w].by);
  const int bz = __builtin_amdgcn_readfirstlane(work[w].bz);
  k_emf_march_body<D, R, false, B0V, UT>(batch_args(args, arg), bx, by, bz);
}

// ---------------------------------------------------------------------------
struct CtArgs {
  Grid g;
  Geo geo;
  const double *B;   // B^n
  double *Bn;        // B^{n+1}
  float *Bn32;       // FP32 copy of B^{n+1} (FP32-bulk form), else null
  const void *E;     // 3 edge fields of R
  const Scal *sc;
};

__device__ __forceinline__ void globalize(CtArgs &a) {
  globalize(a.geo);
  gptr(a.B);
  gptr(a.Bn);
  gptr(a.Bn32);
  gptr(a.E);
  gptr(a.sc);
}

// Constrained-transport face update (reference src/solver_solve.c:262-306),
// in FP64 for both forms (EMFs of type R are widened first), so the discrete
// div B stays at FP64 roundoff.
template <typename R>
__device__ __forceinline__ void k_ct_body(const CtArgs &a, const int bx, const int by,
                                     const int bz) {
  if (halted(a.sc)) return;
  const Grid &g = a.g;
  (void)by;
  int k, j;
  if (!row_point(g.s[2], g.e[2] + 1, g.s[1], g.e[1] + 1, bx, k, j)) return;
  const int i = g.s[0] + bz;
  const long c = i * g.st[0] + j * g.st[1] + k;
  const long SI = g.st[0], SJ = g.st[1];
  const double dt = a.sc->dt;
  const R *Ei = (const R *)a.E, *Ej = Ei + g.fs, *Ek = Ei + 2 * g.fs;
  const Geo &G = a.geo;
  const bool ai = i <= g.e[0], aj = j <= g.e[1], ak = k <= g.e[2];
  // every input of the point (edge fields, face B, metric factors) is read
  // before the first store, so the wave waits for memory once; neighbour
  // indices stay inside the box for i, j, k <= e + 1
  const double ei = (double)Ei[c], ei_j = (double)Ei[c + SJ], ei_k = (double)Ei[c + 1];
  const double ej = (double)Ej[c], ej_i = (double)Ej[c + SI], ej_k = (double)Ej[c + 1];
  const double ek = (double)Ek[c], ek_i = (double)Ek[c + SI], ek_j = (double)Ek[c + SJ];
  const double bi = a.B[c], bj = a.B[g.fs + c], bk = a.B[2 * g.fs + c];
  const double li = geo3(G, GK_G_EDGE_I, i, j, k), li_j = geo3(G, GK_G_EDGE_I, i, j + 1, k),
               li_k = geo3(G, GK_G_EDGE_I, i, j, k + 1);
  const double lj = geo3(G, GK_G_EDGE_J, i, j, k), lj_i = geo3(G, GK_G_EDGE_J, i + 1, j, k),
               lj_k = geo3(G, GK_G_EDGE_J, i, j, k + 1);
  const double lk = geo3(G, GK_G_EDGE_K, i, j, k), lk_i = geo3(G, GK_G_EDGE_K, i + 1, j, k),
               lk_j = geo3(G, GK_G_EDGE_K, i, j + 1, k);
  const double si = geo3(G, GK_G_FACE_I, i, j, k), sj = geo3(G, GK_G_FACE_J, i, j, k),
               sk = geo3(G, GK_G_FACE_K, i, j, k);
  if (aj && ak) {  // B_i on faces i = is..ie+1
    const double d = (ek_j * lk_j - ek * lk) - (ej_k * lj_k - ej * lj);
    const double b = bi - dt * d / si;
    st_nt(a.Bn + c, b);
    if (a.Bn32) st_nt(a.Bn32 + c, (float)b);
  }
  if (ai && ak) {  // B_j on faces j = js..je+1
    const double d = (ei_k * li_k - ei * li) - (ek_i * lk_i - ek * lk);
    const double b = bj - dt * d / sj;
    st_nt(a.Bn + g.fs + c, b);
    if (a.Bn32) st_nt(a.Bn32 + g.fs + c, (float)b);
  }
  if (ai && aj) {  // B_k on faces k = ks..ke+1
    const double d = (ej_i * lj_i - ej * lj) - (ei_j * li_j - ei * li);
    const double b = bk - dt * d / sk;
    st_nt(a.Bn + 2 * g.fs + c, b);
    if (a.Bn32) st_nt(a.Bn32 + 2 * g.fs + c, (float)b);
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_ct(CtArgs a, int gx, int gy, int gz) {
  int bx, by, bz;
  (void)gy;
  by = 0;
  if (!xcd_patch(gx, gz, bx, bz)) return;
  k_ct_body<R>(a, bx, by, bz);
}

// one launch for many grids: block w runs work-list entry w (argument set, block)
template <typename R>
__global__ __launch_bounds__(256) void k_ct_batch(const CtArgs *__restrict__ args,
                                                    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  // uniform per block: keep the argument set in scalar registers
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  const int bx = __builtin_amdgcn_readfirstlane(work[w].bx);
  const int by = __builtin_amdgcn_readfirstlane(work[w].by);
  const int bz = __builtin_amdgcn_readfirstlane(work[w].bz);
  k_ct_body<R>(batch_args(args, arg), bx, by, bz);
}

// ---------------------------------------------------------------------------
struct CellArgs {
  Grid g;
  Geo geo;
  B0c b0;
  const double *b0a;       // per-point B0 (b0_mode 2), else null
  int spherical;
  const double *X, *Xp;    // X^n, X^{n-1}
  double *Xn;              // X^{n+1} (aliases Xp: each cell reads before it writes)
  float *Xn32;             // FP32 copy of X^{n+1} (FP32-bulk form), else null
  const void *M, *Mp;      // cell B at n, n-1, of type R (FP64, or the FP32 copies)
  double *Mn;              // cell B at n+1 of the active cells (FP64 form), else null
  float *Mn32;             // its FP32 copy (FP32-bulk form), else null
  const double *Bn;        // face B at n+1
  const void *F[3];        // 9 fields of R per direction
  double rho_floor, p_floor;  // positivity floors
  double grav[3];          // uniform gravity G01..G03
  Scal *sc;
  int skip[4];             // j rows [skip[0], skip[1]], [skip[2], skip[3]] left out of
                           // the fused dt / div B reduction (pole ring rows,
                           // reduced again after the ring average); empty if lo > hi
};

__device__ __forceinline__ void globalize(CellArgs &a) {
  globalize(a.geo);
  gptr(a.b0a);
  gptr(a.X);
  gptr(a.Xp);
  gptr(a.Xn);
  gptr(a.Xn32);
  gptr(a.M);
  gptr(a.Mp);
  gptr(a.Mn);
  gptr(a.Mn32);
  gptr(a.Bn);
  gptr(a.F[0]);
  gptr(a.F[1]);
  gptr(a.F[2]);
  gptr(a.sc);
}

// Cell update (reference src/solver_solve.c:178-255, 314-484; positivity
// src/solver_func.c:2117-2135; spherical sources
// src/spherical/spherical_general.c:605-688) plus the next step's CFL dt and
// |div B| (src/solver_func.c:255-343).
template <typename R, bool B0V>
__device__ __forceinline__ void k_cell_body(const CellArgs &a, const int bx, const int by,
                                       const int bz) {
  if (halted(a.sc)) return;
  const Grid &g = a.g;
  const Geo &G = a.geo;
  int k, j;
  const bool valid = row_point(g.s[2], g.e[2], g.s[1], g.e[1], bx, k, j);
  // each thread walks CELL_CHUNK cells along i, so the i-face values a cell
  // shares with the next (fluxes, face B) are re-read from cache, not HBM
  const int i0 = g.s[0] + bz * CELL_CHUNK;
  const int i1 = min(i0 + CELL_CHUNK, g.e[0] + 1);
  double dtc = INFINITY, divb = 0.0;
  bool bad = false;
  if (valid)
#pragma unroll 1
  for (int i = i0; i < i1; i++) {
    const long c = i * g.st[0] + j * g.st[1] + k;
    const int idx[3] = {i, j, k};
    const double dt = a.sc->dt;
    const double CA = g.CA;
    const double V = geo3(G, GK_G_VOL, i, j, k);
    const double iV = 1.0 / V;
    // cell widths of the CFL dt, read here: loads after the state stores
    // below could not be moved ahead of them
    const double dxi = geo3(G, GK_G_DXI, i, j, k), dxj = geo3(G, GK_G_DXJ, i, j, k),
                 dxk = geo3(G, GK_G_DXK, i, j, k);
    double A0[3], A1[3];
    A0[0] = geo3(G, GK_G_FACE_I, i, j, k);
    A1[0] = geo3(G, GK_G_FACE_I, i + 1, j, k);
    A0[1] = geo3(G, GK_G_FACE_J, i, j, k);
    A1[1] = geo3(G, GK_G_FACE_J, i, j + 1, k);
    A0[2] = geo3(G, GK_G_FACE_K, i, j, k);
    A1[2] = geo3(G, GK_G_FACE_K, i, j, k + 1);
    // flux divergence, accumulated i then j then k
    double dU[9];
#pragma unroll
    for (int m = 0; m < 9; m++) dU[m] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const R *Fd = (const R *)a.F[d];
      const long sd = g.st[d];
#pragma unroll
      for (int m = 0; m < 9; m++) {
        double v = (double)Fd[m * g.fs + c + sd] * A1[d] - (double)Fd[m * g.fs + c] * A0[d];
        // FP64 divides every term by V as the reference does; FP32-bulk
        // (fluxes already rounded to FP32) multiplies by 1/V
        v = sizeof(R) == 4 ? v * iV : v / V;
        dU[m] = d == 0 ? v : dU[m] + v;
      }
    }
    // background field at the cell: Lorentz force, cell-centre B0 and the
    // face normals of the div B check
    double lor[3], cb0[3], nb0[3][2];
    if constexpr (B0V) {
#pragma unroll
      for (int m = 0; m < 3; m++) {
        lor[m] = a.b0a[(B0A_LOR + m) * g.fs + c];
        cb0[m] = a.b0a[(B0A_CELL + m) * g.fs + c];
        nb0[m][0] = a.b0a[(B0A_FACE + 5 * m) * g.fs + c];
        nb0[m][1] = a.b0a[(B0A_FACE + 5 * m) * g.fs + c + g.st[m]];
      }
    } else {
#pragma unroll
      for (int m = 0; m < 3; m++) {
        lor[m] = a.b0.lor[m];
        cb0[m] = a.b0.cell[m];
        nb0[m][0] = nb0[m][1] = a.b0.face[m][m];
      }
    }
#pragma unroll
    for (int m = 0; m < 3; m++) dU[6 + m] += lor[m];

    double x[6];
#pragma unroll
    for (int m = 0; m < 6; m++) x[m] = ld_nt(a.X + m * g.fs + c);
    double bold[3];
#pragma unroll
    for (int m = 0; m < 3; m++) bold[m] = (double)ld_nt((const R *)a.M + m * g.fs + c);
    // face B at n+1 on the cell's six faces (cell-centred B and div B below),
    // loaded with the other inputs so the wave waits for memory once
    double bf[3][2];
#pragma unroll
    for (int m = 0; m < 3; m++) {
      bf[m][0] = a.Bn[m * g.fs + c];
      bf[m][1] = a.Bn[m * g.fs + c + g.st[m]];
    }
    // cell B at n+1 (same formula as the reference)
    double bnew[3];
#pragma unroll
    for (int m = 0; m < 3; m++) bnew[m] = cell_b_of(bf[m][0], bf[m][1], G, idx[m], m);

    double SH[3] = {0.0, 0.0, 0.0}, SB[3] = {0.0, 0.0, 0.0};
    if (a.spherical) {
      const double cc = a.sc->c;
      double xp[6];
#pragma unroll
      for (int m = 0; m < 6; m++) xp[m] = a.Xp[m * g.fs + c];
      double rh = x[0] + cc * (x[0] - xp[0]);
      double v2h = x[2] + cc * (x[2] - xp[2]);
      double v3h = x[3] + cc * (x[3] - xp[3]);
      double ph = x[4] + cc * (x[4] - xp[4]);
      double M22 = ph + rh * v2h * v2h;
      double M33 = ph + rh * v3h * v3h;
      const R *Fi = (const R *)a.F[0], *Fj = (const R *)a.F[1];
      const long SI = g.st[0], SJ = g.st[1];
      const double dx1 = G.g[GK_G_EDGE_I][0][i];  // dx1
      const double x1c = G.g[GK_G_DXJ][0][i];     // x1c (cell-centre radius)
      const double s0 = G.g[GK_G_EDGE_K][1][j], s1 = G.g[GK_G_EDGE_K][1][j + 1];  // sin(x2)
      SH[0] = 0.5 * (M22 + M33) * (A1[0] - A0[0]) / V;
      double s21 = -(Fi[2 * g.fs + c] * A0[0] + Fi[2 * g.fs + c + SI] * A1[0]) * dx1 / x1c / 2.0 / V;
      double s22 = (A1[1] - A0[1]) / V * M33;
      SH[1] = s21 + s22;
      SH[2] = -(Fi[3 * g.fs + c] * A0[0] + Fi[3 * g.fs + c + SI] * A1[0]) * dx1 / x1c / 2.0 / V;
      SH[2] += -(Fj[3 * g.fs + c] * A0[1] + Fj[3 * g.fs + c + SJ] * A1[1]) * (s1 - s0) / (s1 + s0) / V;
      double bh[3];
#pragma unroll
      for (int m = 0; m < 3; m++) {
        double bp = (double)((const R *)a.Mp)[m * g.fs + c];
        bh[m] = bold[m] + cc * (bold[m] - bp);
      }
      const double *c0 = cb0;
      double Bsq = bh[0] * bh[0] + bh[1] * bh[1] + bh[2] * bh[2] +
                   2 * (bh[0] * c0[0] + bh[1] * c0[1] + bh[2] * c0[2]);
      double M22b = (0.5 * Bsq - (bh[1] + c0[1]) * (bh[1]) - bh[1] * c0[1]);
      double M33b = (0.5 * Bsq - (bh[2] + c0[2]) * (bh[2]) - bh[2] * c0[2]);
      SB[0] = 0.5 * (M22b + M33b) * (A1[0] - A0[0]) / V;
      double b21 = -(Fi[7 * g.fs + c] * A0[0] + Fi[7 * g.fs + c + SI] * A1[0]) * dx1 / x1c / 2.0 / V;
      double b22 = (A1[1] - A0[1]) / V * M33b;
      SB[1] = b21 + b22;
      SB[2] = -(Fi[8 * g.fs + c] * A0[0] + Fi[8 * g.fs + c + SI] * A1[0]) * dx1 / x1c / 2.0 / V;
      SB[2] += -(Fj[8 * g.fs + c] * A0[1] + Fj[8 * g.fs + c + SJ] * A1[1]) * (s1 - s0) / (s1 + s0) / V;
    }

    // conservative variables at n (get_derived_variables, src/solver_func.c:356-382)
    const double rho0 = x[0];
    double ru0[3] = {rho0 * x[1], rho0 * x[2], rho0 * x[3]};
    double eng0 = 0.5 * rho0 * (x[1] * x[1] + x[2] * x[2] + x[3] * x[3]) + x[4] / (GAMMA - 1.0);
    double ent0 = entropy_over(x[5], rho0);
    // conservative update and cons -> prim
    double rho = rho0 - dt * dU[0];
    double ru[3];
#pragma unroll
    for (int m = 0; m < 3; m++) ru[m] = ru0[m] - dt * dU[1 + m] + dt * SH[m];
    double eng = eng0 - dt * dU[4];
    double ent = ent0 - dt * dU[5];
    double v1 = ru[0] / rho, v2 = ru[1] / rho, v3 = ru[2] / rho;
    double p = (eng - 0.5 * (v1 * v1 + v2 * v2 + v3 * v3) * rho) * (GAMMA - 1.0);
    double ps = entropy_times(ent, rho);
    double rhof = rho;
    if (rhof < a.rho_floor) {
      rhof = a.rho_floor;
      p = a.p_floor;
    }
    if (p < a.p_floor) p = a.p_floor;

    // time-centred total field
    double mt[3];
#pragma unroll
    for (int m = 0; m < 3; m++) mt[m] = 0.5 * (bnew[m] + bold[m]) + cb0[m];
    double alf = (mt[0] * mt[0] + mt[1] * mt[1] + mt[2] * mt[2]) / rhof / (CA * CA);
    double perp = 1 / (1 + alf);
    double dva = alf * (-dt * dU[0]);
    double vt[3];
#pragma unroll
    for (int m = 0; m < 3; m++)
      vt[m] = alf * (-dt * dU[1 + m] + dt * SH[m] - (-dt * dU[0]) * x[1 + m]);
    double bdv = mt[0] * vt[0] + mt[1] * vt[1] + mt[2] * vt[2];
    bdv = bdv / (mt[0] * mt[0] + mt[1] * mt[1] + mt[2] * mt[2] + 1e-10);
    double vn[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
      double rv = ru0[m] + perp * ((-dt * dU[1 + m] + dt * SH[m]) + (-dt * dU[6 + m] + dt * SB[m]) +
                                   dva * x[1 + m] + mt[m] * bdv);
      rv += dt * rho * a.grav[m];  // gravity (src/solver_solve.c:469-471)
      vn[m] = rv / rho;
    }
    double out[6] = {rhof, vn[0], vn[1], vn[2], p, ps};
#pragma unroll
    for (int m = 0; m < 6; m++) {
      st_nt(a.Xn + m * g.fs + c, out[m]);
      bad |= !isfinite(out[m]);
    }
    if (a.Xn32) {
#pragma unroll
      for (int m = 0; m < 6; m++) st_nt(a.Xn32 + m * g.fs + c, (float)out[m]);
    }
    // cell-centred B at n+1 (the ghost cells' is formed after the ghost fill)
#pragma unroll
    for (int m = 0; m < 3; m++) {
      if (a.Mn) st_nt(a.Mn + m * g.fs + c, bnew[m]);
      if (a.Mn32) st_nt(a.Mn32 + m * g.fs + c, (float)bnew[m]);
    }
    // next step's CFL dt from the new state (magtot = cell B + B0)
    double dcell = cfl_dt(rhof, vn[0], vn[1], vn[2], p, bnew[0] + cb0[0], bnew[1] + cb0[1],
                          bnew[2] + cb0[2], dxi, dxj, dxk, CA);
    const bool red = !(j >= a.skip[0] && j <= a.skip[1]) && !(j >= a.skip[2] && j <= a.skip[3]);
    if (red && dcell == dcell) dtc = fmin(dtc, dcell);  // fmin in the reference ignores NaN
    // div B of the new face field (with the B0 face normals)
    double db = ((bf[0][1] + nb0[0][1]) * A1[0] - (bf[0][0] + nb0[0][0]) * A0[0]) / V +
                ((bf[1][1] + nb0[1][1]) * A1[1] - (bf[1][0] + nb0[1][0]) * A0[1]) / V +
                ((bf[2][1] + nb0[2][1]) * A1[2] - (bf[2][0] + nb0[2][0]) * A0[2]) / V;
    const double adb = fabs(db);
    if (adb != adb)
      bad = true;
    else if (red)
      divb = fmax(divb, adb);
  }
  block_reduce(dtc, divb, a.sc, bx * 31 + by * 7 + bz);
  if (bad) atomicOr(&a.sc->nanflag, 1u);
}

template <typename R, bool B0V>
__global__ __launch_bounds__(256) void k_cell(CellArgs a, int gx, int gy, int gz) {
  int bx, by, bz;
  (void)gy;
  by = 0;
  if (!xcd_patch(gx, gz, bx, bz)) return;
  k_cell_body<R, B0V>(a, bx, by, bz);
}

// one launch for many grids: block w runs work-list entry w (argument set, block)
template <typename R, bool B0V>
__global__ __launch_bounds__(256) void k_cell_batch(const CellArgs *__restrict__ args,
                                                    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  // uniform per block: keep the argument set in scalar registers
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  const int bx = __builtin_amdgcn_readfirstlane(work[w].bx);
  const int by = __builtin_amdgcn_readfirstlane(work[w].by);
  const int bz = __builtin_amdgcn_readfirstlane(work[w].bz);
  k_cell_body<R, B0V>(batch_args(args, arg), bx, by, bz);
}

// CFL dt and |div B| of the current state (prologue; src/solver_func.c:255-343)
struct DtArgs {
  Grid g;
  Geo geo;
  B0c b0;
  const double *b0a;   // per-point B0 (b0_mode 2), else null
  const double *X, *M, *B;
  const double *dxkj;  // j factor of the phi width in the CFL dt
  int j0, j1;          // rows reduced (blockIdx.y covers them from j0)
  Scal *sc;
};

__device__ __forceinline__ void globalize(DtArgs &a) {
  globalize(a.geo);
  gptr(a.b0a);
  gptr(a.X);
  gptr(a.M);
  gptr(a.B);
  gptr(a.dxkj);
  gptr(a.sc);
}

template <bool B0V>
__device__ __forceinline__ void k_dt_body(const DtArgs &a, const int bx, const int by,
                                         const int bz) {
  if (halted(a.sc)) return;
  const Grid &g = a.g;
  const Geo &G = a.geo;
  const int k = g.s[2] + bx * blockDim.x + threadIdx.x;
  const int j = a.j0 + by * blockDim.y + threadIdx.y;
  const int i = g.s[0] + bz;
  double dtc = INFINITY, divb = 0.0;
  if (k <= g.e[2] && j <= a.j1) {
    const long c = i * g.st[0] + j * g.st[1] + k;
    double b[3], nb0[3][2];
#pragma unroll
    for (int m = 0; m < 3; m++) {
      if constexpr (B0V) {
        b[m] = a.M[m * g.fs + c] + a.b0a[(B0A_CELL + m) * g.fs + c];
        nb0[m][0] = a.b0a[(B0A_FACE + 5 * m) * g.fs + c];
        nb0[m][1] = a.b0a[(B0A_FACE + 5 * m) * g.fs + c + g.st[m]];
      } else {
        b[m] = a.M[m * g.fs + c] + a.b0.cell[m];
        nb0[m][0] = nb0[m][1] = a.b0.face[m][m];
      }
    }
    // phi width: the ring-widened one on averaged rows (get_dt_ring,
    // src/spherical/spherical_general.c:700-735); the j factor carries it
    const double dxk = G.g[GK_G_DXK][0][i] * a.dxkj[j] * G.g[GK_G_DXK][2][k];
    dtc = cfl_dt(a.X[c], a.X[g.fs + c], a.X[2 * g.fs + c], a.X[3 * g.fs + c], a.X[4 * g.fs + c],
                 b[0], b[1], b[2], geo3(G, GK_G_DXI, i, j, k), geo3(G, GK_G_DXJ, i, j, k), dxk,
                 g.CA);
    if (!(dtc == dtc)) dtc = INFINITY;
    const double V = geo3(G, GK_G_VOL, i, j, k);
    const double *Bf = a.B;
    double db = ((Bf[c + g.st[0]] + nb0[0][1]) * geo3(G, GK_G_FACE_I, i + 1, j, k) -
                 (Bf[c] + nb0[0][0]) * geo3(G, GK_G_FACE_I, i, j, k)) / V +
                ((Bf[g.fs + c + g.st[1]] + nb0[1][1]) * geo3(G, GK_G_FACE_J, i, j + 1, k) -
                 (Bf[g.fs + c] + nb0[1][0]) * geo3(G, GK_G_FACE_J, i, j, k)) / V +
                ((Bf[2 * g.fs + c + 1] + nb0[2][1]) * geo3(G, GK_G_FACE_K, i, j, k + 1) -
                 (Bf[2 * g.fs + c] + nb0[2][0]) * geo3(G, GK_G_FACE_K, i, j, k)) / V;
    divb = fabs(db);
  }
  block_reduce(dtc, divb, a.sc, bx * 31 + by * 7 + bz);
}

template <bool B0V>
__global__ __launch_bounds__(256) void k_dt(DtArgs a) {
  k_dt_body<B0V>(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// one launch for many grids: block w runs work-list entry w
template <bool B0V>
__global__ __launch_bounds__(256) void k_dt_batch(const DtArgs *__restrict__ args,
    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  k_dt_body<B0V>(batch_args(args, arg), __builtin_amdgcn_readfirstlane(work[w].bx),
      __builtin_amdgcn_readfirstlane(work[w].by), __builtin_amdgcn_readfirstlane(work[w].bz));
}

// ---------------------------------------------------------------------------
// Ghost fill.  Each axis maps a ghost index to its source: periodic images
// (src/setup_mpi.c:62-253: high ghosts <- first active, low ghosts <- last
// active; a face-normal field keeps its upper boundary face, which CT
// computes), the symmetric mirror (src/boundary_conditions.c) or, on a
// theta side at a pole, the mirror across the pole rotated by pi in phi.  Applying
// the three axis maps at once equals the reference's sequential i, j, k
// passes, because each later pass copies whole rows of the earlier ones, and
// every source point is interior in all three axes, so no thread reads a
// point another thread writes.
struct GhostArgs {
  Grid g;
  double *X;    // 6 cell fields, or null (not filled)
  double *B;    // 3 face fields, or null (not filled)
  float *X32;   // FP32 copies (FP32-bulk form), else null
  float *B32;
  int bc[2][6];
  const Scal *sc;  // run control of a step's launch (null outside a step)
};

__device__ __forceinline__ void globalize(GhostArgs &a) {
  gptr(a.X);
  gptr(a.B);
  gptr(a.X32);
  gptr(a.B32);
  gptr(a.sc);
}

__device__ __forceinline__ bool mirror_kind(int kind) {
  return kind == GK_BC_SYMMETRIC || kind == GK_BC_POLE || kind == GK_BC_REFLECTIVE;
}
__device__ __forceinline__ int gmap(int x, int s, int e, int klo, int khi, int o) {
  if (x < s) {
    if (klo == GK_BC_PERIODIC) return x + (e - s + 1);
    if (mirror_kind(klo)) return 2 * s - 1 - x + o;
    if (klo == GK_BC_EXTRAPOLATED) return s;
    return x;
  }
  if (x > e + o && x <= e + 4 + o) {
    if (khi == GK_BC_PERIODIC) return x - (e - s + 1);
    if (mirror_kind(khi)) return 2 * e + 1 + o - x;
    if (khi == GK_BC_EXTRAPOLATED) return e + o;
  }
  return x;
}

// true when x is a ghost of a fixed side for a field with face offset o:
// the ghost holds constants (gas rho = 1, v = 0, p = 1; face B = 0) and p_S
// follows the maps of the other axes (the fixed axis maps x to itself)
__device__ __forceinline__ bool fixed_ghost(int x, int s, int e, int klo, int khi, int o) {
  return (x < s && klo == GK_BC_FIXED) || (x > e + o && x <= e + 4 + o && khi == GK_BC_FIXED);
}

// true when x is a ghost of a reflective side (cell fields): the velocity
// component normal to that side changes sign
__device__ __forceinline__ bool reflect_ghost(int x, int s, int e, int klo, int khi) {
  return (x < s && klo == GK_BC_REFLECTIVE) || (x > e && x <= e + 4 && khi == GK_BC_REFLECTIVE);
}

// Pole boundary (src/spherical/spherical_boundary.c:23-138): a ghost row j
// beyond a pole takes the row mirrored across it (the symmetric map above) at
// phi + pi, i.e. k -> s + (k - s + nk/2) mod nk for every k of the box, with
// the theta and phi components negated.  True when row j of a field with
// face offset o along j is such a ghost row.
__device__ __forceinline__ bool pole_row(const Grid &g, int j, int klo, int khi, int o) {
  return (j < g.s[1] && klo == GK_BC_POLE) ||
         (j > g.e[1] + o && j <= g.e[1] + 4 + o && khi == GK_BC_POLE);
}
__device__ __forceinline__ int pole_k(const Grid &g, int k) {
  const int nk = g.e[2] - g.s[2] + 1;
  return g.s[2] + (k - g.s[2] + nk / 2) % nk;
}

// The points of the box outside the active cells, walked as three slabs:
// i outside [s0, e0] (all j, k), then j outside [s1, e1] (i active, all k),
// then k outside [s2, e2] (i, j active); k fastest in each, 256 per block.
// The box is N - shrink points per axis (shrink 1 leaves out the upper face
// slot, for cell quantities formed from both faces of a cell).
__device__ __forceinline__ bool ghost_point(const Grid &g, int bx, int &i, int &j, int &k,
                                           int shrink = 0) {
  const int n0 = g.e[0] - g.s[0] + 1, n1 = g.e[1] - g.s[1] + 1, n2 = g.e[2] - g.s[2] + 1;
  const int N0 = g.N[0] - shrink, N1 = g.N[1] - shrink, N2 = g.N[2] - shrink;
  long p = (long)bx * 256 + threadIdx.y * 64 + threadIdx.x;
  const long ca = (long)(N0 - n0) * N1 * N2, cb = (long)n0 * (N1 - n1) * N2;
  const long cc = (long)n0 * n1 * (N2 - n2);
  if (p < ca) {
    k = (int)(p % N2);
    const long t = p / N2;
    j = (int)(t % N1);
    const int q = (int)(t / N1);
    i = q < g.s[0] ? q : q + n0;
    return true;
  }
  p -= ca;
  if (p < cb) {
    k = (int)(p % N2);
    const long t = p / N2;
    const int q = (int)(t % (N1 - n1));
    i = g.s[0] + (int)(t / (N1 - n1));
    j = q < g.s[1] ? q : q + n1;
    return true;
  }
  p -= cb;
  if (p < cc) {
    const int q = (int)(p % (N2 - n2));
    const long t = p / (N2 - n2);
    j = g.s[1] + (int)(t % n1);
    i = g.s[0] + (int)(t / n1);
    k = q < g.s[2] ? q : q + n2;
    return true;
  }
  return false;
}

__device__ __forceinline__ void k_ghost_body(const GhostArgs &a, const int bx, const int by,
                                        const int bz) {
  if (halted(a.sc)) return;
  (void)by;
  (void)bz;
  const Grid &g = a.g;
  int i, j, k;
  if (!ghost_point(g, bx, i, j, k)) return;
  const int x[3] = {i, j, k};
  const long c = i * g.st[0] + j * g.st[1] + k;
  // sources are active points, which this launch never writes, so every
  // value is read before the first store (a load after a store could not be
  // moved ahead of it)
  // pole ghost rows: rotated source column, v_theta, v_phi (cell fields 2, 3)
  // and B_theta, B_phi (face fields 1, 2) negated
  const bool px = pole_row(g, j, a.bc[0][2], a.bc[0][3], 0);
  bool rg[3];  // reflective ghost along axis d: normal velocity negated
  bool fx = false;  // in a fixed ghost slab: rho, v, p constant
#pragma unroll
  for (int d = 0; d < 3; d++) {
    rg[d] = reflect_ghost(x[d], g.s[d], g.e[d], a.bc[0][2 * d], a.bc[0][2 * d + 1]);
    fx = fx || fixed_ghost(x[d], g.s[d], g.e[d], a.bc[0][2 * d], a.bc[0][2 * d + 1], 0);
  }
  const long sx = gmap(i, g.s[0], g.e[0], a.bc[0][0], a.bc[0][1], 0) * g.st[0] +
                  gmap(j, g.s[1], g.e[1], a.bc[0][2], a.bc[0][3], 0) * g.st[1] +
                  (px ? pole_k(g, k) : gmap(k, g.s[2], g.e[2], a.bc[0][4], a.bc[0][5], 0));
  long sb[3];
  bool pb[3], fb[3];
#pragma unroll
  for (int f = 0; f < 3; f++) {
    int m[3];
    fb[f] = false;  // in a fixed ghost slab of this face field: 0
#pragma unroll
    for (int d = 0; d < 3; d++) {
      m[d] = gmap(x[d], g.s[d], g.e[d], a.bc[1][2 * d], a.bc[1][2 * d + 1], d == f ? 1 : 0);
      fb[f] = fb[f] || fixed_ghost(x[d], g.s[d], g.e[d], a.bc[1][2 * d], a.bc[1][2 * d + 1],
                                   d == f ? 1 : 0);
    }
    pb[f] = pole_row(g, j, a.bc[1][2], a.bc[1][3], f == 1 ? 1 : 0);
    if (pb[f]) m[2] = pole_k(g, k);
    sb[f] = m[0] * g.st[0] + m[1] * g.st[1] + m[2];
  }
  double xv[6], bv[3];
  float xw[6], bw[3];
  if (sx != c) {
    if (a.X) {
#pragma unroll
      for (int m = 0; m < 6; m++) xv[m] = a.X[m * g.fs + sx];
      if (px) {
        xv[2] = -xv[2];
        xv[3] = -xv[3];
      }
#pragma unroll
      for (int d = 0; d < 3; d++)
        if (rg[d]) xv[1 + d] = -xv[1 + d];
    }
    if (a.X32) {
#pragma unroll
      for (int m = 0; m < 6; m++) xw[m] = a.X32[m * g.fs + sx];
      if (px) {
        xw[2] = -xw[2];
        xw[3] = -xw[3];
      }
#pragma unroll
      for (int d = 0; d < 3; d++)
        if (rg[d]) xw[1 + d] = -xw[1 + d];
    }
  }
#pragma unroll
  for (int f = 0; f < 3; f++) {
    if (sb[f] != c) {
      const bool neg = pb[f] && f != 0;
      if (a.B) bv[f] = neg ? -a.B[f * g.fs + sb[f]] : a.B[f * g.fs + sb[f]];
      if (a.B32) bw[f] = neg ? -a.B32[f * g.fs + sb[f]] : a.B32[f * g.fs + sb[f]];
    }
  }
  if (fx) {  // fixed: the constants replace rho, v, p; p_S as mapped
    const double fixv[5] = {1.0, 0.0, 0.0, 0.0, 1.0};
#pragma unroll
    for (int m = 0; m < 5; m++) {
      xv[m] = fixv[m];
      xw[m] = (float)fixv[m];
    }
  }
  if (sx != c || fx) {
    const int nm = sx != c ? 6 : 5;  // p_S mapped to itself is left alone
    if (a.X) {
#pragma unroll
      for (int m = 0; m < 6; m++)
        if (m < nm) a.X[m * g.fs + c] = xv[m];
    }
    if (a.X32) {
#pragma unroll
      for (int m = 0; m < 6; m++)
        if (m < nm) a.X32[m * g.fs + c] = xw[m];
    }
  }
#pragma unroll
  for (int f = 0; f < 3; f++) {
    if (fb[f]) {
      bv[f] = 0.0;
      bw[f] = 0.0f;
    }
    if (sb[f] != c || fb[f]) {
      if (a.B) a.B[f * g.fs + c] = bv[f];
      if (a.B32) a.B32[f * g.fs + c] = bw[f];
    }
  }
}

__global__ __launch_bounds__(256) void k_ghost(GhostArgs a, int gx, int gy, int gz) {
  int bx, by, bz;
  if (!xcd_block<true>(gx, gy, gz, bx, by, bz)) return;
  k_ghost_body(a, bx, by, bz);
}

// one launch for many grids: block w runs work-list entry w (argument set, block)
__global__ __launch_bounds__(256) void k_ghost_batch(const GhostArgs *__restrict__ args,
                                                    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  // uniform per block: keep the argument set in scalar registers
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  const int bx = __builtin_amdgcn_readfirstlane(work[w].bx);
  const int by = __builtin_amdgcn_readfirstlane(work[w].by);
  const int bz = __builtin_amdgcn_readfirstlane(work[w].bz);
  k_ghost_body(batch_args(args, arg), bx, by, bz);
}

// cell-centred B over cells 0..N-2 of each axis (get_derived_variables,
// src/solver_func.c:310-322)
struct MagsArgs {
  Grid g;
  Geo geo;
  const double *B;
  double *M;   // or null (not written)
  float *M32;  // FP32 copy (FP32-bulk form), else null
  const Scal *sc;  // run control of a step's launch (null outside a step)
};

__device__ __forceinline__ void globalize(MagsArgs &a) {
  globalize(a.geo);
  gptr(a.B);
  gptr(a.M);
  gptr(a.M32);
  gptr(a.sc);
}

template <bool SHELL = false>
__device__ __forceinline__ void k_mags_body(const MagsArgs &a, const int bx, const int by,
                                       const int bz) {
  if (halted(a.sc)) return;
  const Grid &g = a.g;
  (void)by;
  int i, j, k;
  if (SHELL) {  // only the points outside the active cells (the cell update forms the rest)
    (void)bz;
    if (!ghost_point(g, bx, i, j, k, 1)) return;
  } else {
    if (!row_point(0, g.N[2] - 2, 0, g.N[1] - 2, bx, k, j)) return;
    i = bz;
  }
  const long c = i * g.st[0] + j * g.st[1] + k;
  const int idx[3] = {i, j, k};
  // all three components before the first store (a load after a store
  // could not be moved ahead of it)
  double b[3];
#pragma unroll
  for (int m = 0; m < 3; m++) b[m] = cell_b(a.B + m * g.fs, a.geo, c, g.st[m], idx[m], m);
#pragma unroll
  for (int m = 0; m < 3; m++) {
    if (a.M) a.M[m * g.fs + c] = b[m];
    if (a.M32) a.M32[m * g.fs + c] = (float)b[m];
  }
}

// reference layout (pitch NK) <-> padded device layout (pitch KP)
__global__ __launch_bounds__(256) void k_pitch(const double *__restrict__ src,
                                               double *__restrict__ dst, long rows, int nk,
                                               int src_pitch, int dst_pitch) {
  const long n = rows * nk;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const long r = q / nk;
    const int k = (int)(q - r * nk);
    dst[r * dst_pitch + k] = src[r * src_pitch + k];
  }
}

// FP32 copy of n FP64 values (FP32-bulk form set-up)
__global__ __launch_bounds__(256) void k_to32(const double *__restrict__ src,
                                              float *__restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    dst[q] = (float)src[q];
}

__global__ __launch_bounds__(256) void k_mags(MagsArgs a, int gx, int gy, int gz) {
  int bx, by, bz;
  (void)gy;
  by = 0;
  if (!xcd_patch(gx, gz, bx, bz)) return;
  k_mags_body(a, bx, by, bz);
}

// cell-centred B of the points outside the active cells (a step's cell
// update writes the active ones)
__global__ __launch_bounds__(256) void k_mags_shell(MagsArgs a, int gx, int gy, int gz) {
  int bx, by, bz;
  if (!xcd_block<false>(gx, gy, gz, bx, by, bz)) return;
  k_mags_body<true>(a, bx, by, bz);
}

// one launch for many grids: block w runs work-list entry w (argument set, block)
__global__ __launch_bounds__(256) void k_mags_batch(const MagsArgs *__restrict__ args,
                                                    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  // uniform per block: keep the argument set in scalar registers
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  const int bx = __builtin_amdgcn_readfirstlane(work[w].bx);
  const int by = __builtin_amdgcn_readfirstlane(work[w].by);
  const int bz = __builtin_amdgcn_readfirstlane(work[w].bz);
  k_mags_body<true>(batch_args(args, arg), bx, by, bz);
}

// ---------------------------------------------------------------------------
// Pole ring averaging (spherical grids whose theta range reaches a pole;
// src/spherical/spherical_RingAverage.c, helpers src/solver_func.c:2136-2257;
// called from src/solver_solve.c:258-260, 308-310, 486-491).  The kernels
// work on the FP64 new state in place (the reference's single-rank gather /
// broadcast copies are identities); in the FP32-bulk form they also refresh
// the FP32 copies of what they change, and the edge fields they read or
// write are the form's (type R).  Every serial sum runs in one
// lane in the reference's order, so results match the reference bit for bit;
// the work is a few rows per pole, negligible beside the step.
constexpr int RING_N = 256;   // azimuthal cells of a ring row, at most
constexpr int RING_B = 128;   // blocks of a ring row, at most

struct RingArgs {
  Grid g;
  Geo geo;
  double *X;          // new primitives (6 cell fields)
  double *B;          // new face B (3 face fields)
  double *M;          // new cell-centred B
  void *E;            // edge fields of type R: E_i, E_j, E_k (ring scratch after CT)
  float *X32, *B32, *M32;  // FP32 copies of X, B, M (FP32-bulk form), else null
  const double *cs;   // cos and sin of the ring angles: cs[k], cs[nk + k]
  int nk, rows;
  int pole[2];        // theta = 0 / theta = pi is a pole
  int nchunk[16];     // blocks of ring row r (0 at the pole)
  const Scal *sc;     // run control
};

__device__ __forceinline__ double ring_fsign(double x) {
  return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0);
}
__device__ __forceinline__ double ring_minmod(double x, double y) {
  return 0.5 * (ring_fsign(x) + ring_fsign(y)) * fmin(fabs(x), fabs(y));
}

// limited PPM edge values of the middle block (src/solver_func.c:2148-2185)
__device__ __forceinline__ void ring_ppm(double fm2, double fm1, double f, double fp1, double fp2,
                                         double &vm, double &vp) {
  const double dvm2 = fm1 - fm2, dvm1 = f - fm1, dvp1 = fp1 - f, dvp2 = fp2 - fp1;
  double dvc = 0.5 * (dvm2 + dvm1);
  double SM = 2.0 * ring_minmod(dvm2, dvm1);
  const double Sm1 = ring_minmod(dvc, SM);
  dvc = 0.5 * (dvm1 + dvp1);
  SM = 2.0 * ring_minmod(dvm1, dvp1);
  const double Sp1 = ring_minmod(dvc, SM);
  dvc = 0.5 * (dvp2 + dvp1);
  SM = 2.0 * ring_minmod(dvp2, dvp1);
  const double Sp2 = ring_minmod(dvc, SM);
  vp = 0.5 * (f + fp1) - (Sp2 - Sp1) / 6.0;
  vm = 0.5 * (f + fm1) - (Sp1 - Sm1) / 6.0;
  double ap = vp - f, am = vm - f;
  if (ap * am >= 0.0) {
    ap = 0.0;
    am = 0.0;
  } else {
    if (fabs(ap) >= 2.0 * fabs(am)) ap = -2.0 * am;
    if (fabs(am) >= 2.0 * fabs(ap)) am = -2.0 * ap;
  }
  vp = f + ap;
  vm = f + am;
}

// Ring rows, one wave per (task, pole, row, i).  Tasks 0..4: primitive
// rho, v1, v2, v3, p of cell row r (HydroRingAverage_base,
// src/spherical/spherical_RingAverage.c:237-377); task 5: the B_j flux of
// the face row beyond cell row r, giving E_i; task 6: the B_i flux of cell
// row r, giving E_j (MagneticRingAverage_base, :380-547).  Velocities and
// fluxes keep their m = 0, 1 Fourier part (FilterBaseMode) and block-average
// the rest (RingBlockAvg: block means, periodic PPM edges, the parabola
// integrated over each cell).  The flux tasks write the change as edge
// fields, cumulative sums of the flux differences inside each block, zero on
// block boundaries, over k = ks .. ke+1.
template <typename R>
__device__ __forceinline__ void k_ring_rows_body(const RingArgs &a, const int bx, const int by,
                                         const int bz) {
  if (halted(a.sc)) return;
  __shared__ double q[RING_N];
  __shared__ double blk[RING_B];
  __shared__ double co[3];
  const Grid &g = a.g;
  const int lane = threadIdx.x;
  const int task = bz % 7, p = bz / 7;
  const int r = by;
  const int i = g.s[0] + bx;
  if (!a.pole[p] || i > g.e[0] + (task == 6 ? 1 : 0)) return;  // uniform per block
  const int N = a.nk, ks = g.s[2];
  const int nch = a.nchunk[r], ncell = N / nch;
  const int jc = p == 0 ? g.s[1] + r : g.e[1] - r;
  const int j = task == 5 ? (p == 0 ? jc + 1 : jc) : jc;
  const long row = i * g.st[0] + j * g.st[1] + ks;
  const bool flux = task >= 5;
  const bool filt = flux || (task >= 1 && task <= 3);
  const double *cs = a.cs, *sn = a.cs + N;
  double *xr = flux ? nullptr : a.X + (long)task * g.fs + row;
  for (int k = lane; k < N; k += 64) {
    if (task == 5)
      q[k] = a.B[g.fs + row + k] * geo3(a.geo, GK_G_FACE_J, i, j, ks + k);
    else if (task == 6)
      q[k] = a.B[row + k] * geo3(a.geo, GK_G_FACE_I, i, j, ks + k);
    else
      q[k] = xr[k];
  }
  __syncthreads();
  if (filt) {  // FilterBaseMode (src/solver_func.c:2187-2212)
    if (lane == 0) {
      const double scl = 2.0 / N;
      double a0 = 0.0, a1 = 0.0, b1 = 0.0;
      for (int k = 0; k < N; k++) {
        a0 += q[k];
        a1 += q[k] * cs[k];
        b1 += q[k] * sn[k];
      }
      a1 *= scl;
      b1 *= scl;
      a0 *= 0.5 * scl;
      co[0] = a0;
      co[1] = a1;
      co[2] = b1;
    }
    __syncthreads();
    for (int k = lane; k < N; k += 64) q[k] -= co[0] + co[1] * cs[k] + co[2] * sn[k];
    __syncthreads();
  }
  for (int b = lane; b < nch; b += 64) {  // block means
    double sum = 0.0;
    for (int c = 0; c < ncell; c++) sum += q[b * ncell + c];
    blk[b] = sum / ncell;
  }
  __syncthreads();
  R *E = (R *)a.E + (task == 5 ? 0 : g.fs) + row;
  float *xr32 = (flux || !a.X32) ? nullptr : a.X32 + (long)task * g.fs + row;
  for (int b = lane; b < nch; b += 64) {
    const double e0 = blk[(b + nch - 2) % nch], e1 = blk[(b + nch - 1) % nch], e2 = blk[b];
    const double e3 = blk[(b + 1) % nch], e4 = blk[(b + 2) % nch];
    double fL, fR;
    ring_ppm(e0, e1, e2, e3, e4, fL, fR);
    const double A = 3.0 * (fL + fR - 2.0 * e2);
    const double Bc = 2.0 * (3.0 * e2 - fR - 2.0 * fL);
    const double C = fL;
    const int k0 = b * ncell;
    const double A3 = A / 3.0 / ncell / ncell;  // the same for every cell of the block
    double sum = 0.0;
    if (flux) E[k0] = (R)0.0;
    for (int c = 0; c < ncell; c++) {
      const double kl = (double)c + 1.0;
      const double v = A3 * (3 * kl * kl - 3 * kl + 1) + 0.5 * Bc * (2 * kl - 1) / ncell + C;
      const int k = k0 + c;
      if (!flux) {
        const double out = filt ? v + (co[0] + co[1] * cs[k] + co[2] * sn[k]) : v;
        xr[k] = out;
        if (xr32) xr32[k] = (float)out;
      } else if (c < ncell - 1) {
        sum += v - q[k];
        E[k + 1] = (R)(task == 5 ? -sum : sum);
      }
    }
  }
  if (flux && lane == 0) E[N] = (R)0.0;
}

template <typename R>
__global__ __launch_bounds__(64) void k_ring_rows(RingArgs a) {
  k_ring_rows_body<R>(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// one launch for many grids: block w runs work-list entry w
template <typename R>
__global__ __launch_bounds__(64) void k_ring_rows_batch(const RingArgs *__restrict__ args,
    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  k_ring_rows_body<R>(args[arg], __builtin_amdgcn_readfirstlane(work[w].bx),
      __builtin_amdgcn_readfirstlane(work[w].by), __builtin_amdgcn_readfirstlane(work[w].bz));
}

// Mean of the n values of a row, summed in order by one lane of the wave
// (the reference's serial loop); the row is staged through LDS by the wave.
template <typename T>
__device__ __forceinline__ double ring_mean(const T *v, int n, double *lds) {
  for (int k = threadIdx.x; k < n; k += 64) lds[k] = (double)v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int k = 0; k < n; k++) m = m + lds[k];
    lds[n] = m / n;
  }
  __syncthreads();
  return lds[n];
}

// E and B pole fixes (src/spherical/spherical_RingAverage.c:153-235): before
// CT, E_i on the pole edges := its mean over the nk + 1 edges of the ring and
// E_k := 0; after CT, B_j on the pole face := half the difference of the next
// face row and its antipode in phi.  One wave per (pole, i).
template <typename R>
__device__ __forceinline__ void k_pole_e_body(const RingArgs &a, const int bx, const int by,
                                         const int bz) {
  if (halted(a.sc)) return;
  (void)bz;  // 2-D launch
  __shared__ double lds[RING_N + 2];
  const Grid &g = a.g;
  const int p = by;
  const int i = g.s[0] + bx;
  if (!a.pole[p] || i > g.e[0] + 1) return;  // uniform per block
  const int jp = p == 0 ? g.s[1] : g.e[1] + 1;
  const long row = i * g.st[0] + jp * g.st[1];
  if (i <= g.e[0]) {
    R *Ei = (R *)a.E + row + g.s[2];
    const double m = ring_mean(Ei, a.nk + 1, lds);
    for (int k = threadIdx.x; k <= a.nk; k += 64) Ei[k] = (R)m;
  }
  R *Ek = (R *)a.E + 2 * g.fs + row + g.s[2];
  for (int k = threadIdx.x; k < a.nk; k += 64) Ek[k] = (R)0.0;
}

template <typename R>
__global__ __launch_bounds__(64) void k_pole_e(RingArgs a) {
  k_pole_e_body<R>(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// one launch for many grids: block w runs work-list entry w
template <typename R>
__global__ __launch_bounds__(64) void k_pole_e_batch(const RingArgs *__restrict__ args,
    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  k_pole_e_body<R>(args[arg], __builtin_amdgcn_readfirstlane(work[w].bx),
      __builtin_amdgcn_readfirstlane(work[w].by), __builtin_amdgcn_readfirstlane(work[w].bz));
}

__device__ __forceinline__ void k_pole_b_body(const RingArgs &a, const int bx, const int by,
                                         const int bz) {
  if (halted(a.sc)) return;
  (void)bz;  // 2-D launch
  const Grid &g = a.g;
  const int p = by;
  const long t = (long)bx * 256 + threadIdx.x;
  const int ni = g.e[0] - g.s[0] + 1;
  if (!a.pole[p] || t >= (long)ni * a.nk) return;
  const int i = g.s[0] + (int)(t / a.nk), k = g.s[2] + (int)(t % a.nk);
  const int jp = p == 0 ? g.s[1] : g.e[1] + 1, jn = p == 0 ? g.s[1] + 1 : g.e[1];
  double *Bj = a.B + g.fs + i * g.st[0];
  const double b = 0.5 * (Bj[jn * g.st[1] + k] - Bj[jn * g.st[1] + pole_k(g, k)]);
  Bj[jp * g.st[1] + k] = b;
  if (a.B32) a.B32[g.fs + i * g.st[0] + jp * g.st[1] + k] = (float)b;
}

__global__ __launch_bounds__(256) void k_pole_b(RingArgs a) {
  k_pole_b_body(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// one launch for many grids: block w runs work-list entry w
__global__ __launch_bounds__(256) void k_pole_b_batch(const RingArgs *__restrict__ args,
    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  k_pole_b_body(args[arg], __builtin_amdgcn_readfirstlane(work[w].bx),
      __builtin_amdgcn_readfirstlane(work[w].by), __builtin_amdgcn_readfirstlane(work[w].bz));
}

// ring edge fields at (i, j, k): E_i on the face rows beyond the averaged
// cell rows, E_j on the averaged cell rows, zero elsewhere
__device__ __forceinline__ bool ring_face_row(const RingArgs &a, int j) {
  const Grid &g = a.g;
  return (a.pole[0] && j >= g.s[1] + 1 && j <= g.s[1] + a.rows) ||
         (a.pole[1] && j >= g.e[1] - a.rows + 1 && j <= g.e[1]);
}
__device__ __forceinline__ bool ring_cell_row(const RingArgs &a, int j) {
  const Grid &g = a.g;
  return (a.pole[0] && j >= g.s[1] && j <= g.s[1] + a.rows - 1) ||
         (a.pole[1] && j >= g.e[1] - a.rows + 1 && j <= g.e[1]);
}
template <typename R>
__device__ __forceinline__ double ring_ei(const RingArgs &a, int i, int j, int k) {
  const Grid &g = a.g;
  if (i > g.e[0] || !ring_face_row(a, j)) return 0.0;
  return (double)((const R *)a.E)[i * g.st[0] + j * g.st[1] + k];
}
template <typename R>
__device__ __forceinline__ double ring_ej(const RingArgs &a, int i, int j, int k) {
  const Grid &g = a.g;
  if (!ring_cell_row(a, j)) return 0.0;
  return (double)((const R *)a.E)[g.fs + i * g.st[0] + j * g.st[1] + k];
}

// Curl update of the ring edge fields (E_k is zero), over the rows they
// reach: cell rows s..s+rows at theta = 0, e-rows..e at theta = pi
// (src/spherical/spherical_RingAverage.c:584-615; elsewhere the reference's
// update subtracts 0 / A and leaves B as it is).  The pole face itself is
// skipped: its area is zero and the next kernel overwrites it.
template <typename R>
__device__ __forceinline__ void k_ring_curl_body(const RingArgs &a, const int bx, const int by,
                                         const int bz) {
  if (halted(a.sc)) return;
  (void)bz;  // 2-D launch
  const Grid &g = a.g;
  const int p = by;
  const int R1 = a.rows + 1, NI1 = g.e[0] - g.s[0] + 2, NK1 = a.nk + 1;
  const long t = (long)bx * 256 + threadIdx.x;
  if (!a.pole[p] || t >= (long)NI1 * R1 * NK1) return;
  const int k = g.s[2] + (int)(t % NK1);
  const long u = t / NK1;
  const int jj = (int)(u % R1);
  const int i = g.s[0] + (int)(u / R1);
  const int j = p == 0 ? g.s[1] + jj : g.e[1] - a.rows + jj;
  const int jp = p == 0 ? g.s[1] : g.e[1] + 1;
  const long c = i * g.st[0] + j * g.st[1] + k;
  const bool ki = k <= g.e[2], ii = i <= g.e[0];
  // every input before the first store
  const double ej = ring_ej<R>(a, i, j, k);
  const double ej_k = ki ? ring_ej<R>(a, i, j, k + 1) : 0.0;
  const double ej_i = ii ? ring_ej<R>(a, i + 1, j, k) : 0.0;
  const double ei = ring_ei<R>(a, i, j, k);
  const double ei_k = ki ? ring_ei<R>(a, i, j, k + 1) : 0.0;
  const double ei_j = ring_ei<R>(a, i, j + 1, k);
  const double bi = a.B[c], bj = a.B[g.fs + c], bk = a.B[2 * g.fs + c];
  if (ki) {  // B_i, faces i = is..ie+1
    const double cE = 1.0 * ((0.0 - 0.0) - (ej_k - ej));
    const double b = bi - cE / geo3(a.geo, GK_G_FACE_I, i, j, k);
    a.B[c] = b;
    if (a.B32) a.B32[c] = (float)b;
  }
  if (ii && ki && j != jp) {  // B_j
    const double cE = 1.0 * ((ei_k - ei) - (0.0 - 0.0));
    const double b = bj - cE / geo3(a.geo, GK_G_FACE_J, i, j, k);
    a.B[g.fs + c] = b;
    if (a.B32) a.B32[g.fs + c] = (float)b;
  }
  if (ii) {  // B_k, faces k = ks..ke+1
    const double cE = 1.0 * ((ej_i - ej) - (ei_j - ei));
    const double b = bk - cE / geo3(a.geo, GK_G_FACE_K, i, j, k);
    a.B[2 * g.fs + c] = b;
    if (a.B32) a.B32[2 * g.fs + c] = (float)b;
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_ring_curl(RingArgs a) {
  k_ring_curl_body<R>(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// one launch for many grids: block w runs work-list entry w
template <typename R>
__global__ __launch_bounds__(256) void k_ring_curl_batch(const RingArgs *__restrict__ args,
    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  k_ring_curl_body<R>(args[arg], __builtin_amdgcn_readfirstlane(work[w].bx),
      __builtin_amdgcn_readfirstlane(work[w].by), __builtin_amdgcn_readfirstlane(work[w].bz));
}

// After the curl update (src/spherical/spherical_RingAverage.c:617-677): the
// pole face B_j := the next face row; B_k of the pole cell row minus its mean
// over the nk + 1 faces of the ring.  One wave per (pole, i).
__device__ __forceinline__ void k_ring_fix_body(const RingArgs &a, const int bx, const int by,
                                         const int bz) {
  if (halted(a.sc)) return;
  (void)bz;  // 2-D launch
  __shared__ double lds[RING_N + 2];
  const Grid &g = a.g;
  const int p = by;
  const int i = g.s[0] + bx;
  if (!a.pole[p] || i > g.e[0]) return;  // uniform per block
  const int jp = p == 0 ? g.s[1] : g.e[1] + 1, jn = p == 0 ? g.s[1] + 1 : g.e[1];
  const int jr = p == 0 ? g.s[1] : g.e[1];
  const long oj = g.fs + i * g.st[0] + g.s[2];
  const long ok = 2 * g.fs + i * g.st[0] + jr * g.st[1] + g.s[2];
  double *Bj = a.B + oj;
  for (int k = threadIdx.x; k < a.nk; k += 64) {
    const double b = Bj[jn * g.st[1] + k];
    Bj[jp * g.st[1] + k] = b;
    if (a.B32) a.B32[oj + jp * g.st[1] + k] = (float)b;
  }
  double *Bk = a.B + ok;
  const double m = ring_mean(Bk, a.nk + 1, lds);
  for (int k = threadIdx.x; k <= a.nk; k += 64) {
    const double b = lds[k] - m;
    Bk[k] = b;
    if (a.B32) a.B32[ok + k] = (float)b;
  }
}

__global__ __launch_bounds__(64) void k_ring_fix(RingArgs a) {
  k_ring_fix_body(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// one launch for many grids: block w runs work-list entry w
__global__ __launch_bounds__(64) void k_ring_fix_batch(const RingArgs *__restrict__ args,
    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  k_ring_fix_body(args[arg], __builtin_amdgcn_readfirstlane(work[w].bx),
      __builtin_amdgcn_readfirstlane(work[w].by), __builtin_amdgcn_readfirstlane(work[w].bz));
}

// cell-centred B of the active cells whose faces the ring average changed
// (cell rows s..s+rows and e-rows..e), as the reference's derived-variable
// pass computes it (src/solver_func.c:310-322)
__device__ __forceinline__ void k_ring_mags_body(const RingArgs &a, const int bx, const int by,
                                         const int bz) {
  if (halted(a.sc)) return;
  (void)bz;  // 2-D launch
  const Grid &g = a.g;
  const int p = by;
  const int R1 = a.rows + 1, ni = g.e[0] - g.s[0] + 1;
  const long t = (long)bx * 256 + threadIdx.x;
  if (!a.pole[p] || t >= (long)ni * R1 * a.nk) return;
  const int k = g.s[2] + (int)(t % a.nk);
  const long u = t / a.nk;
  const int jj = (int)(u % R1);
  const int i = g.s[0] + (int)(u / R1);
  const int j = p == 0 ? g.s[1] + jj : g.e[1] - a.rows + jj;
  const long c = i * g.st[0] + j * g.st[1] + k;
  const int idx[3] = {i, j, k};
  double b[3];
#pragma unroll
  for (int m = 0; m < 3; m++) b[m] = cell_b(a.B + m * g.fs, a.geo, c, g.st[m], idx[m], m);
#pragma unroll
  for (int m = 0; m < 3; m++) {
    a.M[m * g.fs + c] = b[m];  // FP64 too in the FP32-bulk form: the band dt reads it
    if (a.M32) a.M32[m * g.fs + c] = (float)b[m];
  }
}

__global__ __launch_bounds__(256) void k_ring_mags(RingArgs a) {
  k_ring_mags_body(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// one launch for many grids: block w runs work-list entry w
__global__ __launch_bounds__(256) void k_ring_mags_batch(const RingArgs *__restrict__ args,
    const WorkItem *__restrict__ work, int nwork) {
  int w;
  if (!xcd_item(nwork, w)) return;
  const int arg = __builtin_amdgcn_readfirstlane(work[w].arg);
  k_ring_mags_body(args[arg], __builtin_amdgcn_readfirstlane(work[w].bx),
      __builtin_amdgcn_readfirstlane(work[w].by), __builtin_amdgcn_readfirstlane(work[w].bz));
}

// plain copy, 16 bytes per lane per iteration (calibration kernel)
// one 8-byte value per thread and one block per 256 values: of the plain
// copies measured (grid-stride loops over 16-byte values with 2048 to 32768
// blocks, with and without streaming hints) the fastest, ≈6.2 TB/s on MI355X
__global__ __launch_bounds__(256) void k_copy(const double *__restrict__ src,
                                              double *__restrict__ dst, size_t n) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = src[q];
}

}  // namespace

// ===========================================================================
// host side
struct gk_solver {
  Grid g;
  Geo geo;
  B0c b0;
  int device;          // HIP device the solver was created on
  bool own_stream;     // the stream came from the pool (else the caller's)
  double *b0a;         // per-point background field (b0_mode 2), else null
  bool ut;             // one reconstruction-table class per transverse axis
  int spherical;
  int bc[2][6];
  int ni, nj, nk;
  long fs;
  hipStream_t stream;
  double *tables;      // device copy of all per-axis tables
  int *itables;        // class maps
  std::vector<void *> field_mem;  // allocations behind the field blocks below
  double *X[2];        // 6 fields each: X[cur] = X^n, X[1-cur] = X^{n-1}
  float *X32[2];       // FP32 copies of X, B, M read by the FP32-bulk stencils
  float *B32[3];
  float *M32[3];
  double *B[3];        // face B rotation: B[b_m1], B[b_0] = B^n, B[b_n]
  double *M[3];        // cell-centred B, rotating with B
  void *F[3];          // 9 face-flux fields per direction (type by precision)
  void *V[3];          // 2 face-velocity fields per direction
  void *E;             // 3 edge fields
  float *tables_f;     // FP32 copies of the reconstruction tables
  double *pin;         // pinned host bounce buffer of one such field, lazily
  long long *ctl_host; // pinned: steps taken, non-finite flag of the last run
  double *stage;       // host transfer staging field (reference layout), lazily
  int precision;       // 0 FP64, 1 FP32-bulk
  size_t esz;          // bytes per intermediate value
  Scal *sc;
  int cur, b_m1, b_0, b_n;
  int ready;
  size_t bytes;
  int timing;
  std::string time_only;  // time only this stage's launches (empty: all)
  int chunk;           // faces per thread in the marching sweeps
  int emf_chunk;       // edges per thread in the marching EMF kernels
  int emf_stage;       // bit D: E_D (D = 0, 1) over whole 64-edge k tiles staged in LDS
  double rho_floor, p_floor;  // positivity floors
  double grav[3];      // uniform gravity
  struct gk_batch *batch;  // batch that owns the stepping, if any
  long long budget;    // batch member: steps it may still take (< 0: no limit)
  int ring;            // pole ring averaging on
  int ring_pole[2];
  int ring_rows;
  int ring_nchunk[16];
  double *ring_tab;    // cos, sin of the ring angles (2 nk), then the j factor
                       // of the CFL phi width (N[1])
  const double *dxkj;  // j factor of the CFL phi width
  std::vector<std::string> names;
  std::vector<double> acc_ms;
  std::vector<int> acc_n;
  std::vector<hipEvent_t> evpool;  // timing events, reused
  std::vector<int> pend;           // kernel slot per recorded event pair
  std::vector<struct Launch> plan; // launch list of the current step
};

#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      fprintf(stderr, "gamera_op: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return GK_ERR_DEVICE;                                                    \
    }                                                                          \
  } while (0)

// Device memory of solvers and batches is stream ordered: hipMallocAsync /
// hipFreeAsync on the owner's stream, from the device's default pool, which
// keeps what is freed for reuse.  Destroying a solver or rebuilding a
// batch's tables then never waits for work that other solvers and batches
// have in flight (a plain hipFree synchronises the whole device, so the
// serving loop's retirements used to stall on the quantum running beside
// them).  Pinned host buffers are recycled for the same reason (hipHostFree
// synchronises the device too).
constexpr int MAX_DEV = 64;
static std::mutex pool_mu;
static bool pool_kept[MAX_DEV];

static void keep_pool(int dev) {
  if (dev < 0 || dev >= MAX_DEV) return;
  std::lock_guard<std::mutex> lock(pool_mu);
  if (pool_kept[dev]) return;
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  pool_kept[dev] = true;
}

// streams are recycled too: destroying one may wait for the device
static std::mutex stream_mu;
static std::vector<std::pair<int, hipStream_t>> stream_free;  // (device, stream)

static hipError_t stream_get(int dev, hipStream_t *out) {
  {
    std::lock_guard<std::mutex> lock(stream_mu);
    for (size_t q = 0; q < stream_free.size(); q++)
      if (stream_free[q].first == dev) {
        *out = stream_free[q].second;
        stream_free.erase(stream_free.begin() + q);
        return hipSuccess;
      }
  }
  return hipStreamCreateWithFlags(out, hipStreamNonBlocking);
}

static void stream_put(int dev, hipStream_t st) {
  if (!st) return;
  std::lock_guard<std::mutex> lock(stream_mu);
  stream_free.push_back({dev, st});
}

static std::mutex pin_mu;
static std::vector<std::pair<void *, size_t>> pin_free;  // recycled buffers
static std::vector<std::pair<void *, size_t>> pin_live;  // handed out
static size_t pin_cached = 0;
constexpr size_t PIN_CACHE_MAX = size_t(2) << 30;

static void *pinned_get(size_t n) {
  std::lock_guard<std::mutex> lock(pin_mu);
  int best = -1;
  for (int q = 0; q < (int)pin_free.size(); q++)
    if (pin_free[q].second >= n && (best < 0 || pin_free[q].second < pin_free[best].second))
      best = q;
  void *p = nullptr;
  size_t sz = n;
  if (best >= 0) {
    p = pin_free[best].first;
    sz = pin_free[best].second;
    pin_cached -= sz;
    pin_free.erase(pin_free.begin() + best);
  } else if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess) {
    return nullptr;
  }
  pin_live.push_back({p, sz});
  return p;
}

static void pinned_put(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> lock(pin_mu);
  for (size_t q = 0; q < pin_live.size(); q++)
    if (pin_live[q].first == p) {
      const size_t sz = pin_live[q].second;
      pin_live.erase(pin_live.begin() + q);
      if (pin_cached + sz <= PIN_CACHE_MAX) {
        pin_free.push_back({p, sz});
        pin_cached += sz;
      } else {
        (void)hipHostFree(p);
      }
      return;
    }
}

static int dev_alloc_bytes(gk_solver *s, void **p, size_t nbytes) {
  if (hipMallocAsync(p, nbytes, s->stream) != hipSuccess) return GK_ERR_NOMEM;
  s->bytes += nbytes;
  if (hipMemsetAsync(*p, 0, nbytes, s->stream) != hipSuccess) return GK_ERR_DEVICE;
  return GK_OK;
}

static int dev_alloc(gk_solver *s, double **p, size_t ndouble) {
  return dev_alloc_bytes(s, (void **)p, ndouble * sizeof(double));
}

// A block of whole fields.  Its start is placed 128 - NGH * esz bytes into
// the allocation, so that the first active cell of every row (k = NGH) begins
// a 128-byte line (the row pitch is a whole number of lines): a wave's 64
// consecutive cells of a row then span 4 lines in FP64 and 2 in FP32, not
// 5 and 3.
static int field_alloc(gk_solver *s, void **p, size_t nbytes, size_t esz) {
  void *raw = nullptr;
  const int rc = dev_alloc_bytes(s, &raw, nbytes + 128);
  if (raw) s->field_mem.push_back(raw);
  if (rc != GK_OK) return rc;
  *p = (char *)raw + (128 - NGH * esz % 128) % 128;
  return GK_OK;
}

static int copy_in(gk_solver *s, double *dst, const double *src);
static int release_bounce(gk_solver *s);

// small host -> device upload through a recycled pinned buffer (a copy from
// pageable memory can wait behind other streams' work); waits for it
static hipError_t upload(gk_solver *s, void *dst, const void *src, size_t bytes) {
  void *pin = pinned_get(bytes);
  if (!pin) return hipErrorOutOfMemory;
  memcpy(pin, src, bytes);
  hipError_t e = hipMemcpyAsync(dst, pin, bytes, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  pinned_put(pin);
  return e;
}

// Default march lengths by grid size (faces per thread in the i and j
// sweeps, edges per thread in the edge-field kernels): smaller grids need
// shorter marches, so that their columns are split over enough threads to
// fill the GPU.  From a sweep of 2..32 faces x 2..16 edges on one MI355X
// (step times against the best pair): 32^3 (2, 2) best; 128x128x8 and
// 64^3 (4, 2) within 3%; 96^3, 128^3 and 192x96x192 (16, 4) in FP64 and
// (8, 4) in FP32-bulk within 3%; 256^3 (32, 16) best.  The pair that was the
// default for every size, (32, 16), had cost 32^3 3.8x and 128x128x8 2.3x.
// gk_autotune times the candidates on the grid itself.
static void default_shapes(gk_solver *s) {
  const double n = (double)s->ni * s->nj * s->nk;
  if (n >= 1.2e7) {
    // 17 edges per thread: 257 edge positions in 15 chunks, whose launches
    // end nearer a whole number of rounds of the wave slots than 16 chunks
    // (256^3 FP64: E_i and E_j tiles 16384 -> 15360 waves, 5 rounds of
    // 3072 slots; 0.3% faster steps in both forms)
    s->chunk = 64;
    s->emf_chunk = 17;
  } else if (n >= 4e6) {
    s->chunk = 32;
    s->emf_chunk = 16;
  } else if (n >= 5e5) {
    s->chunk = s->precision == 1 ? 8 : 16;
    s->emf_chunk = 4;
  } else if (n >= 1e5) {
    s->chunk = 4;
    s->emf_chunk = 2;
  } else {
    s->chunk = 2;
    s->emf_chunk = 2;
  }
}

extern "C" int gk_create(const gk_grid_desc *d, gk_solver **out) {
  if (!d || !out) return GK_ERR_ARG;
  *out = nullptr;
  if (d->ni < 5 || d->nj < 5 || d->nk < 5) return GK_ERR_ARG;
  if (d->coord != 0 && d->coord != 2) return GK_ERR_ARG;
  if (d->precision != 0 && d->precision != 1) return GK_ERR_ARG;
  if (d->b0_mode < 0 || d->b0_mode > 2) return GK_ERR_ARG;
  static_assert(GK_B0_FIELDS == B0A_COUNT, "per-point B0 fields");
  if (d->b0_mode == 2)
    for (int q = 0; q < GK_B0_FIELDS; q++)
      if (!d->b0_field[q]) return GK_ERR_ARG;
  // floors: >= 0 and finite (0 selects the reference's initial value);
  // gravity finite
  if (!(d->rho_floor >= 0.0 && d->rho_floor < INFINITY) ||
      !(d->p_floor >= 0.0 && d->p_floor < INFINITY))
    return GK_ERR_ARG;
  for (int m = 0; m < 3; m++)
    if (!std::isfinite(d->gravity[m])) return GK_ERR_ARG;
  for (int q = 0; q < GK_G_COUNT; q++)
    for (int a = 0; a < 3; a++)
      if (!d->geom[q][a]) return GK_ERR_ARG;
  for (int a = 0; a < 3; a++) {
    if (!d->recw[a]) return GK_ERR_ARG;
    for (int b = 0; b < 3; b++)
      if (!d->cell_coef[a][b]) return GK_ERR_ARG;
    for (int b = 0; b < 2; b++)
      if (!d->cls[a][b] || d->ncls[a][b] < 1) return GK_ERR_ARG;
  }
  for (int a = 0; a < 2; a++)
    for (int n = 0; n < 6; n++) {
      if (d->bc[a][n] < 0 || d->bc[a][n] > 6) return GK_ERR_ARG;
      if (a == 1 && d->bc[a][n] == GK_BC_REFLECTIVE) return GK_ERR_ARG;  // gas only
      // the pole boundary: theta sides at a pole of a ring-averaged grid
      if (d->bc[a][n] == GK_BC_POLE &&
          (n < 2 || n > 3 || !d->ring || !d->ring_pole[n - 2]))
        return GK_ERR_ARG;
    }
  if (d->ring) {
    // spherical grids, periodic in phi, nk <= 256 azimuthal cells; the
    // averaged rows of the two poles must not meet
    if (d->coord != 2) return GK_ERR_ARG;
    if (d->bc[0][4] != GK_BC_PERIODIC || d->bc[1][4] != GK_BC_PERIODIC) return GK_ERR_ARG;
    if (d->nk < 8 || d->nk > 256 || d->nk % 2 || d->ring_rows < 1 || d->ring_rows > 16)
      return GK_ERR_ARG;
    for (int r = 0; r < d->ring_rows; r++)
      if (d->ring_nchunk[r] < 2 || d->ring_nchunk[r] > 128 || d->nk % d->ring_nchunk[r])
        return GK_ERR_ARG;
    const int np = (d->ring_pole[0] ? 1 : 0) + (d->ring_pole[1] ? 1 : 0);
    if (d->nj < np * (d->ring_rows + 1) + 1) return GK_ERR_ARG;
  }
  for (int dim = 0; dim < 3; dim++) {  // periodicity is a property of the axis
    bool p0 = d->bc[0][2 * dim] == GK_BC_PERIODIC, p1 = d->bc[0][2 * dim + 1] == GK_BC_PERIODIC;
    bool q0 = d->bc[1][2 * dim] == GK_BC_PERIODIC, q1 = d->bc[1][2 * dim + 1] == GK_BC_PERIODIC;
    if (p0 != p1 || q0 != q1 || p0 != q0) return GK_ERR_ARG;
  }

  gk_solver *s = new gk_solver();
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return GK_ERR_DEVICE;
  }
  keep_pool(s->device);
  // every upload of the set-up runs on the solver's own stream, so creating
  // a solver does not wait for work other solvers or batches have in flight
  s->own_stream = d->stream == nullptr;
  if (!s->own_stream)
    s->stream = (hipStream_t)d->stream;
  else if (stream_get(s->device, &s->stream) != hipSuccess) {
    delete s;
    return GK_ERR_DEVICE;
  }
  s->ni = d->ni;
  s->nj = d->nj;
  s->nk = d->nk;
  const int n[3] = {d->ni, d->nj, d->nk};
  for (int a = 0; a < 3; a++) {
    s->g.N[a] = n[a] + 2 * NGH + 1;
    s->g.s[a] = NGH;
    s->g.e[a] = NGH + n[a] - 1;
  }
  // row pitch: a whole number of 128-byte lines in the arithmetic type
  const long KA = d->precision == 1 ? 32 : 16;
  const long KP = (s->g.N[2] + KA - 1) / KA * KA;
  s->g.st[2] = 1;
  s->g.st[1] = KP;
  s->g.st[0] = KP * s->g.N[1];
  s->g.fs = s->g.st[0] * s->g.N[0];
  s->g.CA = d->CA;
  s->fs = s->g.fs;
  s->spherical = d->coord == 2;
  s->precision = d->precision;
  s->esz = d->precision == 1 ? sizeof(float) : sizeof(double);
  memcpy(s->bc, d->bc, sizeof(s->bc));
  s->rho_floor = d->rho_floor > 0.0 ? d->rho_floor : RHO_FLOOR;
  s->p_floor = d->p_floor > 0.0 ? d->p_floor : P_FLOOR;
  for (int m = 0; m < 3; m++) s->grav[m] = d->gravity[m];
  s->budget = -1;
  s->b0 = B0c{};
  if (d->b0_mode == 1) {
    memcpy(s->b0.cell, d->b0_cell, sizeof(s->b0.cell));
    memcpy(s->b0.face, d->b0_face, sizeof(s->b0.face));
    memcpy(s->b0.edge, d->b0_edge, sizeof(s->b0.edge));
    memcpy(s->b0.lor, d->b0_lorentz, sizeof(s->b0.lor));
  }

  // pack the per-axis tables into one device buffer (class maps in a second)
  std::vector<double> h;
  std::vector<int> hi;
  size_t off_g[GK_G_COUNT][3], off_cc[3][3], off_w[3], off_c[3][2];
  for (int q = 0; q < GK_G_COUNT; q++)
    for (int a = 0; a < 3; a++) {
      off_g[q][a] = h.size();
      h.insert(h.end(), d->geom[q][a], d->geom[q][a] + s->g.N[a]);
    }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      off_cc[a][b] = h.size();
      h.insert(h.end(), d->cell_coef[a][b], d->cell_coef[a][b] + s->g.N[a]);
    }
  bool cls_ok = true;
  for (int a = 0; a < 3; a++) {
    off_w[a] = h.size();
    const size_t npair = (size_t)d->ncls[a][0] * d->ncls[a][1];
    const size_t nf = s->g.N[a];
    if (a < 2) {  // face-major, as given
      h.insert(h.end(), d->recw[a], d->recw[a] + nf * npair * GK_RECW_STRIDE);
      s->geo.tfs[a] = (long)npair * GK_RECW_STRIDE;
      s->geo.tps[a] = GK_RECW_STRIDE;
    } else {      // class-major
      for (size_t p = 0; p < npair; p++)
        for (size_t f = 0; f < nf; f++) {
          const double *row = d->recw[a] + (f * npair + p) * GK_RECW_STRIDE;
          h.insert(h.end(), row, row + GK_RECW_STRIDE);
        }
      s->geo.tfs[a] = GK_RECW_STRIDE;
      s->geo.tps[a] = (long)nf * GK_RECW_STRIDE;
    }
    const int tb = a == 0 ? 1 : 0, tc = a == 2 ? 1 : 2;
    const int tax[2] = {tb, tc};
    for (int q = 0; q < 2; q++) {
      off_c[a][q] = hi.size();
      for (int t = 0; t < s->g.N[tax[q]]; t++) {
        int v = d->cls[a][q][t];
        if (v < 0 || v >= d->ncls[a][q]) cls_ok = false;
        hi.push_back(v);
      }
      s->geo.ncls[a][q] = d->ncls[a][q];
    }
  }
  s->ut = true;
  for (int a = 0; a < 3; a++)
    for (int q = 0; q < 2; q++) s->ut = s->ut && d->ncls[a][q] == 1;
  int rc = GK_OK;
  auto fail = [&](int code) {
    gk_destroy(s);
    return code;
  };
  if (!cls_ok) return fail(GK_ERR_ARG);
  if ((rc = dev_alloc(s, &s->tables, h.size())) != GK_OK) return fail(rc);
  if (upload(s, s->tables, h.data(), h.size() * sizeof(double)) != hipSuccess)
    return fail(GK_ERR_DEVICE);
  if (hipMallocAsync((void **)&s->itables, hi.size() * sizeof(int), s->stream) != hipSuccess)
    return fail(GK_ERR_NOMEM);
  if (upload(s, s->itables, hi.data(), hi.size() * sizeof(int)) != hipSuccess)
    return fail(GK_ERR_DEVICE);
  for (int q = 0; q < GK_G_COUNT; q++)
    for (int a = 0; a < 3; a++) s->geo.g[q][a] = s->tables + off_g[q][a];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) s->geo.cc[a][b] = s->tables + off_cc[a][b];
  for (int a = 0; a < 3; a++) {
    s->geo.recw[a] = s->tables + off_w[a];
    s->geo.cls[a][0] = s->itables + off_c[a][0];
    s->geo.cls[a][1] = s->itables + off_c[a][1];
    s->geo.recwf[a] = nullptr;
  }
  // ring tables: cos and sin of the ring angles, evaluated on the host as the
  // reference does (src/solver_func.c:2187-2212), and the j factor of the
  // CFL phi width, widened nk / nchunk times on averaged rows
  // (src/spherical/spherical_RingAverage.c:118-147; powers of two, so the
  // product equals the reference's scaled width exactly)
  std::vector<double> hr;  // kept until the final stream synchronisation
  s->dxkj = s->geo.g[GK_G_DXK][1];
  if (d->ring) {
    s->ring = 1;
    s->ring_pole[0] = d->ring_pole[0] ? 1 : 0;
    s->ring_pole[1] = d->ring_pole[1] ? 1 : 0;
    s->ring_rows = d->ring_rows;
    memcpy(s->ring_nchunk, d->ring_nchunk, sizeof(s->ring_nchunk));
    const int nk = d->nk;
    const double dth = 2.0 * 3.14159265358979323846 / nk;
    hr.resize(2 * (size_t)nk + s->g.N[1]);
    for (int k = 0; k < nk; k++) {
      const double th = dth * (k + 0.5);
      hr[k] = std::cos(th);
      hr[nk + k] = std::sin(th);
    }
    double *wj = hr.data() + 2 * nk;
    for (int j = 0; j < s->g.N[1]; j++) wj[j] = d->geom[GK_G_DXK][1][j];
    for (int p = 0; p < 2; p++) {
      if (!s->ring_pole[p]) continue;
      for (int r = 0; r < s->ring_rows; r++) {
        const int j = p == 0 ? s->g.s[1] + r : s->g.e[1] - r;
        wj[j] = (nk / s->ring_nchunk[r]) * d->geom[GK_G_DXK][1][j];
      }
    }
    if ((rc = dev_alloc(s, &s->ring_tab, hr.size())) != GK_OK) return fail(rc);
    if (upload(s, s->ring_tab, hr.data(), hr.size() * sizeof(double)) != hipSuccess)
      return fail(GK_ERR_DEVICE);
    s->dxkj = s->ring_tab + 2 * nk;
  }
  std::vector<float> hf;  // kept until the final stream synchronisation
  if (s->precision == 1) {  // FP32 copies of the reconstruction tables
    hf.resize(h.size());
    for (size_t q = 0; q < h.size(); q++) hf[q] = (float)h[q];
    if ((rc = dev_alloc_bytes(s, (void **)&s->tables_f, hf.size() * sizeof(float))) != GK_OK)
      return fail(rc);
    if (upload(s, s->tables_f, hf.data(), hf.size() * sizeof(float)) != hipSuccess)
      return fail(GK_ERR_DEVICE);
    for (int a = 0; a < 3; a++) s->geo.recwf[a] = s->tables_f + off_w[a];
  }

  const size_t fs = (size_t)s->fs;
  const size_t d8 = sizeof(double), f4 = sizeof(float);
  for (int b = 0; b < 2; b++)
    if ((rc = field_alloc(s, (void **)&s->X[b], 6 * fs * d8, d8)) != GK_OK) return fail(rc);
  for (int b = 0; b < 3; b++) {
    if ((rc = field_alloc(s, (void **)&s->B[b], 3 * fs * d8, d8)) != GK_OK) return fail(rc);
    if ((rc = field_alloc(s, (void **)&s->M[b], 3 * fs * d8, d8)) != GK_OK) return fail(rc);
  }
  if (s->precision == 1) {
    for (int b = 0; b < 2; b++)
      if ((rc = field_alloc(s, (void **)&s->X32[b], 6 * fs * f4, f4)) != GK_OK) return fail(rc);
    for (int b = 0; b < 3; b++) {
      if ((rc = field_alloc(s, (void **)&s->B32[b], 3 * fs * f4, f4)) != GK_OK) return fail(rc);
      if ((rc = field_alloc(s, (void **)&s->M32[b], 3 * fs * f4, f4)) != GK_OK) return fail(rc);
    }
  }
  for (int d2 = 0; d2 < 3; d2++) {
    if ((rc = field_alloc(s, &s->F[d2], 9 * fs * s->esz, s->esz)) != GK_OK) return fail(rc);
    if ((rc = field_alloc(s, &s->V[d2], 2 * fs * s->esz, s->esz)) != GK_OK) return fail(rc);
  }
  if ((rc = field_alloc(s, &s->E, 3 * fs * s->esz, s->esz)) != GK_OK) return fail(rc);
  if (d->b0_mode == 2) {  // the background field per point (FP64 in both forms)
    void *b0a = nullptr;
    if ((rc = field_alloc(s, &b0a, (size_t)B0A_COUNT * fs * sizeof(double), sizeof(double))) !=
        GK_OK)
      return fail(rc);
    s->b0a = (double *)b0a;
    for (int q = 0; q < B0A_COUNT; q++)
      if ((rc = copy_in(s, s->b0a + q * fs, d->b0_field[q])) != GK_OK) return fail(rc);
  }
  if (hipMallocAsync((void **)&s->sc, sizeof(Scal), s->stream) != hipSuccess)
    return fail(GK_ERR_NOMEM);
  // pinned: steps taken and non-finite flag of the last run, then room for a
  // copy of the device scalars (gk_get_scalars)
  if (!(s->ctl_host = (long long *)pinned_get(2 * sizeof(long long) + sizeof(Scal))))
    return fail(GK_ERR_NOMEM);
  Scal z{};
  z.budget = -1;
  z.t_stop = INFINITY;
  z.t_event = INFINITY;
  memcpy(s->ctl_host + 2, &z, sizeof z);  // through pinned memory
  if (hipMemcpyAsync(s->sc, s->ctl_host + 2, sizeof(Scal), hipMemcpyHostToDevice, s->stream) !=
      hipSuccess)
    return fail(GK_ERR_DEVICE);
  if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(GK_ERR_DEVICE);
  default_shapes(s);
  s->emf_stage = s->precision == 1 ? 2 : 3;  // measured fastest at 256^3
  s->cur = 0;
  s->b_m1 = 0;
  s->b_0 = 1;
  s->b_n = 2;
  if ((rc = release_bounce(s)) != GK_OK) return fail(rc);
  *out = s;
  return GK_OK;
}

extern "C" int gk_destroy(gk_solver *s) {
  if (!s) return GK_OK;
  if (s->batch) return GK_ERR_STATE;  // remove it from its batch first
  if (s->stream) {
    (void)hipStreamSynchronize(s->stream);  // the solver's own work only
    void *dev[] = {s->tables, s->tables_f, s->ring_tab, s->stage, s->itables, s->sc};
    for (void *p : dev)
      if (p) (void)hipFreeAsync(p, s->stream);
    for (void *p : s->field_mem) (void)hipFreeAsync(p, s->stream);
  }
  pinned_put(s->pin);
  pinned_put(s->ctl_host);
  for (hipEvent_t e : s->evpool) (void)hipEventDestroy(e);
  if (s->own_stream) stream_put(s->device, s->stream);
  delete s;
  return GK_OK;
}

size_t gk_device_bytes(const gk_solver *s) { return s ? s->bytes : 0; }

// Host reference layout (pitch NK) <-> device layout (pitch KP).  One
// contiguous copy through a device staging field plus a re-pitch kernel: a
// pitched copy from pageable memory would move one short row at a time.
static int staging(gk_solver *s, double **out) {
  if (!s->stage) {
    const size_t n = (size_t)s->g.N[0] * s->g.N[1] * s->g.N[2];
    if (hipMallocAsync((void **)&s->stage, n * sizeof(double), s->stream) != hipSuccess)
      return GK_ERR_NOMEM;
    s->bytes += n * sizeof(double);
  }
  *out = s->stage;
  return GK_OK;
}

static void host_copy(double *dst, const double *src, size_t n);

// Host -> device: the caller's array is copied into the pinned bounce buffer
// (a copy from pageable memory can wait behind other streams' work), then to
// the staging field and re-pitched.
static int copy_in(gk_solver *s, double *dst, const double *src) {
  double *st;
  int rc = staging(s, &st);
  if (rc != GK_OK) return rc;
  const long rows = (long)s->g.N[0] * s->g.N[1];
  const int NK = s->g.N[2];
  const size_t n = (size_t)rows * NK;
  if (!s->pin && !(s->pin = (double *)pinned_get(n * sizeof(double)))) return GK_ERR_NOMEM;
  HIPCHK(hipStreamSynchronize(s->stream));  // the bounce buffer's last copy is done
  host_copy(s->pin, src, n);
  HIPCHK(hipMemcpyAsync(st, s->pin, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
  k_pitch<<<2048, 256, 0, s->stream>>>(st, dst, rows, NK, NK, (int)s->g.st[1]);
  HIPCHK(hipGetLastError());
  return GK_OK;
}

// Hand the bounce buffer back to the pool once the transfers that use it are
// done (a solver keeps none between calls).
static int release_bounce(gk_solver *s) {
  if (!s->pin) return GK_OK;
  HIPCHK(hipStreamSynchronize(s->stream));
  pinned_put(s->pin);
  s->pin = nullptr;
  return GK_OK;
}

// Host copy of n doubles, split over a few threads for large fields (a
// single thread copies pinned -> pageable memory at well under the PCIe rate).
static void host_copy(double *dst, const double *src, size_t n) {
  const unsigned hc = std::thread::hardware_concurrency();
  const size_t nt = n < ((size_t)1 << 22) ? 1 : std::min<size_t>(8, hc ? hc : 1);
  if (nt == 1) {
    memcpy(dst, src, n * sizeof(double));
    return;
  }
  std::vector<std::thread> th;
  const size_t per = (n + nt - 1) / nt;
  for (size_t t = 0; t < nt; t++) {
    const size_t a = t * per, b = std::min(n, a + per);
    if (a < b) th.emplace_back([=] { memcpy(dst + a, src + a, (b - a) * sizeof(double)); });
  }
  for (std::thread &x : th) x.join();
}

// Device -> host: re-pitch into the staging field, copy it to a pinned
// bounce buffer (full PCIe rate), then into the caller's array.
static int copy_out(gk_solver *s, double *dst, const double *src) {
  double *st;
  int rc = staging(s, &st);
  if (rc != GK_OK) return rc;
  const long rows = (long)s->g.N[0] * s->g.N[1];
  const int NK = s->g.N[2];
  const size_t n = (size_t)rows * NK;
  if (!s->pin && !(s->pin = (double *)pinned_get(n * sizeof(double)))) return GK_ERR_NOMEM;
  k_pitch<<<2048, 256, 0, s->stream>>>(src, st, rows, NK, (int)s->g.st[1], NK);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(s->pin, st, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  // the staging field and the bounce buffer are reused by the next copy
  HIPCHK(hipStreamSynchronize(s->stream));
  host_copy(dst, s->pin, n);
  return GK_OK;
}

static dim3 box_grid(int nk, int nj, int ni) {
  return dim3((nk + 63) / 64, (nj + 3) / 4, ni);
}

// blocks of 256 points over the (row, k) plane of a launch
static int row_blocks(int klo, int khi, int rlo, int rhi) {
  return ((khi - klo + 1) * (rhi - rlo + 1) + 255) / 256;
}

// ghost fill: 256 points per block over the points outside the active cells
static dim3 ghost_grid(const Grid &g) {
  long n = 1, t = 1;
  for (int d = 0; d < 3; d++) {
    n *= g.e[d] - g.s[d] + 1;
    t *= g.N[d];
  }
  return dim3((unsigned)((t - n + 255) / 256), 1, 1);
}

// blocks a launch decoded by xcd_patch needs (whole 8 x 8 patches)
static int patch_blocks(const dim3 &g) {
  return (int)((g.x + PATCH - 1) / PATCH * PATCH * ((g.z + PATCH - 1) / PATCH) * PATCH);
}

// cell-centred B of the points outside the active cells, in the box minus
// the upper face slot
static dim3 mags_shell_grid(const Grid &g) {
  long n = 1, t = 1;
  for (int d = 0; d < 3; d++) {
    n *= g.e[d] - g.s[d] + 1;
    t *= g.N[d] - 1;
  }
  return dim3((unsigned)((t - n + 255) / 256), 1, 1);
}

// cell-centred B over the box minus the upper face slot
static dim3 mags_grid(const Grid &g) {
  return dim3(row_blocks(0, g.N[2] - 2, 0, g.N[1] - 2), 1, g.N[0] - 1);
}

static void launch_ghost(gk_solver *s, double *X, double *B) {
  GhostArgs a;
  a.g = s->g;
  a.X = X;
  a.B = B;
  a.X32 = nullptr;
  a.B32 = nullptr;
  memcpy(a.bc, s->bc, sizeof(a.bc));
  a.sc = nullptr;
  const dim3 gr = ghost_grid(s->g);
  const int tot = gr.x * gr.y * gr.z;
  k_ghost<<<dim3((tot + NXCD - 1) / NXCD * NXCD), dim3(64, 4), 0, s->stream>>>(a, gr.x, gr.y, gr.z);
}

static void launch_mags(gk_solver *s, const double *B, double *M) {
  MagsArgs a;
  a.g = s->g;
  a.geo = s->geo;
  a.B = B;
  a.M = M;
  a.M32 = nullptr;
  a.sc = nullptr;
  const dim3 gr = mags_grid(s->g);
  const int tot = patch_blocks(gr);
  k_mags<<<dim3((tot + NXCD - 1) / NXCD * NXCD), dim3(64, 4), 0, s->stream>>>(a, gr.x, gr.y, gr.z);
}

// FP32-bulk form: the steps leave the ghosts of the FP64 primitives unfilled
// (only the FP32 copies' ghosts are read); fill both time levels before the
// FP64 state is read back or copied to FP32
static void fill_x64_ghosts(gk_solver *s) {
  if (s->precision != 1) return;
  for (int b = 0; b < 2; b++) launch_ghost(s, s->X[b], nullptr);
}

// FP32 copies of both time levels (FP32-bulk form), after the state is set
static void refresh_copies32(gk_solver *s) {
  if (s->precision != 1) return;
  fill_x64_ghosts(s);
  // FP32-bulk steps keep only the FP32 copy of the cell-centred B: rebuild
  // the FP64 one from the face field (cell B is a function of face B alone)
  for (int b = 0; b < 3; b++) launch_mags(s, s->B[b], s->M[b]);
  const size_t fs = s->fs;
  const dim3 grid(2048), blk(256);
  for (int b = 0; b < 2; b++) hipLaunchKernelGGL(k_to32, grid, blk, 0, s->stream, s->X[b], s->X32[b], 6 * fs);
  for (int b = 0; b < 3; b++) {
    hipLaunchKernelGGL(k_to32, grid, blk, 0, s->stream, s->B[b], s->B32[b], 3 * fs);
    hipLaunchKernelGGL(k_to32, grid, blk, 0, s->stream, s->M[b], s->M32[b], 3 * fs);
  }
}

extern "C" int gk_set_state(gk_solver *s, const double *const prim[6], const double *const bface[3],
                            const double *const prim_prev[6], const double *const bface_prev[3]) {
  if (!s || !prim || !bface) return GK_ERR_ARG;
  const size_t fs = s->fs;
  int rc;
  for (int m = 0; m < 6; m++) {
    if (!prim[m]) return GK_ERR_ARG;
    if ((rc = copy_in(s, s->X[s->cur] + m * fs, prim[m])) != GK_OK) return rc;
  }
  for (int m = 0; m < 3; m++) {
    if (!bface[m]) return GK_ERR_ARG;
    if ((rc = copy_in(s, s->B[s->b_0] + m * fs, bface[m])) != GK_OK) return rc;
  }
  const bool hist = prim_prev && bface_prev;
  if (hist) {
    for (int m = 0; m < 6; m++) {
      if (!prim_prev[m]) return GK_ERR_ARG;
      if ((rc = copy_in(s, s->X[1 - s->cur] + m * fs, prim_prev[m])) != GK_OK) return rc;
    }
    for (int m = 0; m < 3; m++) {
      if (!bface_prev[m]) return GK_ERR_ARG;
      if ((rc = copy_in(s, s->B[s->b_m1] + m * fs, bface_prev[m])) != GK_OK) return rc;
    }
  }
  s->ready = hist ? 2 : 1;
  HIPCHK(hipStreamSynchronize(s->stream));
  return release_bounce(s);
}

extern "C" int gk_prologue(gk_solver *s, double dt0_restart) {
  if (!s || !s->ready) return GK_ERR_STATE;
  const size_t fs = s->fs;
  launch_ghost(s, s->X[s->cur], s->B[s->b_0]);
  launch_mags(s, s->B[s->b_0], s->M[s->b_0]);
  if (s->ready == 1) {  // fresh start: X_p := X (src/solver_solve.c:55-92)
    HIPCHK(hipMemcpyAsync(s->X[1 - s->cur], s->X[s->cur], 6 * fs * sizeof(double),
                          hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->B[s->b_m1], s->B[s->b_0], 3 * fs * sizeof(double),
                          hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->M[s->b_m1], s->M[s->b_0], 3 * fs * sizeof(double),
                          hipMemcpyDeviceToDevice, s->stream));
  } else {
    launch_ghost(s, s->X[1 - s->cur], s->B[s->b_m1]);
    launch_mags(s, s->B[s->b_m1], s->M[s->b_m1]);
  }
  // Ghost faces that no step writes (a side without a boundary condition)
  // keep their initial values at every time level, as in the reference's
  // single in-place array; the X buffers already hold them, and the third
  // face-B buffer starts as a copy of the current one.
  HIPCHK(hipMemcpyAsync(s->B[s->b_n], s->B[s->b_0], 3 * fs * sizeof(double),
                        hipMemcpyDeviceToDevice, s->stream));
  refresh_copies32(s);
  // scalars through the solver's pinned buffer (pageable copies can wait
  // behind other streams' work)
  Scal *pin = (Scal *)(s->ctl_host + 2);
  HIPCHK(hipMemcpyAsync(pin, s->sc, sizeof(Scal), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  Scal cur;
  memcpy(&cur, pin, sizeof cur);
  Scal z{};
  for (int q = 0; q < NSLOT; q++) {
    z.smin[q] = DBL_INF_BITS;
    z.smax[q] = 0ULL;
  }
  z.time = cur.time;  // keep the simulated time across prologues
  z.budget = -1;
  z.t_stop = INFINITY;
  z.t_event = INFINITY;
  memcpy(pin, &z, sizeof z);
  HIPCHK(hipMemcpyAsync(s->sc, pin, sizeof(Scal), hipMemcpyHostToDevice, s->stream));
  DtArgs a;
  a.g = s->g;
  a.geo = s->geo;
  a.b0 = s->b0;
  a.b0a = s->b0a;
  a.X = s->X[s->cur];
  a.M = s->M[s->b_0];
  a.B = s->B[s->b_0];
  a.dxkj = s->dxkj;
  a.j0 = s->g.s[1];
  a.j1 = s->g.e[1];
  a.sc = s->sc;
  if (s->b0a)
    hipLaunchKernelGGL(k_dt<true>, box_grid(s->nk, s->nj, s->ni), dim3(64, 4), 0, s->stream, a);
  else
    hipLaunchKernelGGL(k_dt<false>, box_grid(s->nk, s->nj, s->ni), dim3(64, 4), 0, s->stream, a);
  if (s->ready == 1) {
    hipLaunchKernelGGL(k_set_dt0, dim3(1), dim3(NSLOT), 0, s->stream, s->sc);
  } else {
    HIPCHK(hipStreamSynchronize(s->stream));  // the pinned scalars above have gone out
    memcpy(pin, &dt0_restart, sizeof(double));
    HIPCHK(hipMemcpyAsync(&s->sc->dt0, pin, sizeof(double), hipMemcpyHostToDevice, s->stream));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  s->ready = 3;
  return GK_OK;
}

namespace {
// Per-kernel timing with events on the solver's stream (only when enabled).
// Events come from a pool and are resolved after the stream synchronises at
// the end of gk_step, so timing does not serialise the host with the GPU.
struct Timer {
  gk_solver *s;
  int slot;
  Timer(gk_solver *s_, const char *name) : s(s_), slot(-1) {
    if (!s->timing) return;
    if (!s->time_only.empty() && s->time_only != name) return;
    for (size_t n = 0; n < s->names.size(); n++)
      if (s->names[n] == name) slot = (int)n;
    if (slot < 0) {
      slot = (int)s->names.size();
      s->names.push_back(name);
      s->acc_ms.push_back(0.0);
      s->acc_n.push_back(0);
    }
    const size_t need = 2 * (s->pend.size() + 1);
    while (s->evpool.size() < need) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        slot = -1;
        return;
      }
      s->evpool.push_back(e);
    }
    (void)hipEventRecord(s->evpool[2 * s->pend.size()], s->stream);
  }
  ~Timer() {
    if (slot < 0) return;
    (void)hipEventRecord(s->evpool[2 * s->pend.size() + 1], s->stream);
    s->pend.push_back(slot);
  }
};

// accumulate the recorded kernel times (stream already synchronised)
void resolve_timers(gk_solver *s) {
  for (size_t n = 0; n < s->pend.size(); n++) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, s->evpool[2 * n], s->evpool[2 * n + 1]) == hipSuccess) {
      s->acc_ms[s->pend[n]] += ms;
      s->acc_n[s->pend[n]] += 1;
    }
  }
  s->pend.clear();
}
}  // namespace

// ---------------------------------------------------------------------------
// One AB2 step as a list of launches.  The same plan drives the single-grid
// path (launched one by one) and the batched path (launches of many grids
// merged per stage into one work-list launch).
// Ids follow the order a step runs them in: a batch launches its stages in
// id order (the k_pole_* and k_ring_* stages exist on pole ring-averaged grids)
enum StageKernel {
  SK_BEGIN = 0, SK_SWEEP_I, SK_VSTRIP_I, SK_SWEEP_J, SK_VSTRIP_J, SK_SWEEP_K, SK_KREM, SK_EMF_I, SK_EMF_J,
  SK_EMF_K, SK_EMF_IS, SK_EMF_JS, SK_POLE_E, SK_CT, SK_POLE_B, SK_CELL, SK_RING_ROWS,
  SK_RING_CURL, SK_RING_FIX, SK_RING_MAGS, SK_GHOST, SK_MAGS, SK_RING_DT, SK_COUNT
};
static const char *const sk_name[SK_COUNT] = {
    "k_begin", "k_sweep_i", "k_vstrip_i", "k_sweep_j", "k_vstrip_j", "k_sweep_k", "k_krem_k", "k_emf_i",
    "k_emf_j", "k_emf_k", "k_emf_i_tiles", "k_emf_j_tiles", "k_pole_e", "k_ct", "k_pole_b",
    "k_cell", "k_ring_rows", "k_ring_curl", "k_ring_fix", "k_ring_mags", "k_ghost", "k_mags",
    "k_ring_dt"};

union StageArgs {
  SweepArgs sw;
  EmfArgs em;
  CtArgs ct;
  CellArgs ce;
  GhostArgs gh;
  MagsArgs mg;
  RingArgs rg;
  DtArgs dt;
  Scal *sc;
};

static size_t stage_arg_size(int kid) {
  switch (kid) {
    case SK_BEGIN: return sizeof(Scal *);
    case SK_SWEEP_I: case SK_VSTRIP_I: case SK_SWEEP_J: case SK_VSTRIP_J: case SK_SWEEP_K:
    case SK_KREM:
      return sizeof(SweepArgs);
    case SK_EMF_I: case SK_EMF_J: case SK_EMF_K: case SK_EMF_IS: case SK_EMF_JS:
      return sizeof(EmfArgs);
    case SK_CT: return sizeof(CtArgs);
    case SK_CELL: return sizeof(CellArgs);
    case SK_GHOST: return sizeof(GhostArgs);
    case SK_RING_DT: return sizeof(DtArgs);
    case SK_POLE_E: case SK_POLE_B: case SK_RING_ROWS: case SK_RING_CURL: case SK_RING_FIX:
    case SK_RING_MAGS:
      return sizeof(RingArgs);
    default: return sizeof(MagsArgs);
  }
}

static const void *stage_arg_ptr(int kid, const StageArgs &u) {
  switch (kid) {
    case SK_BEGIN: return &u.sc;
    case SK_SWEEP_I: case SK_VSTRIP_I: case SK_SWEEP_J: case SK_VSTRIP_J: case SK_SWEEP_K:
    case SK_KREM:
      return &u.sw;
    case SK_EMF_I: case SK_EMF_J: case SK_EMF_K: case SK_EMF_IS: case SK_EMF_JS:
      return &u.em;
    case SK_CT: return &u.ct;
    case SK_CELL: return &u.ce;
    case SK_GHOST: return &u.gh;
    case SK_RING_DT: return &u.dt;
    case SK_POLE_E: case SK_POLE_B: case SK_RING_ROWS: case SK_RING_CURL: case SK_RING_FIX:
    case SK_RING_MAGS:
      return &u.rg;
    default: return &u.mg;
  }
}

struct Launch {
  int kid;
  dim3 grid;
  StageArgs u;
  bool rest;  // leftover edge column of E_i or E_j (after whole 64-edge tiles)
};

// history rotation: X[cur] = X^n, X[1-cur] = X^{n-1}; B/M[b_m1], [b_0], [b_n]
struct Rot {
  int cur, b_m1, b_0, b_n;
};
static Rot rot_of(const gk_solver *s) { return Rot{s->cur, s->b_m1, s->b_0, s->b_n}; }
static Rot rot_next(Rot r) { return Rot{1 - r.cur, r.b_0, r.b_n, r.b_m1}; }

// blocks of 256 points over the launch box of the direct sweep
static int point_blocks(const SweepArgs &v) {
  const long n = (long)(v.hi[2] - v.lo[2] + 1) * (v.hi[1] - v.lo[1] + 1) * (v.hi[0] - v.lo[0] + 1);
  return (int)((n + 255) / 256);
}

static void plan_step(const gk_solver *s, const Rot &r, std::vector<Launch> &out) {
  out.clear();
  const Grid &g = s->g;
  const int *S = g.s, *Ee = g.e;
  const double *X = s->X[r.cur], *Xp = s->X[1 - r.cur];
  const double *Bm1 = s->B[r.b_m1], *B = s->B[r.b_0];
  const double *Mm1 = s->M[r.b_m1], *M = s->M[r.b_0];
  double *Bn = s->B[r.b_n], *Mn = s->M[r.b_n];
  double *Xn = s->X[1 - r.cur];
  const int ch = s->chunk;
  // the stencil kernels read the state in their compute type
  const bool f32 = s->precision == 1;
  const void *Xs = f32 ? (const void *)s->X32[r.cur] : (const void *)X;
  const void *Xps = f32 ? (const void *)s->X32[1 - r.cur] : (const void *)Xp;
  const void *Ms = f32 ? (const void *)s->M32[r.b_0] : (const void *)M;
  const void *Mps = f32 ? (const void *)s->M32[r.b_m1] : (const void *)Mm1;
  const void *Bs = f32 ? (const void *)s->B32[r.b_0] : (const void *)B;
  const void *Bps = f32 ? (const void *)s->B32[r.b_m1] : (const void *)Bm1;
  Launch L;
  memset(&L, 0, sizeof L);
  L.kid = SK_BEGIN;
  L.grid = dim3(1, 1, 1);
  L.u.sc = s->sc;
  out.push_back(L);

  // directional sweeps: face fluxes + face velocities
  for (int D = 0; D < 3; D++) {
    SweepArgs a;
    memset(&a, 0, sizeof a);
    a.g = g;
    a.geo = s->geo;
    a.b0 = s->b0;
    a.b0a = s->b0a;
  a.b0a = s->b0a;
    a.X = Xs;
    a.Xp = Xps;
    a.M = Ms;
    a.Mp = Mps;
    a.B = Bs;
    a.F = s->F[D];
    a.V = s->V[D];
    a.sc = s->sc;
    a.nch = march_chunks(Ee[D] + 1 - S[D] + 1, ch);
    for (int d = 0; d < 3; d++) {
      a.lo[d] = S[d];
      a.hi[d] = Ee[d];
      a.flo[d] = S[d];
      a.fhi[d] = Ee[d];
    }
    a.hi[D] = Ee[D] + 1;
    a.fhi[D] = Ee[D] + 1;
    // transverse extension for the edge-velocity chain:
    // i-sweep over all j, j-sweep over all k, k-sweep over all i
    const int X2 = D == 0 ? 1 : (D == 1 ? 2 : 0);
    a.lo[X2] = 0;
    a.hi[X2] = g.N[X2] - 2;
    // on grids whose i and j marches still fill every wave slot (1024 SIMDs
    // x 64 lanes x the march's waves per SIMD: one in FP64, two in FP32)
    // with twice the march length, they walk twice the length: there the
    // window prologue costs more than the longer march loses in parallelism
    // (256^3 at the default of 64: 128-129 faces per thread, 2 chunks of
    // 1024 waves each); smaller grids stepped alone keep the parallelism
    auto long_march = [&](SweepArgs &m, int cols) {
      const int n2 = march_chunks(m.hi[D] - m.lo[D] + 1, std::min(2 * ch, 4096));
      if ((long)cols * n2 < 1024L * 64 * (f32 ? 2 : 1)) return false;
      m.nch = n2;
      return true;
    };
    if (D == 0) {
      // on such grids the i march covers the active rows only (j inside the
      // box), so that the launch is a whole number of waves per chunk
      // (256^3: 1024 waves of 64 k points), and the j ghost rows, which only
      // need face velocities, go to velocity-only strips, as the k ghost
      // columns of the j sweep do; on smaller grids (and batch members) the
      // two extra launches would cost more than they save
      SweepArgs m = a;
      m.lo[1] = S[1];
      m.hi[1] = Ee[1];
      if (!long_march(m, (m.hi[2] - m.lo[2] + 1) * (m.hi[1] - m.lo[1] + 1))) {
        L.kid = SK_SWEEP_I;
        L.grid = dim3(row_blocks(a.lo[2], a.hi[2], a.lo[1], a.hi[1]), 1, a.nch);
        L.u.sw = a;
        out.push_back(L);
        continue;
      }
      L.kid = SK_SWEEP_I;
      L.grid = dim3(row_blocks(m.lo[2], m.hi[2], m.lo[1], m.hi[1]), 1, m.nch);
      L.u.sw = m;
      out.push_back(L);
      {
        // both strips (NGH rows below the box and NGH above) in one box of
        // 2 NGH virtual rows, the upper ones shifted past the box
        SweepArgs v = a;
        v.lo[1] = 0;
        v.hi[1] = 2 * NGH - 1;
        v.gap = Ee[1] + 1 - NGH;
        L.kid = SK_VSTRIP_I;
        L.grid = dim3(point_blocks(v), 1, 1);
        L.u.sw = v;
        out.push_back(L);
      }
    } else if (D == 1) {
      SweepArgs m = a;  // full faces: k over the active range
      m.lo[2] = S[2];
      m.hi[2] = Ee[2];
      long_march(m, (m.hi[2] - m.lo[2] + 1) * (m.hi[0] - m.lo[0] + 1));
      L.kid = SK_SWEEP_J;
      L.grid = dim3(row_blocks(m.lo[2], m.hi[2], m.lo[0], m.hi[0]), 1, m.nch);
      L.u.sw = m;
      out.push_back(L);
      // velocity-only strips k in [0, ks-1] and [ke+1, NK-2], as one box
      // of 2 NGH virtual columns, the upper ones shifted past the box
      {
        SweepArgs v = a;
        v.lo[2] = 0;
        v.hi[2] = 2 * NGH - 1;
        v.gap = Ee[2] + 1 - NGH;
        L.kid = SK_VSTRIP_J;
        L.grid = dim3(point_blocks(v), 1, 1);
        L.u.sw = v;
        out.push_back(L);
      }
    } else {
      // whole 64-face tiles in k_sweepk; a remainder of fewer than 32 faces
      // per row goes to the direct sweep instead of a mostly empty tile
      const int wk = a.hi[2] - a.lo[2] + 1;
      const int rem = wk % 64;
      SweepArgs m = a;
      const bool split = wk >= 64 && rem > 0 && rem < 32;
      if (split) m.hi[2] = a.hi[2] - rem;
      L.kid = SK_SWEEP_K;
      L.grid = dim3((m.hi[2] - m.lo[2] + 64) / 64, m.hi[1] - m.lo[1] + 1, m.hi[0] - m.lo[0] + 1);
      L.u.sw = m;
      out.push_back(L);
      if (split) {
        SweepArgs v = a;
        v.lo[2] = m.hi[2] + 1;
        L.kid = SK_KREM;  // k_kcol: one block per column and 64 rows
        const long nrow = (long)(v.hi[1] - v.lo[1] + 1) * (v.hi[0] - v.lo[0] + 1);
        L.grid = dim3((unsigned)((nrow + 63) / 64 * rem), 1, 1);
        L.u.sw = v;
        out.push_back(L);
      }
    }
  }
  // edge electric fields
  for (int D = 0; D < 3; D++) {
    EmfArgs a;
    memset(&a, 0, sizeof a);
    a.g = g;
    a.geo = s->geo;
    a.b0 = s->b0;
    a.b0a = s->b0a;
  a.b0a = s->b0a;
    a.X = Xs;
    a.Xp = Xps;
    a.B = Bs;
    a.Bp = Bps;
    // E_i uses the j-sweep velocities, E_j the k-sweep, E_k the i-sweep
    a.V = s->V[D == 0 ? 1 : (D == 1 ? 2 : 0)];
    a.E = (char *)s->E + (size_t)D * s->fs * s->esz;
    a.sc = s->sc;
    for (int d = 0; d < 3; d++) {
      a.lo[d] = S[d];
      a.hi[d] = Ee[d] + (d == D ? 0 : 1);
    }
    const int MA = D == 1 ? 0 : 1, RA = D == 1 ? 1 : 0;
    a.nch = march_chunks(a.hi[MA] - a.lo[MA] + 1, s->emf_chunk);
    const int nf = a.nch;
    // staged E_i / E_j (bit D of emf_stage; by default both in FP64, E_j in
    // FP32-bulk): whole 64-edge tiles along k in the staged form, the rest
    // (all of it when there is no whole tile) in the marching form
    const int ntile = (D < 2 && ((s->emf_stage >> D) & 1)) ? (a.hi[2] - a.lo[2] + 1) / 64 : 0;
    if (ntile > 0) {
      EmfArgs t = a;
      t.hi[2] = a.lo[2] + 64 * ntile - 1;
      L.kid = D == 0 ? SK_EMF_IS : SK_EMF_JS;
      L.grid = dim3(ntile, a.hi[RA] - a.lo[RA] + 1, nf);
      L.u.em = t;
      out.push_back(L);
      a.lo[2] = t.hi[2] + 1;
      // the leftover columns are few points: march one edge per thread so
      // that they spread over many waves
      a.nch = a.hi[MA] - a.lo[MA] + 1;
    }
    if (a.lo[2] <= a.hi[2]) {
      L.kid = SK_EMF_I + D;
      L.grid = dim3(row_blocks(a.lo[2], a.hi[2], a.lo[RA], a.hi[RA]), 1, a.nch);
      L.u.em = a;
      L.rest = ntile > 0;
      out.push_back(L);
      L.rest = false;
    }
  }
  // pole ring averaging: arguments shared by its kernels
  RingArgs ra;
  memset(&ra, 0, sizeof ra);
  if (s->ring) {
    ra.g = g;
    ra.geo = s->geo;
    ra.X = Xn;
    ra.B = Bn;
    ra.M = Mn;
    ra.E = s->E;
    ra.X32 = f32 ? s->X32[1 - r.cur] : nullptr;
    ra.B32 = f32 ? s->B32[r.b_n] : nullptr;
    ra.M32 = f32 ? s->M32[r.b_n] : nullptr;
    ra.cs = s->ring_tab;
    ra.nk = s->nk;
    ra.rows = s->ring_rows;
    ra.pole[0] = s->ring_pole[0];
    ra.pole[1] = s->ring_pole[1];
    memcpy(ra.nchunk, s->ring_nchunk, sizeof(ra.nchunk));
    ra.sc = s->sc;
    L.kid = SK_POLE_E;  // E_i ring mean and E_k := 0 on the pole edges
    L.grid = dim3(s->ni + 1, 2, 1);
    L.u.rg = ra;
    out.push_back(L);
  }
  {
    CtArgs a;
    memset(&a, 0, sizeof a);
    a.g = g;
    a.geo = s->geo;
    a.B = B;
    a.Bn = Bn;
    a.Bn32 = f32 ? s->B32[r.b_n] : nullptr;
    a.E = s->E;
    a.sc = s->sc;
    L.kid = SK_CT;
    L.grid = dim3(row_blocks(S[2], Ee[2] + 1, S[1], Ee[1] + 1), 1, s->ni + 1);
    L.u.ct = a;
    out.push_back(L);
  }
  if (s->ring) {
    L.kid = SK_POLE_B;  // B_j on the pole faces (their area is zero)
    L.grid = dim3((unsigned)(((long)s->ni * s->nk + 255) / 256), 2, 1);
    L.u.rg = ra;
    out.push_back(L);
  }
  {
    CellArgs a;
    memset(&a, 0, sizeof a);
    a.g = g;
    a.geo = s->geo;
    a.b0 = s->b0;
    a.b0a = s->b0a;
  a.b0a = s->b0a;
    a.spherical = s->spherical;
    a.X = X;
    a.Xp = Xp;
    a.Xn = Xn;
    a.Xn32 = f32 ? s->X32[1 - r.cur] : nullptr;
    // the cell-centred B of the Boris correction and the spherical sources:
    // the FP32-bulk form reads the FP32 copies and keeps no FP64 cell B
    a.M = Ms;
    a.Mp = Mps;
    a.Mn = f32 ? nullptr : Mn;
    a.Mn32 = f32 ? s->M32[r.b_n] : nullptr;
    a.Bn = Bn;
    for (int d = 0; d < 3; d++) a.F[d] = s->F[d];
    a.rho_floor = s->rho_floor;
    a.p_floor = s->p_floor;
    for (int m = 0; m < 3; m++) a.grav[m] = s->grav[m];
    a.sc = s->sc;
    a.skip[0] = a.skip[2] = 1;  // empty ranges
    a.skip[1] = a.skip[3] = 0;
    if (s->ring) {
      if (s->ring_pole[0]) {
        a.skip[0] = S[1];
        a.skip[1] = S[1] + s->ring_rows;
      }
      if (s->ring_pole[1]) {
        a.skip[2] = Ee[1] - s->ring_rows;
        a.skip[3] = Ee[1];
      }
    }
    L.kid = SK_CELL;
    L.grid = dim3(row_blocks(S[2], Ee[2], S[1], Ee[1]), 1, (s->ni + CELL_CHUNK - 1) / CELL_CHUNK);
    L.u.ce = a;
    out.push_back(L);
  }
  if (s->ring) {
    // hydro rows and ring edge fields, curl update, pole-row fixes, then the
    // cell-centred B of the rows whose faces changed
    const long band = (long)(s->ring_rows + 1);
    L.kid = SK_RING_ROWS;
    L.grid = dim3(s->ni + 1, s->ring_rows, 14);
    L.u.rg = ra;
    out.push_back(L);
    L.kid = SK_RING_CURL;
    L.grid = dim3((unsigned)(((long)(s->ni + 1) * band * (s->nk + 1) + 255) / 256), 2, 1);
    out.push_back(L);
    L.kid = SK_RING_FIX;
    L.grid = dim3(s->ni, 2, 1);
    out.push_back(L);
    L.kid = SK_RING_MAGS;
    L.grid = dim3((unsigned)(((long)s->ni * band * s->nk + 255) / 256), 2, 1);
    out.push_back(L);
  }
  {
    GhostArgs a;
    memset(&a, 0, sizeof a);
    a.g = g;
    // the FP32-bulk stencils read the ghosts of X32; the FP64 state is only
    // read at active cells during a step, so its ghosts are filled when the
    // state is read back (fill_x64_ghosts)
    a.X = f32 ? nullptr : Xn;
    a.B = Bn;
    a.X32 = f32 ? s->X32[1 - r.cur] : nullptr;
    a.B32 = f32 ? s->B32[r.b_n] : nullptr;
    memcpy(a.bc, s->bc, sizeof(a.bc));
    a.sc = s->sc;
    L.kid = SK_GHOST;
    L.grid = ghost_grid(g);
    L.u.gh = a;
    out.push_back(L);
  }
  {
    MagsArgs a;
    memset(&a, 0, sizeof a);
    a.g = g;
    a.geo = s->geo;
    a.B = Bn;
    a.M = f32 ? nullptr : Mn;  // FP32-bulk: only the FP32 copy is read
    a.M32 = f32 ? s->M32[r.b_n] : nullptr;
    a.sc = s->sc;
    L.kid = SK_MAGS;  // ghost cells only (k_mags_shell)
    L.grid = mags_shell_grid(g);
    L.u.mg = a;
    out.push_back(L);
  }
  if (s->ring) {
    // the next CFL dt and |div B| of the rows the ring average changed (the
    // cell update's fused reduction left them out), with the ring-widened phi
    // width (get_dt_ring); launched once per pole band
    DtArgs a;
    memset(&a, 0, sizeof a);
    a.g = g;
    a.geo = s->geo;
    a.b0 = s->b0;
    a.b0a = s->b0a;
  a.b0a = s->b0a;
    a.X = Xn;
    a.M = Mn;
    a.B = Bn;
    a.dxkj = s->dxkj;
    a.sc = s->sc;
    L.kid = SK_RING_DT;
    L.grid = box_grid(s->nk, s->ring_rows + 1, s->ni);
    for (int p = 0; p < 2; p++) {
      if (!s->ring_pole[p]) continue;
      a.j0 = p == 0 ? S[1] : Ee[1] - s->ring_rows;
      a.j1 = a.j0 + s->ring_rows;
      L.u.dt = a;
      out.push_back(L);
    }
  }
}

template <typename R, bool B0V, bool UT>
static void launch_one(const Launch &L, hipStream_t st) {
  const dim3 blk(64, 4);
  const int gx = L.grid.x, gy = L.grid.y, gz = L.grid.z;
  // decoded by xcd_patch (cell update, CT, cell B) or xcd_block (the others)
  const bool patch = L.kid == SK_CT || L.kid == SK_CELL;
  const int total = patch ? patch_blocks(L.grid) : gx * gy * gz;
  const dim3 g1((total + NXCD - 1) / NXCD * NXCD);
  switch (L.kid) {
    case SK_BEGIN: k_begin<<<1, NSLOT, 0, st>>>(L.u.sc); break;
    case SK_SWEEP_I: k_march<0, R, B0V, UT><<<g1, blk, 0, st>>>(L.u.sw, gx, gy, gz); break;
    case SK_VSTRIP_I: k_sweep<0, R, B0V><<<g1, blk, 0, st>>>(L.u.sw, gx, gy, gz); break;
    case SK_SWEEP_J: k_march<1, R, B0V, UT><<<g1, blk, 0, st>>>(L.u.sw, gx, gy, gz); break;
    case SK_VSTRIP_J: k_sweep<1, R, B0V><<<g1, blk, 0, st>>>(L.u.sw, gx, gy, gz); break;
    case SK_SWEEP_K: k_sweepk<R, B0V><<<g1, dim3(64), 0, st>>>(L.u.sw, gx, gy, gz); break;
    case SK_KREM: k_kcol<R, B0V><<<g1, dim3(64), 0, st>>>(L.u.sw, gx, gy, gz); break;
    case SK_EMF_I: k_emf_march<0, R, B0V, UT><<<g1, blk, 0, st>>>(L.u.em, gx, gy, gz); break;
    case SK_EMF_J: k_emf_march<1, R, B0V, UT><<<g1, blk, 0, st>>>(L.u.em, gx, gy, gz); break;
    case SK_EMF_K: k_emf_march<2, R, B0V, UT><<<g1, blk, 0, st>>>(L.u.em, gx, gy, gz); break;
    case SK_EMF_IS: k_emf_stage<0, R, B0V, UT><<<g1, dim3(64), 0, st>>>(L.u.em, gx, gy, gz); break;
    case SK_EMF_JS: k_emf_stage<1, R, B0V, UT><<<g1, dim3(64), 0, st>>>(L.u.em, gx, gy, gz); break;
    case SK_CT: k_ct<R><<<g1, blk, 0, st>>>(L.u.ct, gx, gy, gz); break;
    case SK_CELL: k_cell<R, B0V><<<g1, blk, 0, st>>>(L.u.ce, gx, gy, gz); break;
    case SK_GHOST: k_ghost<<<g1, blk, 0, st>>>(L.u.gh, gx, gy, gz); break;
    case SK_MAGS: k_mags_shell<<<g1, blk, 0, st>>>(L.u.mg, gx, gy, gz); break;
    case SK_POLE_E: k_pole_e<R><<<L.grid, dim3(64), 0, st>>>(L.u.rg); break;
    case SK_POLE_B: k_pole_b<<<L.grid, dim3(256), 0, st>>>(L.u.rg); break;
    case SK_RING_ROWS: k_ring_rows<R><<<L.grid, dim3(64), 0, st>>>(L.u.rg); break;
    case SK_RING_CURL: k_ring_curl<R><<<L.grid, dim3(256), 0, st>>>(L.u.rg); break;
    case SK_RING_FIX: k_ring_fix<<<L.grid, dim3(64), 0, st>>>(L.u.rg); break;
    case SK_RING_MAGS: k_ring_mags<<<L.grid, dim3(256), 0, st>>>(L.u.rg); break;
    case SK_RING_DT: k_dt<B0V><<<L.grid, dim3(64, 4), 0, st>>>(L.u.dt); break;
  }
}

template <typename R, bool B0V, bool UT>
static void launch_batch(int kid, const void *args, const WorkItem *work, int nwork, int nblocks,
                         hipStream_t st) {
  const dim3 blk(64, 4);
  switch (kid) {
    case SK_BEGIN: k_begin_batch<<<nwork, NSLOT, 0, st>>>((Scal *const *)args, nwork); break;
    case SK_SWEEP_I: k_march_batch<0, R, B0V, UT><<<nblocks, blk, 0, st>>>((const SweepArgs *)args, work, nwork); break;
    case SK_VSTRIP_I: k_sweep_batch<0, R, B0V><<<nblocks, blk, 0, st>>>((const SweepArgs *)args, work, nwork); break;
    case SK_SWEEP_J: k_march_batch<1, R, B0V, UT><<<nblocks, blk, 0, st>>>((const SweepArgs *)args, work, nwork); break;
    case SK_VSTRIP_J: k_sweep_batch<1, R, B0V><<<nblocks, blk, 0, st>>>((const SweepArgs *)args, work, nwork); break;
    case SK_SWEEP_K: k_sweepk_batch<R, B0V><<<nblocks, dim3(64), 0, st>>>((const SweepArgs *)args, work, nwork); break;
    case SK_KREM: k_kcol_batch<R, B0V><<<nblocks, dim3(64), 0, st>>>((const SweepArgs *)args, work, nwork); break;
    case SK_EMF_I: k_emf_march_batch<0, R, B0V, UT><<<nblocks, blk, 0, st>>>((const EmfArgs *)args, work, nwork); break;
    case SK_EMF_J: k_emf_march_batch<1, R, B0V, UT><<<nblocks, blk, 0, st>>>((const EmfArgs *)args, work, nwork); break;
    case SK_EMF_K: k_emf_march_batch<2, R, B0V, UT><<<nblocks, blk, 0, st>>>((const EmfArgs *)args, work, nwork); break;
    case SK_EMF_IS: k_emf_stage_batch<0, R, B0V, UT><<<nblocks, dim3(64), 0, st>>>((const EmfArgs *)args, work, nwork); break;
    case SK_EMF_JS: k_emf_stage_batch<1, R, B0V, UT><<<nblocks, dim3(64), 0, st>>>((const EmfArgs *)args, work, nwork); break;
    case SK_CT: k_ct_batch<R><<<nblocks, blk, 0, st>>>((const CtArgs *)args, work, nwork); break;
    case SK_CELL: k_cell_batch<R, B0V><<<nblocks, blk, 0, st>>>((const CellArgs *)args, work, nwork); break;
    case SK_GHOST: k_ghost_batch<<<nblocks, blk, 0, st>>>((const GhostArgs *)args, work, nwork); break;
    case SK_MAGS: k_mags_batch<<<nblocks, blk, 0, st>>>((const MagsArgs *)args, work, nwork); break;
    case SK_POLE_E: k_pole_e_batch<R><<<nblocks, dim3(64), 0, st>>>((const RingArgs *)args, work, nwork); break;
    case SK_POLE_B: k_pole_b_batch<<<nblocks, dim3(256), 0, st>>>((const RingArgs *)args, work, nwork); break;
    case SK_RING_ROWS: k_ring_rows_batch<R><<<nblocks, dim3(64), 0, st>>>((const RingArgs *)args, work, nwork); break;
    case SK_RING_CURL: k_ring_curl_batch<R><<<nblocks, dim3(256), 0, st>>>((const RingArgs *)args, work, nwork); break;
    case SK_RING_FIX: k_ring_fix_batch<<<nblocks, dim3(64), 0, st>>>((const RingArgs *)args, work, nwork); break;
    case SK_RING_MAGS: k_ring_mags_batch<<<nblocks, dim3(256), 0, st>>>((const RingArgs *)args, work, nwork); break;
    case SK_RING_DT: k_dt_batch<B0V><<<nblocks, dim3(64, 4), 0, st>>>((const DtArgs *)args, work, nwork); break;
  }
}

// Launch groups: a solver's precision, whether it reads B0 per point, and
// whether its reconstruction tables have one class per transverse axis
// select the kernel instantiations it runs.
constexpr int NGROUP = 8;

static int launch_group(const gk_solver *s) {
  return (s->precision == 1 ? 1 : 0) + (s->b0a ? 2 : 0) + (s->ut ? 4 : 0);
}

static void launch_one_group(int grp, const Launch &L, hipStream_t st) {
  switch (grp) {
    case 0: launch_one<double, false, false>(L, st); break;
    case 1: launch_one<float, false, false>(L, st); break;
    case 2: launch_one<double, true, false>(L, st); break;
    case 3: launch_one<float, true, false>(L, st); break;
    case 4: launch_one<double, false, true>(L, st); break;
    case 5: launch_one<float, false, true>(L, st); break;
    case 6: launch_one<double, true, true>(L, st); break;
    default: launch_one<float, true, true>(L, st); break;
  }
}

// call f(R *, B0V, UT) with the launch group's template arguments as tags
template <class F>
static void with_group(int grp, F &&f) {
  using T = std::true_type;
  using N = std::false_type;
  switch (grp) {
    case 0: f((double *)nullptr, N(), N()); break;
    case 1: f((float *)nullptr, N(), N()); break;
    case 2: f((double *)nullptr, T(), N()); break;
    case 3: f((float *)nullptr, T(), N()); break;
    case 4: f((double *)nullptr, N(), T()); break;
    case 5: f((float *)nullptr, N(), T()); break;
    case 6: f((double *)nullptr, T(), T()); break;
    default: f((float *)nullptr, T(), T()); break;
  }
}

// the k sweep's tiles and its leftover k columns in one launch
static void launch_sweepk_rem(int grp, const Launch &T, const Launch &C, hipStream_t st) {
  const int nrem = (int)C.grid.x, nrem8 = (nrem + NXCD - 1) / NXCD * NXCD;
  const int gx = T.grid.x, gy = T.grid.y, gz = T.grid.z;
  const dim3 g1(nrem8 + (gx * gy * gz + NXCD - 1) / NXCD * NXCD);
  with_group(grp, [&](auto r, auto b0v, auto) {
    using R = std::remove_pointer_t<decltype(r)>;
    k_sweepk_rem<R, decltype(b0v)::value><<<g1, dim3(64), 0, st>>>(T.u.sw, C.u.sw, gx, gy, gz, nrem, nrem8);
  });
}

// E_i's and E_j's leftover edge columns in one launch
static_assert(2 * sizeof(EmfArgs) + 2 * sizeof(dim3) <= 4096, "kernel arguments");
static void launch_emf_cols(int grp, const EmfArgs &a0, const EmfArgs &a1, dim3 g0, dim3 gj,
                            hipStream_t st) {
  const int total = (int)(g0.x * g0.y * g0.z + gj.x * gj.y * gj.z);
  const dim3 g1((total + NXCD - 1) / NXCD * NXCD);
  with_group(grp, [&](auto r, auto b0v, auto ut) {
    using R = std::remove_pointer_t<decltype(r)>;
    k_emf_cols<R, decltype(b0v)::value, decltype(ut)::value><<<g1, dim3(64, 4), 0, st>>>(a0, a1, g0, gj);
  });
}

static void launch_batch_group(int grp, int kid, const void *args, const WorkItem *work, int nwork,
                               int nblocks, hipStream_t st) {
  switch (grp) {
    case 0: launch_batch<double, false, false>(kid, args, work, nwork, nblocks, st); break;
    case 1: launch_batch<float, false, false>(kid, args, work, nwork, nblocks, st); break;
    case 2: launch_batch<double, true, false>(kid, args, work, nwork, nblocks, st); break;
    case 3: launch_batch<float, true, false>(kid, args, work, nwork, nblocks, st); break;
    case 4: launch_batch<double, false, true>(kid, args, work, nwork, nblocks, st); break;
    case 5: launch_batch<float, false, true>(kid, args, work, nwork, nblocks, st); break;
    case 6: launch_batch<double, true, true>(kid, args, work, nwork, nblocks, st); break;
    default: launch_batch<float, true, true>(kid, args, work, nwork, nblocks, st); break;
  }
}

static void one_step(gk_solver *s) {
  std::vector<Launch> &plan = s->plan;
  plan_step(s, rot_of(s), plan);
  const int gp = launch_group(s);
  // E_i's and E_j's leftover edge columns (when both run over tiles) in one
  // launch, at the place of E_j's: the edge-field launches are independent
  // of each other.  256^3 FP64: 0.057 + 0.056 ms apart, 0.086 ms together.
  auto col = [](const Launch &L) { return L.rest && (L.kid == SK_EMF_I || L.kid == SK_EMF_J); };
  int ncol = 0;
  for (const Launch &L : plan) ncol += col(L) ? 1 : 0;
  for (size_t q = 0; q < plan.size(); q++) {
    const Launch &L = plan[q];
    // the k sweep's leftover columns (k_kcol) start first, beside its tiles:
    // 256^3 FP64 1.127 + 0.087 ms apart
    if (L.kid == SK_SWEEP_K && q + 1 < plan.size() && plan[q + 1].kid == SK_KREM) {
      Timer t(s, sk_name[SK_SWEEP_K]);
      launch_sweepk_rem(gp, L, plan[q + 1], s->stream);
      q++;
      continue;
    }
    if (ncol == 2 && col(L)) {
      if (L.kid == SK_EMF_I) continue;  // launched with E_j's column
      const Launch *Li = nullptr;
      for (const Launch &M : plan)
        if (col(M) && M.kid == SK_EMF_I) Li = &M;
      Timer t(s, "k_emf_cols");
      launch_emf_cols(gp, Li->u.em, L.u.em, Li->grid, L.grid, s->stream);
      continue;
    }
    Timer t(s, sk_name[L.kid]);
    launch_one_group(gp, L, s->stream);
  }
  const Rot r = rot_next(rot_of(s));
  s->cur = r.cur;
  s->b_m1 = r.b_m1;
  s->b_0 = r.b_0;
  s->b_n = r.b_n;
}

// Arm the grid, enqueue n steps and wait for them; the device stops the run
// early when the budget or a stop time is reached (later steps of the
// enqueued run then do nothing).  *done receives the steps taken; the host
// history rotation follows them.
static int run_steps(gk_solver *s, int n, long long budget, double t_stop, double t_event,
                     int *done) {
  if (s->timing)
    for (size_t q = 0; q < s->acc_ms.size(); q++) {
      s->acc_ms[q] = 0.0;
      s->acc_n[q] = 0;
    }
  hipLaunchKernelGGL(k_arm, dim3(1), dim3(64), 0, s->stream, s->sc, budget, t_stop, t_event);
  const Rot r0 = rot_of(s);
  for (int t = 0; t < n; t++) {
    one_step(s);
    if (s->timing && s->pend.size() >= 4096) {  // bound the event pool
      HIPCHK(hipStreamSynchronize(s->stream));
      resolve_timers(s);
    }
  }
  HIPCHK(hipGetLastError());
  // steps taken and the non-finite flag, read after the run drains
  HIPCHK(hipMemcpyAsync(s->ctl_host, &s->sc->steps, sizeof(long long), hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipMemcpyAsync(s->ctl_host + 1, &s->sc->nanflag, sizeof(unsigned int),
                        hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (s->timing) resolve_timers(s);
  long long took = 0;
  unsigned int bad = 0;
  memcpy(&took, s->ctl_host, sizeof took);
  memcpy(&bad, s->ctl_host + 1, sizeof bad);
  if (took < n) {  // the run stopped early: the rotation follows the steps taken
    Rot r = r0;
    for (long long t = 0; t < took % 6; t++) r = rot_next(r);
    s->cur = r.cur;
    s->b_m1 = r.b_m1;
    s->b_0 = r.b_0;
    s->b_n = r.b_n;
  }
  if (done) *done = (int)took;
  return bad ? GK_ERR_NAN : GK_OK;
}

extern "C" int gk_step(gk_solver *s, int n) {
  if (!s) return GK_ERR_ARG;
  if (s->ready != 3 || s->batch) return GK_ERR_STATE;
  if (n < 0) return GK_ERR_ARG;
  return run_steps(s, n, -1, INFINITY, INFINITY, nullptr);
}

extern "C" int gk_step_until(gk_solver *s, int n, double t_stop, double t_event, int *done) {
  if (!s) return GK_ERR_ARG;
  if (s->ready != 3 || s->batch) return GK_ERR_STATE;
  if (n < 0 || t_stop != t_stop || t_event != t_event) return GK_ERR_ARG;
  return run_steps(s, n, -1, t_stop, t_event, done);
}

extern "C" int gk_get_state(gk_solver *s, double *prim[6], double *bface[3], double *prim_prev[6],
                            double *bface_prev[3]) {
  if (!s) return GK_ERR_ARG;
  const size_t fs = s->fs;
  int rc;
  if (prim || prim_prev) fill_x64_ghosts(s);
  for (int m = 0; m < 6; m++) {
    if (prim && prim[m] && (rc = copy_out(s, prim[m], s->X[s->cur] + m * fs)) != GK_OK) return rc;
    if (prim_prev && prim_prev[m] &&
        (rc = copy_out(s, prim_prev[m], s->X[1 - s->cur] + m * fs)) != GK_OK)
      return rc;
  }
  for (int m = 0; m < 3; m++) {
    if (bface && bface[m] && (rc = copy_out(s, bface[m], s->B[s->b_0] + m * fs)) != GK_OK) return rc;
    if (bface_prev && bface_prev[m] &&
        (rc = copy_out(s, bface_prev[m], s->B[s->b_m1] + m * fs)) != GK_OK)
      return rc;
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return release_bounce(s);
}

// ---------------------------------------------------------------------------
// Read-back that allocates nothing (gk_get_state_noalloc).  gk_get_state
// stages each field through a device staging field and a pinned host buffer
// that it allocates on first use; this variant stages through `scratch`, a
// file-scope device array of NOALLOC_SCRATCH_BYTES that every GPU the library
// runs on holds for the life of the process, and copies from there straight
// into the caller's arrays.  In host code `scratch` names a host-side
// stand-in, not device memory: its device address on each GPU comes from
// hipGetSymbolAddress, looked up once per device (with that device current)
// and kept in scratch_dev[]; only that address is ever passed to a kernel or
// a copy.  Every call on one GPU shares the array, so calls that overlap on
// one GPU (other threads, other streams) race; each field is written in full
// before it is read, so nothing needs clearing between calls.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;  // one field of up to 4Mi points
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex scratch_mu;
static void *scratch_dev[MAX_DEV];

static int scratch_on(int dev, double **out) {
  if (dev < 0 || dev >= MAX_DEV) return GK_ERR_ARG;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "gamera_op: hipGetSymbolAddress(scratch) failed: %s\n",
              hipGetErrorString(e));
      return GK_ERR_DEVICE;
    }
    scratch_dev[dev] = p;
  }
  *out = (double *)scratch_dev[dev];
  return GK_OK;
}

// one device field -> the caller's host array (reference layout) via scratch
static int copy_out_scratch(gk_solver *s, double *st, double *dst, const double *src) {
  const long rows = (long)s->g.N[0] * s->g.N[1];
  const int NK = s->g.N[2];
  k_pitch<<<2048, 256, 0, s->stream>>>(src, st, rows, NK, (int)s->g.st[1], NK);
  HIPCHK(hipGetLastError());
  // stream order keeps the next field's re-pitch behind this copy
  HIPCHK(hipMemcpyAsync(dst, st, (size_t)rows * NK * sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  return GK_OK;
}

extern "C" int gk_get_state_noalloc(gk_solver *s, double *prim[6], double *bface[3],
                                    double *prim_prev[6], double *bface_prev[3]) {
  if (!s) return GK_ERR_ARG;
  const size_t n = (size_t)s->g.N[0] * s->g.N[1] * s->g.N[2];
  if (n * sizeof(double) > NOALLOC_SCRATCH_BYTES) {
    fprintf(stderr,
            "gamera_op: gk_get_state_noalloc: a field of %zu points (%zu bytes) exceeds the "
            "%zu-byte scratch; use gk_get_state\n",
            n, n * sizeof(double), NOALLOC_SCRATCH_BYTES);
    return GK_ERR_ARG;
  }
  double *st = nullptr;
  int rc = scratch_on(s->device, &st);
  if (rc != GK_OK) return rc;
  const size_t fs = s->fs;
  if (prim || prim_prev) fill_x64_ghosts(s);
  for (int m = 0; m < 6; m++) {
    if (prim && prim[m] && (rc = copy_out_scratch(s, st, prim[m], s->X[s->cur] + m * fs)) != GK_OK)
      return rc;
    if (prim_prev && prim_prev[m] &&
        (rc = copy_out_scratch(s, st, prim_prev[m], s->X[1 - s->cur] + m * fs)) != GK_OK)
      return rc;
  }
  for (int m = 0; m < 3; m++) {
    if (bface && bface[m] && (rc = copy_out_scratch(s, st, bface[m], s->B[s->b_0] + m * fs)) != GK_OK)
      return rc;
    if (bface_prev && bface_prev[m] &&
        (rc = copy_out_scratch(s, st, bface_prev[m], s->B[s->b_m1] + m * fs)) != GK_OK)
      return rc;
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return GK_OK;
}

extern "C" int gk_debug_fetch(gk_solver *s, int which, int field, double *host) {
  if (!s || !host || field < 0) return GK_ERR_ARG;
  const char *src = nullptr;
  size_t esz = s->esz;
  const size_t fb = (size_t)s->fs * s->esz;
  if (which >= 0 && which < 3 && field < 9) src = (const char *)s->F[which] + field * fb;
  if (which >= 3 && which < 6 && field < 2) src = (const char *)s->V[which - 3] + field * fb;
  if (which == 6 && field < 3) src = (const char *)s->E + field * fb;
  if (which == 7 && field < 3) {
    // FP32-bulk steps keep only the FP32 copy: form the FP64 one first
    if (s->precision == 1) launch_mags(s, s->B[s->b_0], s->M[s->b_0]);
    src = (const char *)(s->M[s->b_0] + field * s->fs);
    esz = sizeof(double);
  }
  if (!src) return GK_ERR_ARG;
  if (esz == sizeof(double)) {
    int rc = copy_out(s, host, (const double *)src);
    if (rc != GK_OK) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    return GK_OK;
  }
  // FP32 intermediates: copy the padded block, widen into the host layout
  std::vector<float> tmp((size_t)s->fs);
  HIPCHK(hipMemcpyAsync(tmp.data(), src, (size_t)s->fs * sizeof(float), hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  const size_t NK = s->g.N[2];
  for (int i = 0; i < s->g.N[0]; i++)
    for (int j = 0; j < s->g.N[1]; j++)
      for (size_t k = 0; k < NK; k++)
        host[((size_t)i * s->g.N[1] + j) * NK + k] = tmp[i * s->g.st[0] + j * s->g.st[1] + k];
  return GK_OK;
}

extern "C" int gk_get_scalars(gk_solver *s, double out[6]) {
  if (!s || !out) return GK_ERR_ARG;
  // through pinned memory: a pageable copy can wait behind other streams' work
  Scal *pin = (Scal *)(s->ctl_host + 2);
  HIPCHK(hipMemcpyAsync(pin, s->sc, sizeof(Scal), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  Scal h;
  memcpy(&h, pin, sizeof h);
  out[0] = h.dt;
  out[1] = h.dt0;
  out[2] = h.time;
  unsigned long long mn = DBL_INF_BITS, mx = 0ULL;  // fold the reduction slots
  for (int q = 0; q < NSLOT; q++) {
    if (h.smin[q] < mn) mn = h.smin[q];
    if (h.smax[q] > mx) mx = h.smax[q];
  }
  memcpy(&out[3], &mx, sizeof(double));
  memcpy(&out[4], &mn, sizeof(double));
  out[5] = (double)h.nanflag;
  return GK_OK;
}

extern "C" int gk_set_time(gk_solver *s, double t) {
  if (!s) return GK_ERR_ARG;
  double *pin = (double *)(s->ctl_host + 2);
  HIPCHK(hipStreamSynchronize(s->stream));
  *pin = t;
  HIPCHK(hipMemcpyAsync(&s->sc->time, pin, sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return GK_OK;
}

extern "C" int gk_enable_timing(gk_solver *s, int on) {
  if (!s) return GK_ERR_ARG;
  s->timing = on;
  s->time_only.clear();
  return GK_OK;
}

extern "C" int gk_time_stage(gk_solver *s, const char *name) {
  if (!s) return GK_ERR_ARG;
  s->timing = 1;
  s->time_only = name ? name : "";
  return GK_OK;
}

extern "C" int gk_kernel_times(gk_solver *s, const char **names, double *ms, int *n) {
  if (!s || !n) return GK_ERR_ARG;
  int cnt = (int)s->names.size();
  int m = *n < cnt ? *n : cnt;
  for (int q = 0; q < m; q++) {
    if (names) names[q] = s->names[q].c_str();
    if (ms) ms[q] = s->acc_n[q] ? s->acc_ms[q] / s->acc_n[q] : 0.0;
  }
  *n = cnt;
  return GK_OK;
}

extern "C" int gk_copy_calibrate(size_t bytes, int reps, double *ms) {
  if (!ms || reps < 1 || bytes < 8) return GK_ERR_ARG;
  const size_t n = bytes / 8;
  if ((n + 255) / 256 > 0x7fffffffu) return GK_ERR_ARG;
  double *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GK_OK;
  if (hipMalloc((void **)&a, n * 8) != hipSuccess || hipMalloc((void **)&b, n * 8) != hipSuccess) {
    rc = GK_ERR_NOMEM;
  } else if (hipMemset(a, 0, n * 8) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
             hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = GK_ERR_DEVICE;
  } else {
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, a, b, n);
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, a, b, n);
    (void)hipEventRecord(e1, st);
    float t = 0.0f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess ||
        hipGetLastError() != hipSuccess)
      rc = GK_ERR_DEVICE;
    *ms = (double)t / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

extern "C" int gk_set_tuning(gk_solver *s, int param, int value) {
  if (!s) return GK_ERR_ARG;
  if (s->batch) return GK_ERR_STATE;  // a batch's launch tables hold its shapes
  if (param == GK_TUNE_EMF_STAGE) {
    if (value < 0 || value > 3) return GK_ERR_ARG;
    s->emf_stage = value;
    return GK_OK;
  }
  if (value < 1 || value > 4096) return GK_ERR_ARG;
  if (param == GK_TUNE_CHUNK) {
    s->chunk = value;
    return GK_OK;
  }
  if (param == GK_TUNE_EMF_CHUNK) {
    s->emf_chunk = value;
    return GK_OK;
  }
  return GK_ERR_ARG;
}

extern "C" int gk_get_tuning(const gk_solver *s, int param) {
  if (!s) return GK_ERR_ARG;
  if (param == GK_TUNE_CHUNK) return s->chunk;
  if (param == GK_TUNE_EMF_CHUNK) return s->emf_chunk;
  if (param == GK_TUNE_EMF_STAGE) return s->emf_stage;
  return GK_ERR_ARG;
}

// ===========================================================================
// Batched stepping of many grids: per stage, one launch whose blocks walk a
// work list covering every grid in the batch (grids of one precision share a
// launch); the launches of a step are captured in HIP graphs.  Each grid keeps
// its own dt, time and history, and runs exactly the arithmetic of gk_step.
struct gk_batch {
  std::vector<gk_solver *> mem;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // members' set-up and read-back (gk_batch_side_stream)
  int device = 0;
  bool dirty = true;
  int phase = 0;  // batch steps since the tables were built, mod 6
  char *dargs = nullptr;
  WorkItem *dwork = nullptr;
  // launch groups: FP64, FP32-bulk, and each again with per-point B0
  size_t args_off[NGROUP][6][SK_COUNT];
  int nargs[NGROUP][SK_COUNT];
  size_t work_off[NGROUP][SK_COUNT];
  int nwork[NGROUP][SK_COUNT];
  int use_graphs = 1;
  hipGraphExec_t g1[6] = {};
  hipGraphExec_t g6 = nullptr;
  int timing = 0;
  std::vector<std::string> names;
  std::vector<double> acc_ms;
  std::vector<int> acc_n;
  // non-finite flags of the members (member order), gathered after each
  // launch of steps into pinned host memory
  Scal **dscs = nullptr;
  unsigned int *dflags = nullptr;
  unsigned int *hflags = nullptr;
  int nflags = 0;
  // device time of runs (gk_batch_run_timing): events around the last run
  hipEvent_t run_ev[2] = {};
  bool run_on = false, run_pending = false;
  double run_ms = 0.0;
};

// fold the last run's device time into the total once its end event is done
static void run_fold(gk_batch *b) {
  if (!b->run_pending) return;
  float ms = 0.0f;
  if (hipEventSynchronize(b->run_ev[1]) == hipSuccess &&
      hipEventElapsedTime(&ms, b->run_ev[0], b->run_ev[1]) == hipSuccess)
    b->run_ms += ms;
  b->run_pending = false;
}

static void batch_free_tables(gk_batch *b) {
  for (int p = 0; p < 6; p++)
    if (b->g1[p]) {
      (void)hipGraphExecDestroy(b->g1[p]);
      b->g1[p] = nullptr;
    }
  if (b->g6) {
    (void)hipGraphExecDestroy(b->g6);
    b->g6 = nullptr;
  }
  void *dev[] = {b->dargs, b->dwork, b->dscs, b->dflags};
  for (void *p : dev)
    if (p) (void)hipFreeAsync(p, b->stream);
  pinned_put(b->hflags);
  b->dargs = nullptr;
  b->dwork = nullptr;
  b->dscs = nullptr;
  b->dflags = nullptr;
  b->hflags = nullptr;
  b->nflags = 0;
}

// A grid's blocks of one stage in the logical order its own launch deals
// them to the XCDs (the decode of xcd_block / xcd_patch in that stage's
// kernel).  A batched launch deals consecutive work-list entries to the same
// XCD, so walking each grid in this order keeps the blocks that share halos,
// chunk prologues and faces in one L2, as in a launch of the grid alone.
static void solo_order(int kid, dim3 gr, int idx, std::vector<WorkItem> &out) {
  const int gx = (int)gr.x, gy = (int)gr.y, gz = (int)gr.z;
  int mode = 0;  // 0: x fastest (xcd_block<false>), 1: z fastest (xcd_block<true>), 2: patches
  switch (kid) {
    case SK_SWEEP_I: case SK_SWEEP_J: case SK_EMF_I: case SK_EMF_J: case SK_EMF_IS:
    case SK_EMF_JS: case SK_GHOST:
      mode = 1;
      break;
    case SK_CT: case SK_CELL:
      mode = gy == 1 ? 2 : 0;
      break;
    default:  // SK_VSTRIP_I, SK_VSTRIP_J, SK_SWEEP_K, SK_KREM, SK_EMF_K, SK_MAGS, the ring stages
      mode = 0;
  }
  if (mode == 2) {
    const int npz = (gz + PATCH - 1) / PATCH;
    const long total = (long)(gx + PATCH - 1) / PATCH * PATCH * npz * PATCH;
    for (long L = 0; L < total; L++) {
      const long pp = L / (PATCH * PATCH), w = L % (PATCH * PATCH);
      const int bz = (int)(pp % npz * PATCH + w % PATCH);
      const int bx = (int)(pp / npz * PATCH + w / PATCH);
      if (bx < gx && bz < gz) out.push_back(WorkItem{idx, bx, 0, bz});
    }
    return;
  }
  const long total = (long)gx * gy * gz;
  for (long L = 0; L < total; L++) {
    int bx, by, bz;
    if (mode == 1) {
      bz = (int)(L % gz);
      const long t = L / gz;
      bx = (int)(t % gx);
      by = (int)(t / gx);
    } else {
      bx = (int)(L % gx);
      const long t = L / gx;
      by = (int)(t % gy);
      bz = (int)(t / gy);
    }
    out.push_back(WorkItem{idx, bx, by, bz});
  }
}

static int batch_build(gk_batch *b) {
  HIPCHK(hipStreamSynchronize(b->stream));  // nothing in flight uses the old tables
  batch_free_tables(b);
  std::vector<char> hargs[NGROUP][6][SK_COUNT];
  std::vector<WorkItem> hwork[NGROUP][SK_COUNT];
  std::vector<Launch> plan;
  // A batch that fills the GPU by itself (4e6 cells or more) needs no short
  // marches to do so: its members march at least 16 faces and 4 edges per
  // thread inside it (their own, autotuned lengths apply when stepped
  // alone).  64-grid FP32-bulk ensemble: 6.39 ms per step with the members'
  // own picks, 6.25 with this floor.
  double cells = 0.0;
  for (gk_solver *s : b->mem) cells += (double)s->ni * s->nj * s->nk;
  const bool full = cells >= 4e6;
  for (gk_solver *s : b->mem) {
    const int gp = launch_group(s);
    Rot r = rot_of(s);
    const int own_chunk = s->chunk, own_emf = s->emf_chunk;
    if (full) {
      s->chunk = std::max(s->chunk, 16);
      s->emf_chunk = std::max(s->emf_chunk, 4);
    }
    for (int p = 0; p < 6; p++) {
      plan_step(s, r, plan);
      for (const Launch &L : plan) {
        const size_t sz = stage_arg_size(L.kid);
        std::vector<char> &A = hargs[gp][p][L.kid];
        const int idx = (int)(A.size() / sz);
        const char *src = (const char *)stage_arg_ptr(L.kid, L.u);
        A.insert(A.end(), src, src + sz);
        if (p == 0 && L.kid != SK_BEGIN) solo_order(L.kid, L.grid, idx, hwork[gp][L.kid]);
      }
      r = rot_next(r);
    }
    s->chunk = own_chunk;
    s->emf_chunk = own_emf;
  }
  size_t abytes = 0, nw = 0;
  for (int gp = 0; gp < NGROUP; gp++)
    for (int k = 0; k < SK_COUNT; k++) {
      for (int p = 0; p < 6; p++) {
        b->args_off[gp][p][k] = abytes;
        abytes += (hargs[gp][p][k].size() + 255) / 256 * 256;
      }
      b->nargs[gp][k] = (int)(hargs[gp][0][k].size() / stage_arg_size(k));
      b->work_off[gp][k] = nw;
      b->nwork[gp][k] = k == SK_BEGIN ? b->nargs[gp][k] : (int)hwork[gp][k].size();
      nw += hwork[gp][k].size();
    }
  if (abytes && hipMallocAsync((void **)&b->dargs, abytes, b->stream) != hipSuccess)
    return GK_ERR_NOMEM;
  if (nw && hipMallocAsync((void **)&b->dwork, nw * sizeof(WorkItem), b->stream) != hipSuccess)
    return GK_ERR_NOMEM;
  for (int gp = 0; gp < NGROUP; gp++)
    for (int k = 0; k < SK_COUNT; k++) {
      for (int p = 0; p < 6; p++)
        if (!hargs[gp][p][k].empty())
          HIPCHK(hipMemcpyAsync(b->dargs + b->args_off[gp][p][k], hargs[gp][p][k].data(),
                                hargs[gp][p][k].size(), hipMemcpyHostToDevice, b->stream));
      if (!hwork[gp][k].empty())
        HIPCHK(hipMemcpyAsync(b->dwork + b->work_off[gp][k], hwork[gp][k].data(),
                              hwork[gp][k].size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                              b->stream));
    }
  const int nm = (int)b->mem.size();
  if (nm) {
    std::vector<Scal *> hs(nm);
    for (int q = 0; q < nm; q++) hs[q] = b->mem[q]->sc;
    if (hipMallocAsync((void **)&b->dscs, nm * sizeof(Scal *), b->stream) != hipSuccess ||
        hipMallocAsync((void **)&b->dflags, nm * sizeof(unsigned int), b->stream) != hipSuccess)
      return GK_ERR_NOMEM;
    if (!(b->hflags = (unsigned int *)pinned_get(nm * sizeof(unsigned int)))) return GK_ERR_NOMEM;
    memset(b->hflags, 0, nm * sizeof(unsigned int));
    HIPCHK(hipMemcpyAsync(b->dscs, hs.data(), nm * sizeof(Scal *), hipMemcpyHostToDevice,
                          b->stream));
    b->nflags = nm;
  }
  HIPCHK(hipStreamSynchronize(b->stream));  // the host copies of the tables go out of scope
  b->phase = 0;
  b->dirty = false;
  return GK_OK;
}

// launches of one batch step at phase p
static void batch_launch_step(gk_batch *b, int p, bool timed) {
  for (int k = 0; k < SK_COUNT; k++)
    for (int gp = 0; gp < NGROUP; gp++) {
      const int nw = b->nwork[gp][k];
      if (!nw) continue;
      const int nblocks = (nw + NXCD - 1) / NXCD * NXCD;  // one block per entry (xcd_item)
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (timed) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, b->stream);
      }
      const void *args = b->dargs + b->args_off[gp][p][k];
      const WorkItem *work = b->dwork + b->work_off[gp][k];
      launch_batch_group(gp, k, args, work, nw, nblocks, b->stream);
      if (timed) {
        (void)hipEventRecord(e1, b->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        std::string nm = std::string(sk_name[k]) + ((gp & 1) ? "/f32" : "/f64") +
                         ((gp & 2) ? "/b0" : "") + ((gp & 4) ? "/ut" : "");
        int slot = -1;
        for (size_t q = 0; q < b->names.size(); q++)
          if (b->names[q] == nm) slot = (int)q;
        if (slot < 0) {
          slot = (int)b->names.size();
          b->names.push_back(nm);
          b->acc_ms.push_back(0.0);
          b->acc_n.push_back(0);
        }
        b->acc_ms[slot] += ms;
        b->acc_n[slot] += 1;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
      }
    }
}

static int batch_capture(gk_batch *b, int first_phase, int nsteps, hipGraphExec_t *out) {
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int t = 0; t < nsteps; t++) batch_launch_step(b, (first_phase + t) % 6, false);
  HIPCHK(hipStreamEndCapture(b->stream, &graph));
  hipError_t e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    *out = nullptr;
    return GK_ERR_DEVICE;
  }
  return GK_OK;
}

extern "C" int gk_batch_create(gk_batch **out) {
  if (!out) return GK_ERR_ARG;
  gk_batch *b = new gk_batch();
  // the batch's stream and its side stream are created together, so that
  // the runtime maps them to different hardware queues: work on the side
  // stream then never waits behind the batch's steps
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return GK_ERR_DEVICE;
  }
  if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess) {
    (void)hipStreamDestroy(b->stream);
    delete b;
    return GK_ERR_DEVICE;
  }
  (void)hipGetDevice(&b->device);
  *out = b;
  return GK_OK;
}

extern "C" int gk_batch_side_stream(gk_batch *b, void **stream) {
  if (!b || !stream) return GK_ERR_ARG;
  *stream = (void *)b->side;
  return GK_OK;
}

extern "C" int gk_batch_destroy(gk_batch *b) {
  if (!b) return GK_OK;
  (void)hipStreamSynchronize(b->stream);
  for (gk_solver *s : b->mem) s->batch = nullptr;
  batch_free_tables(b);
  for (hipEvent_t e : b->run_ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(b->stream);
  // solvers created on the side stream may outlive the batch: the stream
  // stays alive in the pool
  (void)hipStreamSynchronize(b->side);
  stream_put(b->device, b->side);
  delete b;
  return GK_OK;
}

extern "C" int gk_batch_add(gk_batch *b, gk_solver *s) {
  if (!b || !s) return GK_ERR_ARG;
  if (s->ready != 3 || s->batch) return GK_ERR_STATE;
  HIPCHK(hipStreamSynchronize(s->stream));
  // a member steps without a budget until gk_batch_set_budget gives it one
  hipLaunchKernelGGL(k_arm, dim3(1), dim3(64), 0, b->stream, s->sc, -1LL, (double)INFINITY,
                     (double)INFINITY);
  HIPCHK(hipGetLastError());
  s->budget = -1;
  b->mem.push_back(s);
  s->batch = b;
  b->dirty = true;
  return GK_OK;
}

extern "C" int gk_batch_set_budget(gk_batch *b, gk_solver *s, long long steps) {
  if (!b || !s || s->batch != b) return GK_ERR_ARG;
  // ordered after the steps already launched: the budget counts from here
  hipLaunchKernelGGL(k_arm, dim3(1), dim3(64), 0, b->stream, s->sc, steps < 0 ? -1LL : steps,
                     (double)INFINITY, (double)INFINITY);
  HIPCHK(hipGetLastError());
  s->budget = steps < 0 ? -1 : steps;
  return GK_OK;
}

extern "C" int gk_batch_status(gk_batch *b, gk_solver *s) {
  if (!b || !s || s->batch != b) return GK_ERR_ARG;
  for (int q = 0; q < (int)b->mem.size() && q < b->nflags; q++)
    if (b->mem[q] == s) return b->hflags[q] ? GK_ERR_NAN : GK_OK;
  return GK_OK;
}

extern "C" long long gk_batch_budget(gk_batch *b, gk_solver *s) {
  if (!b || !s || s->batch != b) return GK_ERR_ARG;
  return s->budget;
}

extern "C" int gk_batch_remove(gk_batch *b, gk_solver *s) {
  if (!b || !s) return GK_ERR_ARG;
  for (size_t q = 0; q < b->mem.size(); q++)
    if (b->mem[q] == s) {
      HIPCHK(hipStreamSynchronize(b->stream));
      b->mem.erase(b->mem.begin() + q);
      s->batch = nullptr;
      b->dirty = true;
      return GK_OK;
    }
  return GK_ERR_ARG;
}

extern "C" int gk_batch_size(const gk_batch *b) { return b ? (int)b->mem.size() : GK_ERR_ARG; }

extern "C" int gk_batch_set_graphs(gk_batch *b, int on) {
  if (!b) return GK_ERR_ARG;
  b->use_graphs = on ? 1 : 0;
  return GK_OK;
}

extern "C" int gk_batch_enable_timing(gk_batch *b, int on) {
  if (!b) return GK_ERR_ARG;
  b->timing = on ? 1 : 0;
  return GK_OK;
}

extern "C" int gk_batch_kernel_times(gk_batch *b, const char **names, double *ms, int *n) {
  if (!b || !n) return GK_ERR_ARG;
  int cnt = (int)b->names.size();
  int m = *n < cnt ? *n : cnt;
  for (int q = 0; q < m; q++) {
    if (names) names[q] = b->names[q].c_str();
    if (ms) ms[q] = b->acc_n[q] ? b->acc_ms[q] / b->acc_n[q] : 0.0;
  }
  *n = cnt;
  return GK_OK;
}

// Launch n batch steps without waiting for them (graphs unless disabled or
// timing); gk_batch_sync waits.
extern "C" int gk_batch_step_async(gk_batch *b, int n) {
  if (!b || n < 0) return GK_ERR_ARG;
  if (b->mem.empty() || n == 0) return GK_OK;
  if (b->dirty) {
    int rc = batch_build(b);
    if (rc != GK_OK) return rc;
  }
  if (b->timing) {
    for (size_t q = 0; q < b->acc_ms.size(); q++) {
      b->acc_ms[q] = 0.0;
      b->acc_n[q] = 0;
    }
  }
  if (b->run_on) {
    run_fold(b);
    HIPCHK(hipEventRecord(b->run_ev[0], b->stream));
  }
  int left = n;
  while (left > 0) {
    if (b->timing || !b->use_graphs) {
      batch_launch_step(b, b->phase, b->timing != 0);
      left -= 1;
      b->phase = (b->phase + 1) % 6;
      continue;
    }
    if (b->phase == 0 && left >= 6) {
      if (!b->g6) {
        int rc = batch_capture(b, 0, 6, &b->g6);
        if (rc != GK_OK) return rc;
      }
      HIPCHK(hipGraphLaunch(b->g6, b->stream));
      left -= 6;
    } else {
      if (!b->g1[b->phase]) {
        int rc = batch_capture(b, b->phase, 1, &b->g1[b->phase]);
        if (rc != GK_OK) return rc;
      }
      HIPCHK(hipGraphLaunch(b->g1[b->phase], b->stream));
      left -= 1;
      b->phase = (b->phase + 1) % 6;
    }
  }
  // non-finite flags of every member, read back with the run
  if (b->nflags) {
    hipLaunchKernelGGL(k_flags, dim3(1), dim3(64), 0, b->stream, b->dscs, b->nflags, b->dflags);
    HIPCHK(hipMemcpyAsync(b->hflags, b->dflags, b->nflags * sizeof(unsigned int),
                          hipMemcpyDeviceToHost, b->stream));
  }
  if (b->run_on) {
    HIPCHK(hipEventRecord(b->run_ev[1], b->stream));
    b->run_pending = true;
  }
  // every grid's history rotation advances by the steps it takes: n, or
  // what is left of its budget (the device stops it there; its later
  // launches do nothing).  A member that stopped short is out of phase with
  // the batch's argument tables, which are rebuilt before the next launch.
  for (gk_solver *s : b->mem) {
    long long took = n;
    if (s->budget >= 0) {
      took = std::min<long long>(n, s->budget);
      s->budget -= took;
    }
    if (took != n) b->dirty = true;
    Rot r = rot_of(s);
    for (long long t = 0; t < took % 6; t++) r = rot_next(r);
    s->cur = r.cur;
    s->b_m1 = r.b_m1;
    s->b_0 = r.b_0;
    s->b_n = r.b_n;
  }
  HIPCHK(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_batch_sync(gk_batch *b) {
  if (!b) return GK_ERR_ARG;
  HIPCHK(hipStreamSynchronize(b->stream));
  run_fold(b);
  for (int q = 0; q < b->nflags; q++)
    if (b->hflags[q]) return GK_ERR_NAN;
  return GK_OK;
}

extern "C" int gk_batch_run_timing(gk_batch *b, int on) {
  if (!b) return GK_ERR_ARG;
  if (on && !b->run_ev[0]) {
    for (hipEvent_t &e : b->run_ev) HIPCHK(hipEventCreate(&e));
  }
  if (on) {
    HIPCHK(hipStreamSynchronize(b->stream));
    b->run_pending = false;
    b->run_ms = 0.0;
  }
  b->run_on = on != 0;
  return GK_OK;
}

extern "C" int gk_batch_run_ms(gk_batch *b, double *ms) {
  if (!b || !ms) return GK_ERR_ARG;
  *ms = b->run_ms;
  return GK_OK;
}

extern "C" int gk_batch_step(gk_batch *b, int n) {
  int rc = gk_batch_step_async(b, n);
  if (rc != GK_OK) return rc;
  return gk_batch_sync(b);
}

// Launch-shape autotuner: times `trial_steps` steps for each candidate march
// length on this grid and keeps the fastest.  The state (both time levels,
// face and cell B, scalars, history rotation) is saved first and restored
// afterwards, so results do not depend on tuning.  Picks are cached per
// (ni, nj, nk, coord, precision) for the life of the process.
struct TuneKey {
  int ni, nj, nk, coord, prec;
  bool operator==(const TuneKey &o) const {
    return ni == o.ni && nj == o.nj && nk == o.nk && coord == o.coord && prec == o.prec;
  }
};
struct TunePick {
  int chunk, emf_chunk, emf_stage;
};
static std::vector<std::pair<TuneKey, TunePick>> g_tune_cache;

extern "C" int gk_autotune(gk_solver *s, int trial_steps, int *best) {
  if (!s || trial_steps < 0) return GK_ERR_ARG;
  if (s->ready != 3 || s->batch) return GK_ERR_STATE;
  const TuneKey key{s->ni, s->nj, s->nk, s->spherical, s->precision};
  for (auto &e : g_tune_cache)
    if (e.first == key) {
      s->chunk = e.second.chunk;
      s->emf_chunk = e.second.emf_chunk;
      s->emf_stage = e.second.emf_stage;
      if (best) *best = e.second.chunk;
      return GK_OK;
    }
  if (trial_steps == 0) {  // cache lookup only: no pick yet, nothing timed
    if (best) *best = 0;
    return GK_OK;
  }
  const size_t fs = s->fs;
  const size_t nX = 12 * fs, nB = 9 * fs, nM = 9 * fs;
  double *save = nullptr;
  // stream-ordered snapshot: allocating and freeing it on the solver's own
  // stream does not wait for work in flight elsewhere on the device (a batch
  // quantum the serving loop launched before admitting this grid)
  if (hipMallocAsync((void **)&save, (nX + nB + nM) * sizeof(double) + sizeof(Scal), s->stream) !=
      hipSuccess)
    return GK_ERR_NOMEM;
  const Rot r0 = rot_of(s);
  auto snap = [&](bool restore) -> int {
    double *p = save;
    const hipMemcpyKind kd = hipMemcpyDeviceToDevice;
    for (int q = 0; q < 2; q++, p += 6 * fs)
      HIPCHK(restore ? hipMemcpyAsync(s->X[q], p, 6 * fs * 8, kd, s->stream)
                     : hipMemcpyAsync(p, s->X[q], 6 * fs * 8, kd, s->stream));
    for (int q = 0; q < 3; q++, p += 3 * fs)
      HIPCHK(restore ? hipMemcpyAsync(s->B[q], p, 3 * fs * 8, kd, s->stream)
                     : hipMemcpyAsync(p, s->B[q], 3 * fs * 8, kd, s->stream));
    for (int q = 0; q < 3; q++, p += 3 * fs)
      HIPCHK(restore ? hipMemcpyAsync(s->M[q], p, 3 * fs * 8, kd, s->stream)
                     : hipMemcpyAsync(p, s->M[q], 3 * fs * 8, kd, s->stream));
    HIPCHK(restore ? hipMemcpyAsync(s->sc, p, sizeof(Scal), kd, s->stream)
                   : hipMemcpyAsync(p, s->sc, sizeof(Scal), kd, s->stream));
    // the FP32 copies always equal the rounded FP64 state: rebuild them
    if (restore) refresh_copies32(s);
    HIPCHK(hipStreamSynchronize(s->stream));
    s->cur = r0.cur;
    s->b_m1 = r0.b_m1;
    s->b_0 = r0.b_0;
    s->b_n = r0.b_n;
    return GK_OK;
  };
  int rc = snap(false);
  const int maxlen = std::max(s->ni, std::max(s->nj, s->nk)) + 1;
  const int keep_timing = s->timing;
  s->timing = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == GK_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = GK_ERR_DEVICE;
  // time trial_steps steps of the current shape (after one warm-up step);
  // on grids of up to ~1.3M cells (steps of about a millisecond or less) the
  // fastest of three repeats, so that launch jitter does not decide the pick
  const int repeats = (long)s->ni * s->nj * s->nk <= (1L << 20) + (1L << 18) ? 3 : 1;
  auto trial = [&](float &ms) -> int {
    // trials step without a budget or stop time (every stepping entry point
    // arms the grid again afterwards)
    hipLaunchKernelGGL(k_arm, dim3(1), dim3(64), 0, s->stream, s->sc, -1LL, (double)INFINITY,
                       (double)INFINITY);
    one_step(s);
    ms = 1e30f;
    for (int rep = 0; rep < repeats; rep++) {  // back to back; one restore after
      float t_ms = 0.0f;
      (void)hipEventRecord(e0, s->stream);
      for (int t = 0; t < trial_steps; t++) one_step(s);
      (void)hipEventRecord(e1, s->stream);
      if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t_ms, e0, e1) != hipSuccess)
        return GK_ERR_DEVICE;
      ms = std::min(ms, t_ms);
    }
    return snap(true);
  };
  // march length of the sweeps, then of the edge-field kernels, then which
  // of E_i and E_j run in the LDS-staged form.  Candidates go from short to
  // long marches (more to fewer threads); a longer one must be 3% faster to
  // be taken, so near-ties keep the more parallel shape, which also suits a
  // grid stepped inside a batch
  const float kPrefer = 0.97f;
  const int cand[] = {8, 16, 32, 64, 128, 256};
  const int cand_e[] = {2, 4, 8, 16, 32};
  int pick = s->chunk, pick_e = s->emf_chunk, pick_s = s->emf_stage;
  float best_ms = 1e30f;
  for (int c : cand) {
    if (rc != GK_OK) break;
    if (c > 2 * maxlen && c != cand[0]) break;
    s->chunk = c;
    float ms = 0.0f;
    if ((rc = trial(ms)) != GK_OK) break;
    if (ms < best_ms * kPrefer) {
      best_ms = ms;
      pick = c;
    }
  }
  s->chunk = pick;
  best_ms = 1e30f;
  for (int c : cand_e) {
    if (rc != GK_OK) break;
    if (c > 2 * maxlen && c != cand_e[0]) break;
    s->emf_chunk = c;
    float ms = 0.0f;
    if ((rc = trial(ms)) != GK_OK) break;
    if (ms < best_ms * kPrefer) {
      best_ms = ms;
      pick_e = c;
    }
  }
  s->emf_chunk = pick_e;
  // the staged forms cover whole 64-edge tiles along k: without one, every
  // setting launches the same kernels
  if (s->nk + 1 >= 64) {
    best_ms = 1e30f;
    const int dflt = s->emf_stage;
    for (int c : {dflt, 0, 1, 2, 3}) {
      if (rc != GK_OK) break;
      if (c == dflt && best_ms < 1e30f) continue;
      s->emf_stage = c;
      float ms = 0.0f;
      if ((rc = trial(ms)) != GK_OK) break;
      if (ms < best_ms * kPrefer) {
        best_ms = ms;
        pick_s = c;
      }
    }
  }
  s->timing = keep_timing;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFreeAsync(save, s->stream);
  (void)hipStreamSynchronize(s->stream);
  if (rc != GK_OK) return rc;
  s->chunk = pick;
  s->emf_chunk = pick_e;
  s->emf_stage = pick_s;
  g_tune_cache.push_back({key, TunePick{pick, pick_e, pick_s}});
  if (best) *best = pick;
  return GK_OK;
}
