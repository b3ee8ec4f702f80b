/*
 * config.h -- solver configuration and data layout of the GAMERA-OP host
 * facade.
 *
 * Problem plugins written for the reference (src/problem.h template) include
 * this header by name and index the state arrays with these enums, so the
 * names, values and array shapes here are an interface that must stay
 * identical to the reference's src/config.h:1-499: NG = 4 ghost layers,
 * arrays [f][i][j][k] with k fastest and NI = ni + 2*NG + 1 points per axis.
 * The host stores only the fields the device path exchanges with the host:
 * gas fields 0..17, gem fields 0..17, the vector-potential scratch LA*, the
 * background-field block and magtot; pointers of other fields are NULL.
 */
#ifndef CONFIG_H
#define CONFIG_H

#include <stdbool.h>
#include <stdio.h>

#define NG 4
#define NO2 NG
#define NO (NO2 * 2)
#define PI 3.14159265358979323846
#define NS 1
#define NS1 NS

#define out_prefix_pattern "%s_%02d-%02d-%02d_"
#define out_file_pattern (out_prefix_pattern "%06d.%s")
#define BASE_DAT_NAME "mhd"
#define BASE_EXT_NAME "mhd"
#define BASE_LOG_NAME "LOG"

typedef enum { BC_NON_PERIODIC = 1, BC_PERIODIC } Boundary_condition_t;

typedef struct {
  int ni_global, nj_global, nk_global;
  int ni, nj, nk;
  int NI, NJ, NK;
  int proc_dims[3];
  int proc_coords[3];
} config_t;

typedef struct {
  Boundary_condition_t Boundary_i;
  Boundary_condition_t Boundary_j;
  Boundary_condition_t Boundary_k;
} Problem_config_t;

typedef struct {
  double mu0, x_Norm, u_Norm, Time_Norm, rho_Norm, p_Norm, B_Norm,
      Magnetic_moment_Norm, Omega_Norm, Sigma_Norm;
} Norm_t;

extern config_t config;
extern int *proc_coords;
extern int hdf_seq_num;
extern int log_seq_num;
extern bool read_restart;
extern Norm_t norm_config;
extern int Nt;
extern double time_stop;
extern double output_interval;
extern FILE *log_file;
extern int pdm_index_gas;
extern int pdm_index_gem;
extern Problem_config_t problem_config;
extern double x1min_global, x1max_global, x2min_global, x2max_global, x3min_global,
    x3max_global;
extern double x1min, x1max, x2min, x2max, x3min, x3max;
extern int doRingAverage;
extern int doRingAverage_thisrank_positivepole;
extern int doRingAverage_thisrank_negativepole;
extern int Nchunk[32];
extern int NAVR;

extern const double gamma_val;
extern const double CFL;
extern const double PDMB;
extern double CA;
extern const double nlf_cs;
extern const double nlf_va;
extern double rho_floor;
extern double p_floor;
extern double G01, G02, G03;
extern double divB_max;
extern const double fac8th[8];
extern const double facPDMU7[7];

extern int is, ie, isg, ieg;
extern int js, je, jsg, jeg;
extern int ks, ke, ksg, keg;
extern double dtout, dt, dt0, dxx, dyy, dzz, time_sim, time_sim_start;

extern double ***x1, ***x2, ***x3;
extern double ***x1c, ***x2c, ***x3c;
extern double ***x1ctr, ***x2ctr, ***x3ctr;
extern double ***dx1, ***dx2, ***dx3;

typedef enum { i_dir = 0, j_dir, k_dir } dir_t;

extern double ****geo;
enum {
  edge_idir = 0, edge_jdir, edge_kdir,
  dxi, dxj, dxk, dxi_ring, dxj_ring, dxk_ring,
  face_idir, face_jdir, face_kdir,
  vol_idir, vol_jdir, vol_kdir, vol_center,
  GF_face_idir, GF_face_jdir, GF_face_kdir,
  GF_vol_idir, GF_vol_jdir, GF_vol_kdir,
  IfaceAedgeJ, IfaceAedgeK, JfaceAedgeK, JfaceAedgeI, KfaceAedgeI, KfaceAedgeJ,
  NF_geo
};
extern int geo_onface_i[NF_geo];
extern int geo_onface_j[NF_geo];
extern int geo_onface_k[NF_geo];

extern double ****rec;
enum {
  w1l_first_direction = 0, w2l_first_direction, w3l_first_direction,
  w4l_first_direction, w5l_first_direction, w6l_first_direction,
  w7l_first_direction,
  w1l_second_direction, w2l_second_direction, w3l_second_direction,
  w4l_second_direction, w5l_second_direction, w6l_second_direction,
  w7l_second_direction,
  w1l_third_direction, w2l_third_direction, w3l_third_direction,
  w4l_third_direction, w5l_third_direction, w6l_third_direction,
  w7l_third_direction,
  w1r_first_direction, w2r_first_direction, w3r_first_direction,
  w4r_first_direction, w5r_first_direction, w6r_first_direction,
  w7r_first_direction,
  w1r_second_direction, w2r_second_direction, w3r_second_direction,
  w4r_second_direction, w5r_second_direction, w6r_second_direction,
  w7r_second_direction,
  w1r_third_direction, w2r_third_direction, w3r_third_direction,
  w4r_third_direction, w5r_third_direction, w6r_third_direction,
  w7r_third_direction,
  NF_rec
};

extern double ****gem;
enum {
  mag_bi = 0, mag_bj, mag_bk,
  mags_b1, mags_b2, mags_b3,
  mag_bi_h, mag_bj_h, mag_bk_h,
  mag_bi_p, mag_bj_p, mag_bk_p,
  mags_b1_h, mags_b2_h, mags_b3_h,
  mags_b1_p, mags_b2_p, mags_b3_p,
  bi_avg, bj_avg, bk_avg, bil, bir, bjl, bjr, bkl, bkr, dvzz,
  alf_ratio, perp_ratio, dv_alf,
  mag0_bi, mag0_bj, mag0_bk,
  magsl_b1, magsl_b2, magsl_b3,
  magsr_b1, magsr_b2, magsr_b3,
  magsf_x1dir, magsf_x2dir, magsf_x3dir,
  dmagsf_x1dir, dmagsf_x2dir, dmagsf_x3dir,
  efield_ei, efield_ej, efield_ek,
  LAi, LAj, LAk,
  B0_b1, B0_b2, B0_b3,
  B0AfaceI_b1, B0AfaceI_b2, B0AfaceI_b3, B0AfaceI_bn, B0AfaceI_bsq,
  B0AfaceI_bn1, B0AfaceI_bn2, B0AfaceI_bn3,
  B0AfaceJ_b1, B0AfaceJ_b2, B0AfaceJ_b3, B0AfaceJ_bn, B0AfaceJ_bsq,
  B0AfaceJ_bn1, B0AfaceJ_bn2, B0AfaceJ_bn3,
  B0AfaceK_b1, B0AfaceK_b2, B0AfaceK_b3, B0AfaceK_bn, B0AfaceK_bsq,
  B0AfaceK_bn1, B0AfaceK_bn2, B0AfaceK_bn3,
  B0AedgeI_b1, B0AedgeI_b2, B0AedgeI_b3,
  B0AedgeJ_b1, B0AedgeJ_b2, B0AedgeJ_b3,
  B0AedgeK_b1, B0AedgeK_b2, B0AedgeK_b3,
  dmagsf_G_x1dir, dmagsf_G_x2dir, dmagsf_G_x3dir,
  S1_B, S2_B, S3_B,
  magtot_b1, magtot_b2, magtot_b3,
  NF_gem
};
extern int gem_onface_i[NF_gem];
extern int gem_onface_j[NF_gem];
extern int gem_onface_k[NF_gem];
extern int NF_gem_prim;
extern int NF_gem_store;

extern double *****gas;
enum {
  gas_rho = 0, gas_v1, gas_v2, gas_v3, gas_p, gas_p_S,
  gas_rho_h, gas_v1_h, gas_v2_h, gas_v3_h, gas_p_h, gas_p_S_h,
  gas_rho_p, gas_v1_p, gas_v2_p, gas_v3_p, gas_p_p, gas_p_S_p,
  v1_interp_l, v2_interp_l, v3_interp_l,
  v1_interp_r, v2_interp_r, v3_interp_r,
  v1_interp, v2_interp, v3_interp,
  v1_avg, v2_avg, v3_avg,
  gasl_rho, gasl_v1, gasl_v2, gasl_v3, gasl_p, gasl_p_S,
  gasr_rho, gasr_v1, gasr_v2, gasr_v3, gasr_p, gasr_p_S,
  gasc_rho, gasc_rhov1, gasc_rhov2, gasc_rhov3, gasc_eng, gasc_entropy,
  gascf_rho, gascf_rhov1, gascf_rhov2, gascf_rhov3, gascf_eng, gascf_entropy,
  gasc0_rho, gasc0_rhov1, gasc0_rhov2, gasc0_rhov3, gasc0_eng, gasc0_entropy,
  dgascf_rho, dgascf_rhov1, dgascf_rhov2, dgascf_rhov3, dgascf_eng,
  dgascf_entropy,
  v1_tmp, v2_tmp, v3_tmp,
  v1_0, v2_0, v3_0,
  rhov1_fluid, rhov2_fluid, rhov3_fluid,
  rhov1_h, rhov2_h, rhov3_h,
  S1_HD, S2_HD, S3_HD,
  bdotv,
  NF_gas
};
extern int gas_onface_i[NF_gas];
extern int gas_onface_j[NF_gas];
extern int gas_onface_k[NF_gas];
extern int NF_gas_prim;
extern int NF_gas_store;

typedef enum { LOW_BOUNDARY = 0, HIGH_BOUNDARY = 1 } boundary_t;

int read_config(const char *filename);
int initialize_config(const int rank, const int size, const char *filename);
void initialize_grid_indices();
int initialize_logging(const char *base_filename, int rank);

#endif
