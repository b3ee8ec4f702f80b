/*
 * gk_host.h -- internal interfaces of the host facade (not part of the
 * plugin template).
 */
#ifndef GK_HOST_H
#define GK_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "gamera_op.h"

/* recorded boundary kinds (per array and face) */
enum {
  GK_BCK_UNSET = 0,
  GK_BCK_SYMMETRIC,
  GK_BCK_REFLECTIVE,
  GK_BCK_FIXED,
  GK_BCK_EXTRAPOLATED,
  GK_BCK_NONE,
  GK_BCK_POLE,
};

/* ring-average blocks per row for nk azimuthal cells
   (src/spherical/spherical_RingAverage.c:55-79): returns the number of
   averaged rows (NAVR) and fills nchunk, or -1 for an unsupported nk */
int gk_ring_table(int nk, int nchunk[16]);
extern int gk_bc_kind[2][6];
extern int gk_bc_record_only;

typedef struct {
  int mode; /* 0 none, 1 one value per array (constants below), 2 per point */
  double cell[3];
  double face[3][8];
  double edge[3][3];
  double lorentz[3];
} gk_b0_info;
extern gk_b0_info gk_b0;
extern int gk_b0_full_arrays;
int gk_b0_active(void);

double ***gk_alloc3(void);
void gk_free3(double ***a);
double ****gk_alloc4(int nf, const unsigned char *keep);
void gk_free4(double ****a, int nf);

int gk_config_set_sizes(int ni_global, int nj_global, int nk_global, const int dims[3]);
int gk_config_finish(void);
void gk_default_decomposition(void);

/* per-axis device tables built from the host geometry */
typedef struct {
  double *geom[GK_G_COUNT][3];
  double *cc[3][3];   /* cell-centre B coefficients: x[n+1]-xctr[n], xctr[n]-x[n], dx[n] */
  double *recw[3];    /* class-indexed reconstruction tables per sweep axis */
  int *cls[3][2];
  int ncls[3][2];
  int exact;          /* 1 when every transverse width class got its own table */
} gk_tables;
int gk_tables_build(gk_tables *t);
void gk_tables_free(gk_tables *t);
int gk_fill_desc(gk_grid_desc *d, const gk_tables *t, int precision);

/* HDF5 dumps, restart and grid files (built when HDF5 is available) */
extern int gk_io_enabled;
int dump_data_hdf5(const char *base);
int dump_extra_data(const char *base);
int read_data_hdf5(const char *filename);
int gk_read_restart_file(const char *base);
int gk_read_grid_file(void);
void get_out_filename(const char *base, char *filename, size_t size, int seq, const char *ext);

#endif
