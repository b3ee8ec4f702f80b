/*
 * curvilinear.h -- coordinate-system hooks (reference src/curvilinear.h:57-120).
 * The coordinate system is chosen at run time (gk_coord_type: 0 Cartesian,
 * 2 spherical) instead of at build time.
 */
#ifndef CURVILINEAR_H
#define CURVILINEAR_H

#include "config.h"

extern int gk_coord_type;

double H1(double x1, double x2, double x3);
double H2(double x1, double x2, double x3);
double H3(double x1, double x2, double x3);
int set_geometry_arrays();
void getweights();

void cartesian_to_curvilinear_coord(double x, double y, double z, double *x1,
                                    double *x2, double *x3);
void curvilinear_to_cartesian_coord(double x1, double x2, double x3, double *x,
                                    double *y, double *z);
void cartesian_to_curvilinear_vector(double x, double y, double z, double vx,
                                     double vy, double vz, double *v1,
                                     double *v2, double *v3);
void curvilinear_to_cartesian_vector(double x1, double x2, double x3,
                                     double v1, double v2, double v3,
                                     double *vx, double *vy, double *vz);

/* pole boundary and ring-average set-up (spherical grids reaching a pole;
   src/curvilinear.h:80-99) */
void gas_bc_pole_j_low();
void gas_bc_pole_j_high();
void gem_bc_pole_j_low();
void gem_bc_pole_j_high();
void set_RingAverage_geo_data();
void set_RingAverage_config();

#endif
