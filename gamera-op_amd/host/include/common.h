/*
 * common.h -- field-function pointers and the boundary-condition library
 * available to problem plugins (reference src/common.h:1-88).
 *
 * On the device path the boundary routines do not loop over arrays each step:
 * the facade calls the plugin's boundary_conditions() once with the host
 * arrays, applies each requested routine there, and records the kind per
 * array/face so the device ghost-fill kernel applies the same mirror every
 * step.  Every routine is supported on the device: symmetric, reflective (gas
 * only, as in the reference), extrapolated, fixed, pole (curvilinear.h) and
 * none.  Fixed leaves p_S untouched, so where a fixed side meets another one
 * p_S depends on the call order; the device reproduces the order i, j, k
 * (low, high) for gas, then gem.
 */
#ifndef COMMON_H
#define COMMON_H

typedef double (*Field)(double x, double y, double z);
extern Field A1, A2, A3, B10, B20, B30, Source1_B0, Source2_B0, Source3_B0;

double Field_Default(double x, double y, double z);
double B10_uniform(double x, double y, double z);
double B20_uniform(double x, double y, double z);
double B30_uniform(double x, double y, double z);

#define GK_BC_DECL(arr, kind)   \
  void arr##_bc_##kind##_i_low();  \
  void arr##_bc_##kind##_i_high(); \
  void arr##_bc_##kind##_j_low();  \
  void arr##_bc_##kind##_j_high(); \
  void arr##_bc_##kind##_k_low();  \
  void arr##_bc_##kind##_k_high();

GK_BC_DECL(gas, symmetric)
GK_BC_DECL(gas, reflective)
GK_BC_DECL(gas, fixed)
GK_BC_DECL(gas, extrapolated)
GK_BC_DECL(gas, none)
GK_BC_DECL(gem, symmetric)
GK_BC_DECL(gem, reflective)
GK_BC_DECL(gem, fixed)
GK_BC_DECL(gem, extrapolated)
GK_BC_DECL(gem, none)

#endif
