/*
 * solver.h -- solver entry points of the host facade (reference
 * src/solver.h:10-68): same names and signatures; solve() runs the time step
 * on the GPU through the gamera_op C ABI.
 */
#ifndef SOLVER_H
#define SOLVER_H

#include <stdbool.h>

#include "common.h"
#include "config.h"

void AinitB(double ***magi, double ***magj, double ***magk, Field A1, Field A2,
            Field A3, Field B10, Field B20, Field B30);
double GaussianLineIntegral(double (*fx)(double, double, double), double xa,
                            double ya, double za, double xb, double yb,
                            double zb);

int initialize_solver(const char *base_name);
int allocate_solver();
int set_domain_decomposition();
int finalize_solver();
double get_dt();
void get_derived_variables();
int solve(int nt);
void set_background_field();
void check_positivity();

double factorial(int k);
double nchoosek(int j, int k);

#endif
