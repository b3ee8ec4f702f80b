/*
 * problem.h -- the problem plugin template (reference src/problem.h:5-11).
 * A plugin defines these three functions; the facade calls
 * set_problem_config() before reading config.yaml, problem_init() on a fresh
 * start and boundary_conditions() to learn which boundary kinds to apply.
 */
#ifndef PROBLEM_H
#define PROBLEM_H

void set_problem_config();
void problem_init();
void boundary_conditions();

#endif
