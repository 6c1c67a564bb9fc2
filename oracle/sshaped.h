/* Relaxations of sin, cos, odd and non-integer powers (test infrastructure
 * only; see sshaped.c).  R = (lw, lb, uw, ub); the return value holds the
 * undefined sides (1 lower, 2 upper). */
#ifndef GH_ORACLE_SSHAPED_H
#define GH_ORACLE_SSHAPED_H
#include "../include/gearhip_program.h"

/* right-endpoint sin table index rounded down, as the reference computed it
 * when it wrote solver/tests/data_64/s_shape_test_data */
#define GHO_SIN_LEGACY_INDEX 1

int gho_relax_sin(double l, double u, int flags, double R[4]);
int gho_relax_cos(double l, double u, int flags, double R[4]);
int gho_relax_pow_odd(int p, double l, double u, double R[4]);
int gho_relax_pow_frac(double p, double l, double u, double R[4]);
void gho_sshape_init(int pow_mask);
int gho_sshape_tables(int which, int p, double *out);
int gho_relax_unary(int op, double c, int N, const double *l, const double *u, double *R,
                    int flags);
#endif
