/* Oracle interface for expression programs (test infrastructure only). */
#ifndef GH_ORACLE_CROWN_DAG_H
#define GH_ORACLE_CROWN_DAG_H
#include "../include/gearhip_program.h"
int gho_dag_crown(const gh_prog *P, int B, const double *xL, const double *xU, double *lb,
                  int threads);
int gho_dag_crown_planes(const gh_prog *P, int B, const double *xL, const double *xU,
                         double *lb, double *lA, double *lbias, uint8_t *undef, int threads);
int gho_dag_ibp(const gh_prog *P, int B, const double *xL, const double *xU, double *lo,
                double *hi);
int gho_dag_eval(const gh_prog *P, int N, const double *X, double *F);
void gho_dag_needs(const gh_prog *P, uint8_t *needs);
int gho_math(int fn, int N, const double *x, double *y);
void gho_dag_set_flags(int flags);
#endif
