/*
 * oracle.h -- CPU restatement of OpFlow's hot path. TEST INFRASTRUCTURE ONLY:
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * never by the product library. See oracle.c for the per-function citations
 * and how each function is pinned.
 */
#ifndef OPFLOW_ORACLE_H
#define OPFLOW_ORACLE_H
#include "../include/opflow_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    const double* x[3];  /* node coordinates over [start-pad, start+nnodes+pad) */
    const double* dx[3]; /* spacings over [start-pad, start+nnodes+pad-1)       */
    int org[3];          /* index of element 0 (= start - pad)                  */
} orc_mesh;

void orc_mesh_uniform(int nnodes, double lo, double hi, int pad, double* x, double* dx);
int orc_desc_finalize(opf_field_desc* f, const int mesh_start[3], const int mesh_nnodes[3]);
void orc_update_padding(const opf_field_desc* f, const orc_mesh* m, double* u);
void orc_ftcs(const opf_field_desc* f, const orc_mesh* m, double* u, double coef, int nsteps, int nthreads);
double orc_reduce(const opf_field_desc* f, const double* u, int red, const int rs[3], const int re[3]);
int orc_num_threads(void);
void orc_set_threads(int n);
long orc_csr_poisson(const opf_field_desc* f, const orc_mesh* m, double c, int pin_last, long* ptr, long* col,
                     double* val, double* rhs);
void orc_poisson_apply(int dim, const int n[3], const double h[3], const int bc[3][2], double sigma, const double* x,
                       double* y);
int orc_pcg_mg(int dim, const int n[3], const double h[3], const int bc[3][2], double sigma, const double* b,
               double* x, double tol, int maxit, double* relres);

#ifdef __cplusplus
}
#endif
#endif
