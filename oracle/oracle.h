/*
 * ORACLE -- test infrastructure, not product code.  CPU FP64 restatement of the reference's
 * self-gravity multigrid, used only as the checker by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg.
 */
#ifndef SG_ORACLE_H
#define SG_ORACLE_H
#include <stddef.h>

#define ORC_STOP_REFERENCE 0     /* mg.F90:91  Resmaxg / rhomax < tol            */
#define ORC_STOP_TRUE_RESIDUAL 1 /* max|src - L psi| / max|src| < tol, after update */
#define ORC_HIST 64

typedef struct {
    double tol;
    int maxit;
    int mode;
    int ncycle, npre, npost;
} orc_params;

typedef struct {
    int iterations;
    int status;          /* 0 converged, 1 not converged */
    double rel_linf;     /* max|r| / max|f| after the last update */
    double rel_l2;       /* ||r||_2 / ||f||_2 after the last update */
    double resmaxg;      /* Resmaxg of the last relax call */
    double hist[ORC_HIST];
} orc_stats;

size_t orc_ghosted_len(int n);
void orc_boundary_u(double *u, int n);
double orc_relax(double *u, const double *rhs, int n, double h);
void orc_resid(double *res, double *u, const double *rhs, int n, double h);
void orc_rstrct_linear(double *uc, const double *uf, int nc);
void orc_rstrct_cubic_weights(double s1d[4]);
void orc_rstrct_cubic(double *uc, const double *uf, int nc);
void orc_interp_trilinear(double *uf, const double *uc, int nf);
void orc_cubic_stencil(int f, int lo, int hi, int idx[4], double a[4]);
void orc_interp_tricubic(double *uf, const double *uc, int nf);
void orc_mg_interp(double *uf, double *uc, int nf, int cubic);
int orc_mg_cycle(double *psi, const double *src, int n, double h, const orc_params *p, orc_stats *st);
void orc_gradient(double *g, double *u, int n, double h);

void orc_to_ghosted(double *ug, const double *dense, int n);
void orc_from_ghosted(double *dense, const double *ug, int n);
double orc_relax_dense(double *u, const double *rhs, int n, double h);
void orc_resid_dense(double *res, const double *u, const double *rhs, int n, double h);
void orc_rstrct_dense(double *uc, const double *uf, int nc, int cubic);
void orc_interp_dense(double *uf, const double *uc, int nf, int cubic);
int orc_mg_cycle_dense(double *psi, const double *src, int n, double h, const orc_params *p, orc_stats *st);
void orc_gradient_dense(double *g, const double *u, int n, double h);

#endif
