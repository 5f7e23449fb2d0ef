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
#define ORC_BC_PERIODIC 0
#define ORC_BC_DIRICHLET 1

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
void orc_set_bc(int bc);
void orc_boundary_u(double *u, int n);
void orc_boundary_fill0(double *u, int n);
void orc_boundary_extrap(double *u, int n);
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
void orc_exact_basis(int n, double h, double *Q, double *lam);
int orc_exact_periodic(double *u, const double *rhs, int n, double h);
int orc_exact_periodic_dense(double *u, const double *rhs, int n, double h);
void orc_gradient(double *g, double *u, int n, double h);

void orc_to_ghosted(double *ug, const double *dense, int n);
void orc_from_ghosted(double *dense, const double *ug, int n);
double orc_relax_dense(double *u, const double *rhs, int n, double h);
void orc_resid_dense(double *res, const double *u, const double *rhs, int n, double h);
void orc_rstrct_dense(double *uc, const double *uf, int nc, int cubic);
void orc_interp_dense(double *uf, const double *uc, int nf, int cubic);
int orc_mg_cycle_dense(double *psi, const double *src, int n, double h, const orc_params *p, orc_stats *st);
void orc_gradient_dense(double *g, const double *u, int n, double h);

/* ---- AMR (fmg.F90 / vmg.F90) ---- */
#define ORC_AMR_PERIODIC 0
#define ORC_AMR_SPHERE 1
#define ORC_AMR_NCODE 6          /* U, RES, RHS, RHO, SRC, PSI (fmg_data.F90:90) */
#define ORC_AMR_MAXLEV 16
#define ORC_AMR_NBR_NONE 0       /* physical boundary face */
#define ORC_AMR_NBR_SAME 1       /* NeighborSameLevel */
#define ORC_AMR_NBR_PARENT 2     /* NeighborParentLevel */

enum {
    ORC_AMR_OP_C2P = 1, ORC_AMR_OP_GHOSTFIX, ORC_AMR_OP_FILL0, ORC_AMR_OP_EXTRAP, ORC_AMR_OP_RELAX,
    ORC_AMR_OP_RESID, ORC_AMR_OP_RSTRCT, ORC_AMR_OP_INTERP, ORC_AMR_OP_ADDINT, ORC_AMR_OP_VMG_FAS,
    ORC_AMR_OP_FMG_LIN, ORC_AMR_OP_SPHERE_BC, ORC_AMR_OP_REMOVE_MEAN, ORC_AMR_OP_VMG_TAU
};

typedef struct {
    int iterations;
    int status;
    double rel_linf;     /* max |src - L psi| / max |src| over leaf cells */
    double rel_l2;       /* volume-weighted L2 ratio over leaf cells */
    double ratio;        /* Resh2maxg / h2absmax(src) of the last iteration (fmg.F90:183) */
    double hist[ORC_HIST];
    double rel_h2;       /* max_l(max|src - L psi| h_l^2) / max_l(max|src| h_l^2) (fmg.F90:172, 1017-1039) */
} orc_amr_stats;

/* status values, as the library's SG_OK / SG_ERR_NOT_CONVERGED */
#define ORC_OK 0
#define ORC_NOT_CONVERGED 3

typedef struct orc_amr orc_amr;
orc_amr *orc_amr_create(int nblk, int nlev, const int *level, const int *ijk, const int *parent,
                        const int *child, const int *nbr, const int base_dims[3], double h0,
                        const double x0[3], int bc);
void orc_amr_destroy(orc_amr *s);
void orc_amr_set_zero_average(orc_amr *s, int on);
size_t orc_amr_block_len(const orc_amr *s, int L);
int orc_amr_get_fill0_level(const orc_amr *s);
void orc_amr_set_fill0_level(orc_amr *s, int v);
double orc_amr_resh2maxg(const orc_amr *s);
void orc_amr_base_stats(const orc_amr *s, long *calls, long *iters);
void orc_amr_set_array(orc_amr *s, int L, int code, const double *src);
void orc_amr_get_array(const orc_amr *s, int L, int code, double *dst);
double orc_amr_op(orc_amr *s, int op, int L, int a, int b, int c, int flag);
void orc_amr_sphere_bc(orc_amr *s, int ju, int jrho);
int orc_amr_solve(orc_amr *s, double *psi, const double *src, double *g, const orc_params *p, orc_amr_stats *st);

/* ---- hydro block memory on either side of the solve (blockmem.c) ---- */
typedef struct {
    int ng, ncomp;       /* ghost layers (N_GHOST_CELL), components (NM) */
    long stride;         /* doubles from one pool slot to the next */
    int mrho, mvx, mvy, mvz, mp, mpsi, mgx, mgy, mgz;   /* component labels (config.h:63-71) */
} orc_blockmem;

typedef struct {
    int source;          /* fmg_prepare_interp time slice of Rho: 0 current, 1 predictor, 2 past, 3 in time */
    double step_max;     /* Step(LevelMax) */
    double step[ORC_AMR_MAXLEV], dstep[ORC_AMR_MAXLEV];   /* Step(l), Dstep(l) */
} orc_timeslice;

void orc_bm_stage(const double *pool, const int *slot_u, const int *slot_u1, const int *slot_u2, int nblk,
                  const int *level, const orc_blockmem *L, const orc_timeslice *ts, double pi4g, double *psi,
                  double *src);
void orc_bm_restore(double *pool, const int *slot_u, int nblk, const orc_blockmem *L, const double *u_blocks,
                    const double *g);
void orc_bm_source_g(double *pool, const int *slot_u, const int *slot_u1, int nblk, const int *level,
                     const double *dv_lev, const double *dt_lev, const orc_blockmem *L, double gamma);
void orc_bm_grid_boundary(double *pool, const int *su, const int *s1, const int *s2, int nblk, const int *level,
                          const int *parent, const int *child, const int *nbr, const orc_blockmem *L, int lev, int mode);
void orc_bm_fg2cg(double *pool, const int *su, int nblk, const int *level, const int *parent, const int *child,
                  const orc_blockmem *L, int lev, double gamma);
void orc_bm_fluxcorrection(double *pool, const int *su, const double *fb, int nblk, const int *level,
                           const int *parent, const int *child, const int *nbr, const double *h_lev,
                           const double *dt_lev, const orc_blockmem *L, int lev, double gamma);

#endif
