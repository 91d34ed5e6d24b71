/*
 * lddmm_oracle.h -- TEST INFRASTRUCTURE ONLY: declarations of the FP64 CPU
 * restatement of the reference's Cpp/src path.  Only tests, the smoke check
 * and the CPU-baseline leg of the benchmark may use it.
 */
#ifndef LDDMM_ORACLE_H
#define LDDMM_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_ERR_ARG (-1)

typedef struct {
	int dim;                 /* 2 or 3 */
	int ndeNum;              /* N */
	int timeNum;             /* T */
	int knlOrder;            /* -1 Gaussian, 0..4 Matern */
	double knlWidth, timeStp, disWgt;
	const double *iniNdeMat; /* D x N */
	const int *elmVtx;       /* V x E, 0-based */
	const int *elmIfo;       /* ifoRows x N */
	int elmNum, ifoRows, tgtNum;
	const double *tgtCen, *tgtDir, *tgtVol; /* D x Et, D x Et, Et */
	char cenType, dirType;
	double cenWidth, dirWidth;
	const double *tgtNde;    /* D x N landmark targets: when non-null the matching term is
	                            l2Dist (MATLAB/src/l2Dist.m:1-5) instead of the varifold */
} orc_problem;

typedef struct {
	int itrMax, memNum, vbsFlg;
	double tolVal, wolfe1, wolfe2;
} orc_lbfgs_opts;

typedef struct {
	double *f, *gnorm, *step; /* itrMax + 1 entries each */
	int *evals;
	int iters, status, totalEvals;
} orc_trace;

typedef void (*orc_objfun)(void *ctx, double *f, double *g, const double *x);

void orc_bessel_p01(double x, double *p0, double *p1);

int orc_compute_kernel(double *K, const double *q, int order, double w, int n, int dim);
int orc_compute_kernel_cross(double *K, const double *qi, const double *qj, int order, double w,
                             int ni, int nj, int dim);
int orc_multiply_kernel(double *v, const double *q, const double *a, int order, double w, int n, int dim);
int orc_multiply_kernel_cross(double *v, const double *qi, const double *qj, const double *a,
                              int order, double w, int ni, int nj, int dim);
int orc_dq_kernel(double *dq, const double *q, const double *l, const double *r,
                  int order, double w, int n, int dim);
int orc_dq_kernel_cross(double *dqi, double *dqj, const double *qi, const double *qj,
                        const double *l, const double *r, int order, double w,
                        int ni, int nj, int dim);

void orc_element_geometry(double *cen, double *dir, double *vol, const double *q,
                          const int *elmVtx, int ne, int dim);
int orc_varifold_terms(double *dd, double *dt, double *tt, const double *q, const int *elmVtx,
                       const double *tgtCen, const double *tgtDir, const double *tgtVol,
                       char cenType, double cenWidth, char dirType, double dirWidth,
                       int ne, int nt, int dim);
int orc_varifold(double *val, double *dq, const double *q, const int *elmVtx, const int *elmIfo,
                 const double *tgtCen, const double *tgtDir, const double *tgtVol,
                 char cenType, double cenWidth, char dirType, double dirWidth,
                 int nq, int ne, int ifoRows, int nt, int dim);
int orc_get_elm_ifo(int *ifo, const int *elmVtx, int vpe, int ne, int nq);

int orc_objgrd(double *fcn, double *grd, const double *alp, const orc_problem *P, double *ndeStkOut);
/* l2Dist: fcn = w/2 |q - tgt|^2, grd = w (q - tgt) (MATLAB/src/l2Dist.m:1-5) */
void orc_l2_dist(double *fcn, double *grd, const double *q, const double *tgt, double w, int n, int dim);

void orc_get_direction(double *d, double H0, const double *g, const double *S, const double *Y,
                       int newIdx, int hisNum, int memNum, double *rec, size_t n);
int orc_line_search(orc_objfun f, void *ctx, const double *x, const double *g, const double *d,
                    double fnow, double c1, double c2, double tol, double *xt, double *gt,
                    double *ft, double *step, int *cnt, size_t n, int quiet);
int orc_lbfgs(orc_objfun f, void *ctx, double *fx, double *gx, double *x, size_t n,
              const orc_lbfgs_opts *o, orc_trace *tr);
int orc_lddmm_run(const orc_problem *P, const double *iniAlpVec, const orc_lbfgs_opts *o,
                  double *optAlpVec, double *fcnOut, orc_trace *tr);

#ifdef __cplusplus
}
#endif

#endif
