/*
 * sanitize_check.c -- TEST INFRASTRUCTURE ONLY: a driver that exercises every
 * entry of the CPU oracle (lddmm_oracle.c) on small 2D and 3D problems, for a
 * build under AddressSanitizer and UndefinedBehaviorSanitizer
 * (oracle/Makefile target `sanitize`; tests/test_oracle_sanitize.py).  It
 * checks nothing numerically (the other oracle tests do): a run is clean
 * when no sanitizer reports an error and the exit status is 0.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "lddmm_oracle.h"

static double urand(unsigned long long *s)
{
	*s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
	return (double)(*s >> 11) * (1.0 / 9007199254740992.0);
}

static double *rnd(int n, double scale, unsigned long long *s)
{
	double *a = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
	for (int i = 0; i < n; ++i) a[i] = scale * (2.0 * urand(s) - 1.0);
	return a;
}

/* a closed polyline (2D) or a strip of triangles (3D) over n points */
static int *elements(int n, int dim, int *ne)
{
	int *e;
	if (dim == 2) {
		*ne = n;
		e = malloc(sizeof(int) * 2 * (size_t)n);
		for (int k = 0; k < n; ++k) {
			e[2 * k] = k;
			e[2 * k + 1] = (k + 1) % n;
		}
	} else {
		*ne = n - 2;
		e = malloc(sizeof(int) * 3 * (size_t)(n - 2));
		for (int k = 0; k < n - 2; ++k) {
			e[3 * k] = k;
			e[3 * k + 1] = k + 1;
			e[3 * k + 2] = k + 2;
		}
	}
	return e;
}

struct quad {
	int n;
};

static void quad_fun(void *ctx, double *f, double *g, const double *x)
{
	const struct quad *q = ctx;
	*f = 0.0;
	for (int i = 0; i < q->n; ++i) {
		*f += 0.5 * (1.0 + i) * x[i] * x[i];
		g[i] = (1.0 + i) * x[i];
	}
}

static int run_dim(int dim, unsigned long long *s)
{
	const int n = 23, nt = 17;
	int rc = 0;
	double *q = rnd(dim * n, 0.5, s), *a = rnd(dim * n, 1.0, s), *l = rnd(dim * n, 1.0, s);
	double *v = malloc(sizeof(double) * dim * n), *dq = malloc(sizeof(double) * dim * n);
	double *K = malloc(sizeof(double) * n * n);
	for (int order = -1; order <= 4; ++order) {
		rc |= orc_compute_kernel(K, q, order, 0.3, n, dim);
		rc |= orc_compute_kernel_cross(K, q, l, order, 0.3, n, n, dim);
		rc |= orc_multiply_kernel(v, q, a, order, 0.3, n, dim);
		rc |= orc_multiply_kernel_cross(v, q, l, a, order, 0.3, n, n, dim);
		rc |= orc_dq_kernel(dq, q, l, a, order, 0.3, n, dim);
		rc |= orc_dq_kernel_cross(dq, v, q, l, l, a, order, 0.3, n, n, dim);
	}
	int ne;
	int *elm = elements(n, dim, &ne);
	const int vpe = dim;
	int rows = 1;
	/* rows of the adjacency table: 1 + max adjacency * stride */
	int *cnt = calloc((size_t)n, sizeof(int));
	for (int k = 0; k < ne * vpe; ++k) cnt[elm[k]]++;
	for (int i = 0; i < n; ++i)
		if (1 + cnt[i] * (dim == 2 ? 2 : 3) > rows) rows = 1 + cnt[i] * (dim == 2 ? 2 : 3);
	int *ifo = calloc((size_t)rows * n, sizeof(int));
	rc |= orc_get_elm_ifo(ifo, elm, vpe, ne, n) != rows;
	double *tc = rnd(dim * nt, 0.6, s), *tu = rnd(dim * nt, 1.0, s), *ta = rnd(nt, 1.0, s);
	for (int j = 0; j < nt; ++j) {
		double nn = 0.0;
		for (int c = 0; c < dim; ++c) nn += tu[dim * j + c] * tu[dim * j + c];
		for (int c = 0; c < dim; ++c) tu[dim * j + c] /= sqrt(nn);
		ta[j] = fabs(ta[j]) + 0.1;
	}
	const char cts[2] = {'G', 'C'}, dts[4] = {'B', 'L', 'O', 'U'};
	for (int ci = 0; ci < 2; ++ci)
		for (int di = 0; di < 4; ++di) {
			double val, dd, dt, tt;
			rc |= orc_varifold(&val, dq, q, elm, ifo, tc, tu, ta, cts[ci], 0.2, dts[di], 0.7, n, ne, rows, nt, dim);
			rc |= orc_varifold(&val, NULL, q, elm, ifo, tc, tu, ta, cts[ci], 0.2, dts[di], 0.7, n, ne, rows, nt,
			                   dim);
			rc |= orc_varifold_terms(&dd, &dt, &tt, q, elm, tc, tu, ta, cts[ci], 0.2, dts[di], 0.7, ne, nt, dim);
		}
	const int T = 4;
	orc_problem P = {dim, n, T, -1, 0.3, 1.0 / (T - 1), 10.0, q, elm, ifo, ne, rows, nt, tc, tu, ta, 'G', 'B',
	                 0.2, 0.0, NULL};
	const int nv = dim * n * (T - 1);
	double *alp = rnd(nv, 0.05, s), *grd = malloc(sizeof(double) * nv), *traj = malloc(sizeof(double) * dim * n * T);
	double f;
	for (int order = -1; order <= 4; ++order) {
		P.knlOrder = order;
		rc |= orc_objgrd(&f, grd, alp, &P, traj);
	}
	P.knlOrder = -1;
	const int itr = 6;
	double tf[7], tg[7], ts[7];
	int te[7];
	orc_trace tr = {tf, tg, ts, te, 0, 0, 0};
	orc_lbfgs_opts o = {itr, 2, 0, 0.0, 0.0, 0.5};
	double *out = malloc(sizeof(double) * nv);
	rc |= orc_lddmm_run(&P, alp, &o, out, &f, &tr) < 0;
	/* l2 landmark matching */
	double *tl = rnd(dim * n, 0.5, s);
	P.tgtNde = tl;
	rc |= orc_objgrd(&f, grd, alp, &P, NULL);
	orc_l2_dist(&f, dq, q, tl, 3.0, n, dim);
	/* the optimiser pieces on a quadratic, including the failure codes */
	struct quad qc = {5};
	double x[5] = {1, -2, 3, -4, 5}, gx[5], xt[5], gt[5], fx, ft, step;
	int c;
	orc_lbfgs_opts oq = {20, 3, 0, 1e-12, 1e-4, 0.9};
	double qf[21], qg[21], qs[21];
	int qe[21];
	orc_trace tq = {qf, qg, qs, qe, 0, 0, 0};
	rc |= orc_lbfgs(quad_fun, &qc, &fx, gx, x, 5, &oq, &tq) < 0;
	double x0[5] = {1, 1, 1, 1, 1}, g0[5], d0[5];
	quad_fun(&qc, &fx, g0, x0);
	for (int i = 0; i < 5; ++i) d0[i] = -g0[i];
	orc_line_search(quad_fun, &qc, x0, g0, d0, fx, 0.9, 0.1, 0.0, xt, gt, &ft, &step, &c, 5, 1);
	orc_line_search(quad_fun, &qc, x0, g0, g0, fx, 0.0, 0.5, 0.0, xt, gt, &ft, &step, &c, 5, 1);
	double S[10], Y[10], rec[2], d[5];
	for (int i = 0; i < 10; ++i) {
		S[i] = urand(s);
		Y[i] = S[i] * (1.0 + (i % 5));
	}
	orc_get_direction(d, 0.5, g0, S, Y, 1, 2, 2, rec, 5);
	double p0, p1;
	for (double xb = 0.0; xb < 800.0; xb = xb * 1.7 + 0.01) orc_bessel_p01(xb, &p0, &p1);
	free(q); free(a); free(l); free(v); free(dq); free(K); free(elm); free(cnt); free(ifo);
	free(tc); free(tu); free(ta); free(alp); free(grd); free(traj); free(out); free(tl);
	return rc;
}

int main(void)
{
	unsigned long long s = 12345;
	int rc = run_dim(2, &s) | run_dim(3, &s);
	if (rc) {
		fprintf(stderr, "an oracle entry returned an error\n");
		return 1;
	}
	printf("sanitize check done\n");
	return 0;
}
