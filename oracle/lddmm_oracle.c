/*
 * lddmm_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * An FP64, single-threaded C restatement of the CPU path of the reference
 * (dnhsieh/LDDMM, Cpp/src).  It is the checker the GPU library is compared
 * against and the timed CPU baseline; no product code may link or call it.
 *
 * Layout everywhere in this file is the reference's host ("Cpp") layout:
 * a point set is D x N column-major, i.e. AoS x0 y0 [z0] x1 y1 ... , and a
 * time stack is D x N x (T-1) (Cpp/src/LDDMM.m:44).  Element/vertex tables
 * are int32, 0-based, V x E column-major (Cpp/src/LDDMM.m:8); the adjacency
 * table is (1 + maxAdj*V) x N (Cpp/src/getElmIfo.m:1-59).
 *
 * Kernel orders: -1 Gaussian, 0..4 Matern (Cpp/src/computeKernel.cpp:23-24).
 * In 2D the Matern family needs the Bessel factors p0 = x^2 K0(x) and
 * p1 = x K1(x); they are evaluated here from the ascending series
 * (Abramowitz & Stegun 9.6.10/9.6.13 and 9.6.11) for x <= 2 and from the
 * Steed/Thompson-Barnett continued fraction for K_nu (Temme 1975) for x > 2.
 *
 * Parity pins (tests/): every kernel-level function is compared with the
 * reference's own compiled kernels (oracle/_ref) on golden vectors in
 * tests/golden; objgrd is pinned by finite differences (the discrete adjoint
 * is exact) and by the closed form E = w V(q0) at alpha = 0; the L-BFGS
 * pieces by a dense BFGS inverse update and by crafted 1-D objectives.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lddmm_oracle.h"

/* ------------------------------------------------------------------------ */
/* 2D Matern Bessel factors                                                  */
/* ------------------------------------------------------------------------ */

static const double EULER_GAMMA = 0.57721566490153286060651209;

/* K0 and K1 for 0 < x <= 2 from the ascending series (A&S 9.6.13, 9.6.11):
 *   K0(x) = -(ln(x/2)) I0(x) + sum_k psi(k+1) t^k / (k!)^2,   t = x^2/4
 *   K1(x) = 1/x + ln(x/2) I1(x)
 *           - (x/4) sum_k (psi(k+1) + psi(k+2)) t^k / (k! (k+1)!)
 * with psi(k+1) = -gamma + H_k.                                           */
static void bessel_k01_series(double x, double *k0, double *k1)
{
	double t = 0.25 * x * x;
	double lg = log(0.5 * x);
	double i0 = 0.0, i1 = 0.0, s0 = 0.0, s1 = 0.0;
	double term0 = 1.0;             /* t^k / (k!)^2          */
	double term1 = 1.0;             /* t^k / (k! (k+1)!)     */
	double hk = 0.0;                /* harmonic number H_k   */
	for (int k = 0; k < 60; ++k) {
		double psi1 = -EULER_GAMMA + hk;              /* psi(k+1) */
		double psi2 = psi1 + 1.0 / (double)(k + 1);   /* psi(k+2) */
		i0 += term0;
		i1 += term1;
		s0 += psi1 * term0;
		s1 += (psi1 + psi2) * term1;
		if (term0 < 1e-18 * fabs(i0) && k > 2) break;
		hk += 1.0 / (double)(k + 1);
		term0 *= t / ((double)(k + 1) * (double)(k + 1));
		term1 *= t / ((double)(k + 1) * (double)(k + 2));
	}
	i1 *= 0.5 * x;                                   /* I1 = (x/2) sum t^k/(k!(k+1)!) */
	*k0 = -lg * i0 + s0;
	*k1 = 1.0 / x + lg * i1 - 0.25 * x * s1;
}

/* K0 and K1 for x > 2: Steed's algorithm for the continued fraction of
 * K_{nu+1}/K_nu together with the Thompson-Barnett sum that normalises K_nu
 * (Thompson & Barnett, J. Comput. Phys. 64 (1986) 490; Temme 1975), nu = 0:
 *   K0(x) = sqrt(pi/(2x)) e^{-x} / S,   K1(x) = K0 (x + 1/2 - h) / x,
 * where h = (1/4) * CF2 and S the normalising series.                     */
static void bessel_k01_cf(double x, double *k0, double *k1)
{
	const double a1 = 0.25;          /* nu^2 - 1/4 with sign flipped, nu = 0 */
	double b = 2.0 * (1.0 + x);
	double d = 1.0 / b;
	double h = d, dh = d;
	double qm1 = 0.0, qc = 1.0;
	double a = -a1;
	double c = a1, q = a1;
	double s = 1.0 + q * dh;
	for (int i = 2; i < 10000; ++i) {
		a -= 2.0 * (double)(i - 1);
		c = -a * c / (double)i;
		double qn = (qm1 - b * qc) / a;
		qm1 = qc;
		qc = qn;
		q += c * qn;
		b += 2.0;
		d = 1.0 / (b + a * d);
		dh = (b * d - 1.0) * dh;
		h += dh;
		double ds = q * dh;
		s += ds;
		if (fabs(ds / s) < 1e-17) break;
	}
	h *= a1;
	double kk0 = sqrt(M_PI / (2.0 * x)) * exp(-x) / s;
	*k0 = kk0;
	*k1 = kk0 * (x + 0.5 - h) / x;
}

/* p0 = x^2 K0(x), p1 = x K1(x); limits p0(0)=0, p1(0)=1 as in
 * Cpp/src/polybesselk.cpp:13-17,54-58.                                    */
void orc_bessel_p01(double x, double *p0, double *p1)
{
	if (x < DBL_EPSILON) { *p0 = 0.0; *p1 = 1.0; return; }
	if (x > 700.0) { *p0 = 0.0; *p1 = 0.0; return; }
	double k0, k1;
	if (x <= 2.0) bessel_k01_series(x, &k0, &k1);
	else bessel_k01_cf(x, &k0, &k1);
	*p0 = x * x * k0;
	*p1 = x * k1;
}

/* ------------------------------------------------------------------------ */
/* Scalar kernel family                                                      */
/* ------------------------------------------------------------------------ */

static int order_ok(int order) { return order >= -1 && order <= 4; }

/* k(d), d = |qi - qj| / sigma.
 * Gaussian exp(-d^2)            Cpp/src/computeKernel.cpp:36-37
 * 3D Matern, order p            Cpp/src/computeKernel.cpp:170,186,202,218,234
 * 2D Matern, order p            Cpp/src/computeKernel.cpp:60,83,105,127,149  */
static double knl_value(int order, int dim, double d)
{
	if (order == -1) return exp(-d * d);
	if (dim == 3) {
		double e = exp(-d);
		switch (order) {
		case 0: return e;
		case 1: return (1.0 + d) * e;
		case 2: return (3.0 + d * (3.0 + d)) / 3.0 * e;
		case 3: return (15.0 + d * (15.0 + d * (6.0 + d))) / 15.0 * e;
		default: return (105.0 + d * (105.0 + d * (45.0 + d * (10.0 + d)))) / 105.0 * e;
		}
	}
	double p0, p1, d2 = d * d;
	orc_bessel_p01(d, &p0, &p1);
	switch (order) {
	case 0: return p1;
	case 1: return 0.5 * (p0 + 2.0 * p1);
	case 2: return (4.0 * p0 + (8.0 + d2) * p1) / 8.0;
	case 3: return ((24.0 + d2) * p0 + 8.0 * (6.0 + d2) * p1) / 48.0;
	default: return (12.0 * (16.0 + d2) * p0 + (384.0 + d2 * (72.0 + d2)) * p1) / 384.0;
	}
}

/* c(d) with grad_{qi} k = c(d) (qi - qj); w2 = sigma^2.
 * Gaussian -2 exp(-d^2)/w2                       Cpp/src/dqKernel.cpp:57
 * 3D Matern p=1..4                               Cpp/src/dqKernel.cpp:215,240,265,290
 * 2D Matern p=1..4                               Cpp/src/dqKernel.cpp:91,122,153,184
 * order 0 has no derivative in the reference: the result stays zero
 * (Cpp/src/dqKernel.cpp:39 memset, no case 0).                            */
static double knl_deriv(int order, int dim, double d, double w2)
{
	if (order == -1) return -2.0 * exp(-d * d) / w2;
	if (order == 0) return 0.0;
	if (dim == 3) {
		double e = exp(-d);
		switch (order) {
		case 1: return -e / w2;
		case 2: return -(1.0 + d) * e / (3.0 * w2);
		case 3: return -(3.0 + d * (3.0 + d)) * e / (15.0 * w2);
		default: return -(15.0 + d * (15.0 + d * (6.0 + d))) * e / (105.0 * w2);
		}
	}
	double p0, p1, d2 = d * d;
	orc_bessel_p01(d, &p0, &p1);
	switch (order) {
	case 1: return -p1 / (2.0 * w2);
	case 2: return -(p0 + 2.0 * p1) / (8.0 * w2);
	case 3: return -(4.0 * p0 + (8.0 + d2) * p1) / (48.0 * w2);
	default: return -((24.0 + d2) * p0 + 8.0 * (6.0 + d2) * p1) / (384.0 * w2);
	}
}

static double dist(const double *a, const double *b, int dim)
{
	double s = 0.0;
	for (int k = 0; k < dim; ++k) {
		double t = a[k] - b[k];
		s += t * t;
	}
	return sqrt(s);
}

static double dot(const double *a, const double *b, int dim)
{
	double s = 0.0;
	for (int k = 0; k < dim; ++k) s += a[k] * b[k];
	return s;
}

/* ------------------------------------------------------------------------ */
/* Pair operators                                                            */
/* ------------------------------------------------------------------------ */

/* Dense kernel matrix, self form: K[i*n + j] (Cpp/src/computeKernel.cpp:21-244). */
int orc_compute_kernel(double *K, const double *q, int order, double w, int n, int dim)
{
	if (!order_ok(order) || (dim != 2 && dim != 3)) return ORC_ERR_ARG;
	for (int i = 0; i < n; ++i)
		for (int j = 0; j < n; ++j)
			K[(size_t)i * n + j] = knl_value(order, dim, dist(q + (size_t)i * dim, q + (size_t)j * dim, dim) / w);
	return 0;
}

/* Dense kernel matrix, cross form: K[j*ni + i] (Cpp/src/computeKernel.cpp:246-470). */
int orc_compute_kernel_cross(double *K, const double *qi, const double *qj, int order, double w,
                             int ni, int nj, int dim)
{
	if (!order_ok(order) || (dim != 2 && dim != 3)) return ORC_ERR_ARG;
	for (int j = 0; j < nj; ++j)
		for (int i = 0; i < ni; ++i)
			K[(size_t)j * ni + i] = knl_value(order, dim, dist(qi + (size_t)i * dim, qj + (size_t)j * dim, dim) / w);
	return 0;
}

/* v_i = sum_j k(qi, qj) a_j  (Cpp/src/multiplyKernel.cpp:22-292). */
int orc_multiply_kernel(double *v, const double *q, const double *a, int order, double w, int n, int dim)
{
	return orc_multiply_kernel_cross(v, q, q, a, order, w, n, n, dim);
}

/* cross form, v over the ni points (Cpp/src/multiplyKernel.cpp:294-564). */
int orc_multiply_kernel_cross(double *v, const double *qi, const double *qj, const double *a,
                              int order, double w, int ni, int nj, int dim)
{
	if (!order_ok(order) || (dim != 2 && dim != 3)) return ORC_ERR_ARG;
	memset(v, 0, sizeof(double) * (size_t)dim * ni);
	for (int i = 0; i < ni; ++i) {
		const double *xi = qi + (size_t)i * dim;
		double *vi = v + (size_t)i * dim;
		for (int j = 0; j < nj; ++j) {
			double k = knl_value(order, dim, dist(xi, qj + (size_t)j * dim, dim) / w);
			const double *aj = a + (size_t)j * dim;
			for (int c = 0; c < dim; ++c) vi[c] += k * aj[c];
		}
	}
	return 0;
}

/* dq_i = sum_j (l_i.r_j + l_j.r_i) c(d_ij) (q_i - q_j), the gradient of
 * l^T K(q) r with respect to q (Cpp/src/dqKernel.cpp:31-305).             */
int orc_dq_kernel(double *dq, const double *q, const double *l, const double *r,
                  int order, double w, int n, int dim)
{
	if (!order_ok(order) || (dim != 2 && dim != 3)) return ORC_ERR_ARG;
	double w2 = w * w;
	memset(dq, 0, sizeof(double) * (size_t)dim * n);
	if (order == 0) return 0;
	for (int i = 0; i < n; ++i) {
		const double *xi = q + (size_t)i * dim, *li = l + (size_t)i * dim, *ri = r + (size_t)i * dim;
		double *gi = dq + (size_t)i * dim;
		for (int j = 0; j < n; ++j) {
			const double *xj = q + (size_t)j * dim, *lj = l + (size_t)j * dim, *rj = r + (size_t)j * dim;
			double c = knl_deriv(order, dim, dist(xi, xj, dim) / w, w2);
			double lr = dot(li, rj, dim) + dot(lj, ri, dim);
			for (int k = 0; k < dim; ++k) gi[k] += lr * c * (xi[k] - xj[k]);
		}
	}
	return 0;
}

/* cross form: gradients of sum_ij l_i^T k(qi,qj) r_j with respect to qi and
 * qj (Cpp/src/dqKernel.cpp:307-591).                                      */
int orc_dq_kernel_cross(double *dqi, double *dqj, const double *qi, const double *qj,
                        const double *l, const double *r, int order, double w,
                        int ni, int nj, int dim)
{
	if (!order_ok(order) || (dim != 2 && dim != 3)) return ORC_ERR_ARG;
	double w2 = w * w;
	memset(dqi, 0, sizeof(double) * (size_t)dim * ni);
	memset(dqj, 0, sizeof(double) * (size_t)dim * nj);
	if (order == 0) return 0;
	for (int i = 0; i < ni; ++i) {
		const double *xi = qi + (size_t)i * dim, *li = l + (size_t)i * dim;
		double *gi = dqi + (size_t)i * dim;
		for (int j = 0; j < nj; ++j) {
			const double *xj = qj + (size_t)j * dim, *rj = r + (size_t)j * dim;
			double *gj = dqj + (size_t)j * dim;
			double c = knl_deriv(order, dim, dist(xi, xj, dim) / w, w2);
			double lr = dot(li, rj, dim);
			for (int k = 0; k < dim; ++k) {
				gi[k] += lr * c * (xi[k] - xj[k]);
				gj[k] += lr * c * (xj[k] - xi[k]);
			}
		}
	}
	return 0;
}

/* ------------------------------------------------------------------------ */
/* Varifold                                                                  */
/* ------------------------------------------------------------------------ */

static int cen_ok(char t) { return t == 'G' || t == 'C'; }
static int dir_ok(char t) { return t == 'B' || t == 'L' || t == 'O' || t == 'U'; }

/* Element centre, unit direction and volume.
 * 2D: c = (q0+q1)/2, a = |q1-q0|, u = (q1-q0)/a  (Cpp/src/varifold2D.cpp:8-32)
 * 3D: c = (q0+q1+q2)/3, n = (q1-q0)x(q2-q0), a = |n|/2, u = n/|n|
 *     (Cpp/src/varifold3D.cpp:8-39)                                        */
void orc_element_geometry(double *cen, double *dir, double *vol, const double *q,
                          const int *elmVtx, int ne, int dim)
{
	for (int e = 0; e < ne; ++e) {
		if (dim == 2) {
			const double *q0 = q + 2 * (size_t)elmVtx[2 * e], *q1 = q + 2 * (size_t)elmVtx[2 * e + 1];
			double *c = cen + 2 * (size_t)e, *u = dir + 2 * (size_t)e;
			c[0] = 0.5 * (q0[0] + q1[0]);
			c[1] = 0.5 * (q0[1] + q1[1]);
			u[0] = q1[0] - q0[0];
			u[1] = q1[1] - q0[1];
			double a = sqrt(u[0] * u[0] + u[1] * u[1]);
			u[0] /= a;
			u[1] /= a;
			vol[e] = a;
		} else {
			const double *q0 = q + 3 * (size_t)elmVtx[3 * e], *q1 = q + 3 * (size_t)elmVtx[3 * e + 1],
			             *q2 = q + 3 * (size_t)elmVtx[3 * e + 2];
			double *c = cen + 3 * (size_t)e, *u = dir + 3 * (size_t)e;
			double e1[3], e2[3];
			for (int k = 0; k < 3; ++k) {
				c[k] = (q0[k] + q1[k] + q2[k]) / 3.0;
				e1[k] = q1[k] - q0[k];
				e2[k] = q2[k] - q0[k];
			}
			u[0] = e1[1] * e2[2] - e1[2] * e2[1];
			u[1] = -e1[0] * e2[2] + e1[2] * e2[0];
			u[2] = e1[0] * e2[1] - e1[1] * e2[0];
			double nn = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
			u[0] /= nn;
			u[1] /= nn;
			u[2] /= nn;
			vol[e] = 0.5 * nn;
		}
	}
}

/* Position kernel and its gradient factor: value k, and g with
 * grad_{c1} k = g (c1 - c2).   'G' exp(-r^2/s^2), 'C' 1/(1+r^2/s^2)
 * (Cpp/src/varifold3D.cpp:41-93).                                         */
static double cen_kernel(char t, double r2, double s2, double *g)
{
	if (t == 'G') {
		double k = exp(-r2 / s2);
		if (g) *g = -2.0 * k / s2;
		return k;
	}
	double k = 1.0 / (1.0 + r2 / s2);
	if (g) *g = -2.0 * k * k / s2;
	return k;
}

/* Orientation kernel value f(ang) and df/dang; ang = u1.u2.
 * 'B' ang^2, 'L' ang, 'O' exp(2(ang-1)/s^2), 'U' exp(2(ang^2-1)/s^2)
 * (Cpp/src/varifold3D.cpp:95-188).                                        */
static double dir_kernel(char t, double ang, double s2, double *df)
{
	double f;
	switch (t) {
	case 'B': f = ang * ang; if (df) *df = 2.0 * ang; return f;
	case 'L': if (df) *df = 1.0; return ang;
	case 'O': f = exp(2.0 * (ang - 1.0) / s2); if (df) *df = 2.0 * f / s2; return f;
	default:  f = exp(2.0 * (ang * ang - 1.0) / s2); if (df) *df = 4.0 * ang * f / s2; return f;
	}
}

/* Sum over all (i, j) pairs of k_c k_u a_i a_j for two element sets. */
static double vf_pair_sum(const double *ci, const double *ui, const double *ai, int ni,
                          const double *cj, const double *uj, const double *aj, int nj,
                          char ct, double cw, char dt, double dw, int dim)
{
	double s = 0.0, cs2 = cw * cw, ds2 = dw * dw;
	for (int i = 0; i < ni; ++i)
		for (int j = 0; j < nj; ++j) {
			double r = dist(ci + (size_t)i * dim, cj + (size_t)j * dim, dim);
			double kc = cen_kernel(ct, r * r, cs2, NULL);
			double ku = dir_kernel(dt, dot(ui + (size_t)i * dim, uj + (size_t)j * dim, dim), ds2, NULL);
			s += kc * ku * ai[i] * aj[j];
		}
	return s;
}

int orc_varifold_terms(double *dd, double *dt, double *tt, const double *q, const int *elmVtx,
                       const double *tgtCen, const double *tgtDir, const double *tgtVol,
                       char cenType, double cenWidth, char dirType, double dirWidth,
                       int ne, int nt, int dim)
{
	if (!cen_ok(cenType) || !dir_ok(dirType) || (dim != 2 && dim != 3)) return ORC_ERR_ARG;
	double *cen = malloc(sizeof(double) * (size_t)dim * ne);
	double *dir = malloc(sizeof(double) * (size_t)dim * ne);
	double *vol = malloc(sizeof(double) * (size_t)ne);
	orc_element_geometry(cen, dir, vol, q, elmVtx, ne, dim);
	if (dd) *dd = vf_pair_sum(cen, dir, vol, ne, cen, dir, vol, ne, cenType, cenWidth, dirType, dirWidth, dim);
	if (dt) *dt = vf_pair_sum(cen, dir, vol, ne, tgtCen, tgtDir, tgtVol, nt, cenType, cenWidth, dirType, dirWidth, dim);
	if (tt) *tt = vf_pair_sum(tgtCen, tgtDir, tgtVol, nt, tgtCen, tgtDir, tgtVol, nt, cenType, cenWidth, dirType, dirWidth, dim);
	free(cen); free(dir); free(vol);
	return 0;
}

/* Varifold distance V = DD - 2 DT + TT and, if dq != NULL, its gradient
 * with respect to the vertices (Cpp/src/varifold3D.cpp:259-435,
 * Cpp/src/varifold2D.cpp:126-284).  The per-element gradient with respect
 * to the centre (dc) and to the unnormalised direction (dd) is accumulated
 * element by element and then gathered onto vertices through the adjacency
 * table elmIfo (Cpp/src/varifold3D.cpp:359-391, Cpp/src/varifold2D.cpp:218-240).
 * The accumulation order (deformed-deformed then deformed-target per
 * element i, then target-target) follows the reference.                   */
int orc_varifold(double *val, double *dq, const double *q, const int *elmVtx, const int *elmIfo,
                 const double *tgtCen, const double *tgtDir, const double *tgtVol,
                 char cenType, double cenWidth, char dirType, double dirWidth,
                 int nq, int ne, int ifoRows, int nt, int dim)
{
	if (!cen_ok(cenType) || !dir_ok(dirType) || (dim != 2 && dim != 3)) return ORC_ERR_ARG;
	double cs2 = cenWidth * cenWidth, ds2 = dirWidth * dirWidth;
	double *cen = malloc(sizeof(double) * (size_t)dim * ne);
	double *dir = malloc(sizeof(double) * (size_t)dim * ne);
	double *vol = malloc(sizeof(double) * (size_t)ne);
	double *dc = calloc((size_t)dim * ne, sizeof(double));
	double *ddm = calloc((size_t)dim * ne, sizeof(double));
	orc_element_geometry(cen, dir, vol, q, elmVtx, ne, dim);

	/* In 3D the unnormalised direction is n with |n| = 2a, in 2D it is the
	 * edge vector with |e| = a: the volume derivative d a / d n is u/2 in 3D
	 * and u in 2D (Cpp/src/varifold3D.cpp:287,295; varifold2D.cpp:154,161). */
	double nlen_per_vol = (dim == 3) ? 2.0 : 1.0;
	double dvol_dn = (dim == 3) ? 0.5 : 1.0;

	double v = 0.0;
	for (int i = 0; i < ne; ++i) {
		const double *ci = cen + (size_t)i * dim, *ui = dir + (size_t)i * dim;
		double ai = vol[i], nlen = nlen_per_vol * ai;
		double *gci = dc + (size_t)i * dim, *gdi = ddm + (size_t)i * dim;
		for (int pass = 0; pass < 2; ++pass) {
			const double *cj0 = pass == 0 ? cen : tgtCen;
			const double *uj0 = pass == 0 ? dir : tgtDir;
			const double *aj0 = pass == 0 ? vol : tgtVol;
			int nj = pass == 0 ? ne : nt;
			double sgn = pass == 0 ? 1.0 : -1.0;
			for (int j = 0; j < nj; ++j) {
				const double *cj = cj0 + (size_t)j * dim, *uj = uj0 + (size_t)j * dim;
				double aj = aj0[j];
				double r = dist(ci, cj, dim);
				double gk;
				double kc = cen_kernel(cenType, r * r, cs2, &gk);
				double ang = dot(ui, uj, dim), df;
				double ku = dir_kernel(dirType, ang, ds2, &df);
				double w = (pass == 0 ? 1.0 : 2.0);
				v += sgn * w * kc * ku * ai * aj;
				for (int k = 0; k < dim; ++k) {
					double dci = gk * (ci[k] - cj[k]);
					double dui = df / nlen * (-ang * ui[k] + uj[k]);
					gci[k] += sgn * 2.0 * dci * ku * ai * aj;
					gdi[k] += sgn * 2.0 * kc * (dui * ai + ku * dvol_dn * ui[k]) * aj;
				}
			}
		}
	}
	v += vf_pair_sum(tgtCen, tgtDir, tgtVol, nt, tgtCen, tgtDir, tgtVol, nt,
	                 cenType, cenWidth, dirType, dirWidth, dim);
	*val = v;

	if (dq) {
		memset(dq, 0, sizeof(double) * (size_t)dim * nq);
		for (int n = 0; n < nq; ++n) {
			const int *col = elmIfo + (size_t)n * ifoRows;
			int adj = col[0];
			double *g = dq + (size_t)n * dim;
			for (int a = 0; a < adj; ++a) {
				if (dim == 2) {
					int e = col[1 + 2 * a], s = col[1 + 2 * a + 1];
					g[0] += 0.5 * dc[2 * (size_t)e] + s * ddm[2 * (size_t)e];
					g[1] += 0.5 * dc[2 * (size_t)e + 1] + s * ddm[2 * (size_t)e + 1];
				} else {
					int e = col[1 + 3 * a], ia = col[1 + 3 * a + 1], ib = col[1 + 3 * a + 2];
					const double *qa = q + 3 * (size_t)ia, *qb = q + 3 * (size_t)ib;
					const double *d = ddm + 3 * (size_t)e, *c = dc + 3 * (size_t)e;
					double ab[3] = {qa[0] - qb[0], qa[1] - qb[1], qa[2] - qb[2]};
					double x[3];
					x[0] = d[1] * ab[2] - d[2] * ab[1];
					x[1] = -d[0] * ab[2] + d[2] * ab[0];
					x[2] = d[0] * ab[1] - d[1] * ab[0];
					for (int k = 0; k < 3; ++k) g[k] += c[k] / 3.0 + x[k];
				}
			}
		}
	}
	free(cen); free(dir); free(vol); free(dc); free(ddm);
	return 0;
}

/* Adjacency table (Cpp/src/getElmIfo.m:1-59) from a 0-based V x E table.
 * Column n: [count, then per adjacent element in ascending element order
 * 2D: elm, sign (-1 for the element's first vertex, +1 otherwise);
 * 3D: elm, qa, qb with (qa, qb) = (v2, v1), (v0, v2), (v1, v0) for local
 * vertex 0, 1, 2].  Returns the row count 1 + maxAdj*V; when ifo is NULL
 * only the row count is computed.                                          */
int orc_get_elm_ifo(int *ifo, const int *elmVtx, int vpe, int ne, int nq)
{
	if (vpe != 2 && vpe != 3) return ORC_ERR_ARG;
	int *cnt = calloc((size_t)nq, sizeof(int));
	for (int e = 0; e < ne; ++e)
		for (int k = 0; k < vpe; ++k) {
			int v = elmVtx[(size_t)e * vpe + k];
			if (v < 0 || v >= nq) { free(cnt); return ORC_ERR_ARG; }
			cnt[v]++;
		}
	int maxAdj = 0;
	for (int n = 0; n < nq; ++n) if (cnt[n] > maxAdj) maxAdj = cnt[n];
	int rows = 1 + maxAdj * vpe;
	if (ifo) {
		memset(ifo, 0, sizeof(int) * (size_t)rows * nq);
		for (int e = 0; e < ne; ++e) {
			const int *ev = elmVtx + (size_t)e * vpe;
			for (int k = 0; k < vpe; ++k) {
				int *col = ifo + (size_t)ev[k] * rows;
				int a = col[0]++;
				int *slot = col + 1 + (size_t)a * vpe;
				slot[0] = e;
				if (vpe == 2) slot[1] = (k == 0) ? -1 : 1;
				else if (k == 0) { slot[1] = ev[2]; slot[2] = ev[1]; }
				else if (k == 1) { slot[1] = ev[0]; slot[2] = ev[2]; }
				else { slot[1] = ev[1]; slot[2] = ev[0]; }
			}
		}
	}
	free(cnt);
	return rows;
}

/* ------------------------------------------------------------------------ */
/* Objective and gradient                                                    */
/* ------------------------------------------------------------------------ */

/* E(alpha) = (h/2) sum_t <alpha_t, K(q_t) alpha_t> + w V(q_{T-1}) with the
 * forward Euler flow q_{t+1} = q_t + h K(q_t) alpha_t, and its gradient by
 * the discrete adjoint (Cpp/src/objgrd.cpp:26-90):
 *   p = -w grad V(q_{T-1});  for t = T-2 .. 0:
 *     g_t = h K(q_t) (alpha_t - p);
 *     p  -= h dq[(alpha_t/2 - p)^T K(q_t) alpha_t].
 * Matrix-free: the kernel matrix is recomputed per product instead of the
 * reference's stored knlStk (Cpp/src/assignStructMemory.cpp:89-91); the
 * arithmetic of each product is the same.  ndeStk (D*N*T) may be NULL.    */
void orc_l2_dist(double *fcn, double *grd, const double *q, const double *tgt, double w, int n, int dim)
{
	double s = 0.0;
	for (size_t k = 0; k < (size_t)n * dim; ++k) {
		double d = q[k] - tgt[k];
		s += d * d;
		grd[k] = w * d;
	}
	*fcn = w * 0.5 * s;
}

int orc_objgrd(double *fcn, double *grd, const double *alp, const orc_problem *P, double *ndeStkOut)
{
	int n = P->ndeNum, T = P->timeNum, dim = P->dim;
	size_t dn = (size_t)dim * n;
	double h = P->timeStp;
	if (!order_ok(P->knlOrder) || T < 2) return ORC_ERR_ARG;
	double *nde = ndeStkOut ? ndeStkOut : malloc(sizeof(double) * dn * T);
	double *vlc = malloc(sizeof(double) * dn);
	double *pm = malloc(sizeof(double) * dn);
	double *amp = malloc(sizeof(double) * dn);
	double *pdot = malloc(sizeof(double) * dn);
	double *dqv = malloc(sizeof(double) * dn);
	int rc = 0;

	memcpy(nde, P->iniNdeMat, sizeof(double) * dn);
	double e = 0.0;
	for (int t = 0; t < T - 1; ++t) {
		const double *qt = nde + (size_t)t * dn, *at = alp + (size_t)t * dn;
		double *qn = nde + (size_t)(t + 1) * dn;
		orc_multiply_kernel(vlc, qt, at, P->knlOrder, P->knlWidth, n, dim);
		double s = 0.0;
		for (size_t k = 0; k < dn; ++k) s += at[k] * vlc[k];
		e += s;
		for (size_t k = 0; k < dn; ++k) qn[k] = qt[k] + h * vlc[k];
	}
	if (P->tgtNde) {
		/* MATLAB/src/objgrd.m:29-34 with disHdl = l2Dist: the handle's value and
		 * gradient already carry the weight */
		double l2;
		orc_l2_dist(&l2, dqv, nde + (size_t)(T - 1) * dn, P->tgtNde, P->disWgt, n, dim);
		*fcn = h * 0.5 * e + l2;
		for (size_t k = 0; k < dn; ++k) pm[k] = -dqv[k];
	} else {
		double vfd;
		rc = orc_varifold(&vfd, dqv, nde + (size_t)(T - 1) * dn, P->elmVtx, P->elmIfo, P->tgtCen, P->tgtDir,
		                  P->tgtVol, P->cenType, P->cenWidth, P->dirType, P->dirWidth, n, P->elmNum,
		                  P->ifoRows, P->tgtNum, dim);
		if (rc) goto done;
		*fcn = h * 0.5 * e + P->disWgt * vfd;
		for (size_t k = 0; k < dn; ++k) pm[k] = -P->disWgt * dqv[k];
	}
	for (int t = T - 2; t >= 0; --t) {
		const double *qt = nde + (size_t)t * dn, *at = alp + (size_t)t * dn;
		double *gt = grd + (size_t)t * dn;
		for (size_t k = 0; k < dn; ++k) amp[k] = at[k] - pm[k];
		orc_multiply_kernel(gt, qt, amp, P->knlOrder, P->knlWidth, n, dim);
		for (size_t k = 0; k < dn; ++k) amp[k] = amp[k] - 0.5 * at[k];
		orc_dq_kernel(pdot, qt, amp, at, P->knlOrder, P->knlWidth, n, dim);
		for (size_t k = 0; k < dn; ++k) pm[k] = pm[k] - h * pdot[k];
	}
	for (size_t k = 0; k < dn * (T - 1); ++k) grd[k] = h * grd[k];

done:
	if (!ndeStkOut) free(nde);
	free(vlc); free(pm); free(amp); free(pdot); free(dqv);
	return rc;
}

/* ------------------------------------------------------------------------ */
/* L-BFGS                                                                    */
/* ------------------------------------------------------------------------ */

static double vdot(const double *a, const double *b, size_t n)
{
	double s = 0.0;
	for (size_t k = 0; k < n; ++k) s += a[k] * b[k];
	return s;
}

/* Two-loop recursion, Nocedal & Wright Algorithm 7.4
 * (Cpp/src/getDirection.cpp:9-66): d = -H g from the hisNum newest pairs of
 * the ring buffers S, Y (newest slot newIdx), initial scaling H0.         */
void orc_get_direction(double *d, double H0, const double *g, const double *S, const double *Y,
                       int newIdx, int hisNum, int memNum, double *rec, size_t n)
{
	memcpy(d, g, sizeof(double) * n);
	for (int h = 0; h < hisNum; ++h) {
		int idx = newIdx - h;
		if (idx < 0) idx += memNum;
		const double *s = S + (size_t)idx * n, *y = Y + (size_t)idx * n;
		double a = vdot(s, d, n) / vdot(s, y, n);
		for (size_t k = 0; k < n; ++k) d[k] -= a * y[k];
		rec[idx] = a;
	}
	for (size_t k = 0; k < n; ++k) d[k] *= H0;
	int old = (hisNum < memNum) ? 0 : newIdx + 1;
	if (old == memNum) old = 0;
	for (int h = 0; h < hisNum; ++h) {
		int idx = old + h;
		if (idx >= memNum) idx -= memNum;
		const double *s = S + (size_t)idx * n, *y = Y + (size_t)idx * n;
		double b = vdot(y, d, n) / vdot(s, y, n);
		for (size_t k = 0; k < n; ++k) d[k] += (rec[idx] - b) * s[k];
	}
	for (size_t k = 0; k < n; ++k) d[k] = -d[k];
}

/* Weak-Wolfe bracketing/bisection line search (Cpp/src/lineSearch.cpp:26-110).
 * Returns 0 on success, 1 not a descent direction, 2 no bracket found,
 * 3 bracketed but not satisfied.                                           */
int orc_line_search(orc_objfun f, void *ctx, const double *x, const double *g, const double *d,
                    double fnow, double c1, double c2, double tol, double *xt, double *gt,
                    double *ft, double *step, int *cnt, size_t n, int quiet)
{
	double lo = 0.0, hi = DBL_MAX, rel = 1.0;
	double slope = vdot(g, d, n);
	*cnt = 0;
	if (slope > -DBL_MIN) {
		if (!quiet) printf("\nNot a descent direction.\nQuit LBFGS.\n\n");
		return 1;
	}
	double dlen = sqrt(vdot(d, d, n));
	int bisMax = (int)fmax(50.0, log2(1e5 * dlen));
	int epdMax = (int)fmax(10.0, log2(1e5 / dlen));
	int bis = 0, epd = 0;
	while (bis <= bisMax && epd <= epdMax) {
		for (size_t k = 0; k < n; ++k) xt[k] = x[k] + rel * d[k];
		f(ctx, ft, gt, xt);
		++*cnt;
		if (sqrt(vdot(gt, gt, n)) < tol) { *step = rel * dlen; return 0; }
		double st = vdot(gt, d, n);
		if (*ft >= fnow + c1 * rel * slope) hi = rel;
		else if (st <= c2 * slope) lo = rel;
		else { *step = rel * dlen; return 0; }
		if (hi == DBL_MAX) { rel *= 2.0; ++epd; }
		else { rel = 0.5 * (lo + hi); ++bis; }
	}
	if (hi == DBL_MAX) {
		if (!quiet) printf("\nLine search failed to bracket a point satisfying weak Wolfe conditions.\n"
		                   "Function may be unbounded below.\nQuit LBFGS.\n\n");
		return 2;
	}
	if (!quiet) printf("\nLine search failed to satisfy weak Wolfe conditions, although\n"
	                   "a point satisfying conditions was bracketed in [%e, %e].\nQuit LBFGS.\n\n", lo, hi);
	return 3;
}

/* L-BFGS driver (Cpp/src/LBFGS.cpp:25-112).  x holds the start on entry and
 * the last accepted iterate on return.  The optional trace arrays hold
 * itrMax+1 entries (iteration 0 = the start).  Returns 0, or the
 * line-search code that ended the run.                                    */
int orc_lbfgs(orc_objfun f, void *ctx, double *fx, double *gx, double *x, size_t n,
              const orc_lbfgs_opts *o, orc_trace *tr)
{
	int m = o->memNum;
	double *S = calloc(n * (size_t)m, sizeof(double));
	double *Y = calloc(n * (size_t)m, sizeof(double));
	double *d = malloc(sizeof(double) * n), *xn = malloc(sizeof(double) * n), *gn = malloc(sizeof(double) * n);
	double *rec = calloc((size_t)m, sizeof(double));
	double fn;
	int status = 0, evals = 1, iters = 0;

	f(ctx, fx, gx, x);
	double glen = sqrt(vdot(gx, gx, n));
	if (o->vbsFlg) {
		printf("%5s   %13s  %13s  %13s  %9s\n", "iter", "f", "|grad f|", "step length", "fcn eval");
		char sep[65] = {0};
		memset(sep, '-', 62);
		printf("%s\n", sep);
		printf("%5d:  %13.6e  %13.6e\n", 0, *fx, glen);
	}
	if (tr) { tr->f[0] = *fx; tr->gnorm[0] = glen; tr->step[0] = 0.0; tr->evals[0] = 1; }

	int newIdx = -1;
	for (int it = 1; it <= o->itrMax; ++it) {
		if (glen < o->tolVal) break;
		double H0 = 1.0;
		if (newIdx != -1) {
			const double *s = S + (size_t)newIdx * n, *y = Y + (size_t)newIdx * n;
			H0 = vdot(s, y, n) / vdot(y, y, n);
		}
		orc_get_direction(d, H0, gx, S, Y, newIdx, it <= m ? it - 1 : m, m, rec, n);
		double step;
		int cnt;
		int le = orc_line_search(f, ctx, x, gx, d, *fx, o->wolfe1, o->wolfe2, o->tolVal, xn, gn, &fn,
		                         &step, &cnt, n, 0);
		evals += cnt;
		if (le != 0) { status = le; break; }
		newIdx = (newIdx == m - 1) ? 0 : newIdx + 1;
		double *s = S + (size_t)newIdx * n, *y = Y + (size_t)newIdx * n;
		for (size_t k = 0; k < n; ++k) { s[k] = xn[k] - x[k]; y[k] = gn[k] - gx[k]; }
		memcpy(x, xn, sizeof(double) * n);
		memcpy(gx, gn, sizeof(double) * n);
		*fx = fn;
		glen = sqrt(vdot(gx, gx, n));
		iters = it;
		if (o->vbsFlg) printf("%5d:  %13.6e  %13.6e  %13.6e  %9d\n", it, *fx, glen, step, cnt);
		if (tr) { tr->f[it] = *fx; tr->gnorm[it] = glen; tr->step[it] = step; tr->evals[it] = cnt; }
	}
	if (tr) { tr->iters = iters; tr->status = status; tr->totalEvals = evals; }
	free(S); free(Y); free(d); free(xn); free(gn); free(rec);
	return status;
}

static void objgrd_cb(void *ctx, double *f, double *g, const double *x)
{
	orc_objgrd(f, g, x, (const orc_problem *)ctx, NULL);
}

/* Full registration with the argument order of the reference's LDDMM2D /
 * LDDMM3D MEX entry (Cpp/src/LDDMM.cpp:16-37,87-88): optAlpVec receives
 * the last accepted iterate.                                              */
int orc_lddmm_run(const orc_problem *P, const double *iniAlpVec, const orc_lbfgs_opts *o,
                  double *optAlpVec, double *fcnOut, orc_trace *tr)
{
	size_t n = (size_t)P->dim * P->ndeNum * (P->timeNum - 1);
	double *g = malloc(sizeof(double) * n);
	double f;
	memcpy(optAlpVec, iniAlpVec, sizeof(double) * n);
	int st = orc_lbfgs(objgrd_cb, (void *)P, &f, g, optAlpVec, n, o, tr);
	if (fcnOut) *fcnOut = f;
	free(g);
	return st;
}
