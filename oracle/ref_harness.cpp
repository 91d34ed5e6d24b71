// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C-linkage wrappers around the reference's own kernel-level C++ functions
// (dnhsieh/LDDMM Cpp/src: computeKernel.cpp, multiplyKernel.cpp,
// dqKernel.cpp, polybesselk.cpp, varifold2D.cpp / varifold3D.cpp), which
// oracle/Makefile compiles where they lie under the reference tree into
// oracle/_ref/libref{2,3}d.so.  Nothing of the reference is copied here:
// these are prototypes of its functions and thin adapters that own the
// scratch arrays its varifold() expects (Cpp/src/varifoldITF.cpp:7-19).
//
// Only Cpp/src files that need nothing beyond the reference's own headers
// are built; objgrd.cpp, LBFGS.cpp, getDirection.cpp and lineSearch.cpp
// include MATLAB's blas.h and LDDMM.cpp includes mex.h, neither of which is
// in this image, so those are not built.
#include <cstdlib>
#include <vector>

void computeKernel(double *, double *, int, double, int, int);
void computeKernel(double *, double *, double *, int, double, int, int, int);
void multiplyKernel(double *, double *, double *, int, double, int, int);
void multiplyKernel(double *, double *, double *, double *, int, double, int, int, int);
void dqKernel(double *, double *, double *, double *, int, double, int, int);
void dqKernel(double *, double *, double *, double *, double *, double *, int, double, int, int, int);
void p0Fcn(double &, double);
void p1Fcn(double &, double);
void varifold(double *, double *, double *, int *, int *, double *, double *, double *,
              char, double, char, double, double *, double *, double *, double *, double *,
              int, int, int, int);
void varifold(double *, double *, int *, double *, double *, double *,
              char, double, char, double, double *, double *, double *, int, int, int);

extern "C" {

void ref_p01(double x, double *p0, double *p1)
{
	p0Fcn(*p0, x);
	p1Fcn(*p1, x);
}

void ref_compute_kernel(double *K, double *q, int order, double w, int n, int dim)
{
	computeKernel(K, q, order, w, n, dim);
}

void ref_compute_kernel_cross(double *K, double *qi, double *qj, int order, double w, int ni, int nj, int dim)
{
	computeKernel(K, qi, qj, order, w, ni, nj, dim);
}

void ref_multiply_kernel(double *v, double *q, double *a, int order, double w, int n, int dim)
{
	multiplyKernel(v, q, a, order, w, n, dim);
}

void ref_multiply_kernel_cross(double *v, double *qi, double *qj, double *a, int order, double w,
                               int ni, int nj, int dim)
{
	multiplyKernel(v, qi, qj, a, order, w, ni, nj, dim);
}

void ref_dq_kernel(double *dq, double *q, double *l, double *r, int order, double w, int n, int dim)
{
	dqKernel(dq, q, l, r, order, w, n, dim);
}

void ref_dq_kernel_cross(double *dqi, double *dqj, double *qi, double *qj, double *l, double *r,
                         int order, double w, int ni, int nj, int dim)
{
	dqKernel(dqi, dqj, qi, qj, l, r, order, w, ni, nj, dim);
}

// dim is fixed by the build (-DDIM2 / -DDIM3); the caller passes it so a
// mismatch can be rejected.
int ref_varifold(double *val, double *dq, double *q, int *elmVtx, int *elmIfo,
                 double *tgtCen, double *tgtDir, double *tgtVol,
                 char cenType, double cenWidth, char dirType, double dirWidth,
                 int nq, int ne, int ifoRows, int nt, int dim)
{
#ifdef DIM2
	if (dim != 2) return -1;
#else
	if (dim != 3) return -1;
#endif
	std::vector<double> cen((size_t)dim * ne), dir((size_t)dim * ne), vol(ne);
	std::vector<double> dc((size_t)dim * ne), dd((size_t)dim * ne);
	if (dq)
		varifold(val, dq, q, elmVtx, elmIfo, tgtCen, tgtDir, tgtVol, cenType, cenWidth, dirType, dirWidth,
		         cen.data(), dir.data(), vol.data(), dc.data(), dd.data(), nq, ne, ifoRows, nt);
	else
		varifold(val, q, elmVtx, tgtCen, tgtDir, tgtVol, cenType, cenWidth, dirType, dirWidth,
		         cen.data(), dir.data(), vol.data(), nq, ne, nt);
	return 0;
}

}
