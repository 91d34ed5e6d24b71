/*
 * lddmm.h -- C-ABI of the MI355X (gfx950) LDDMM landmark/varifold
 * registration library (liblddmm_amd.so).
 *
 * Plain C: no HIP, torch or C++ types in any signature.  Every entry point
 * returns an int status: 0 = ok; 1, 2, 3 = the reference's line-search
 * codes (not a descent direction / no bracket / bracketed but weak Wolfe
 * not met); negative = LDDMM_ERR_*.  The last error message of the calling
 * thread is available from lddmm_last_error().
 *
 * Layouts.  "Host layout" is the reference's Cpp MEX layout
 * (Cpp/src/LDDMM.m:7-14): a point set is D x N column-major (AoS
 * x0 y0 [z0] x1 ...), a momentum stack is D x N x (T-1), element tables are
 * int32 V x E 0-based, the adjacency table (1 + maxAdj*V) x N.  "Device
 * layout" is the reference's CUDA MEX layout (CUDA/src/LDDMM.m:7-19): N x D
 * column-major (SoA), stacks N x D x (T-1), elements E x V, adjacency
 * N x (1 + maxAdj*V).
 *
 * Kernel orders: -1 Gaussian, 0..4 Matern (Cpp/src/computeKernel.cpp:23-24),
 * in 2D and 3D (2D Matern through the Bessel factors x^2 K0(x), x K1(x)).
 */
#ifndef LDDMM_AMD_H
#define LDDMM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDDMM_OK 0
#define LDDMM_ERR_ARG (-1)
#define LDDMM_ERR_ALLOC (-2)
#define LDDMM_ERR_HIP (-3)
#define LDDMM_ERR_UNSUPPORTED (-4)

#define LDDMM_PRECISION_FP64 0   /* FP64 throughout: the reference's arithmetic */
#define LDDMM_PRECISION_FP32 1   /* FP32 pair terms with FP64 coordinate differences, FP64
                                    cross-tile sums, energies, reductions and L-BFGS state */

/* Options beyond the reference's 22 arguments; NULL = defaults. */
typedef struct lddmm_options {
    int precision;            /* LDDMM_PRECISION_FP64 (default) or LDDMM_PRECISION_FP32 */
    int no_fp64_refine;       /* FP32 runs: 0 (default) = once the FP32 phase stalls (line-search
                                 failure or relative decrease below refine_rel_tol), re-evaluate in
                                 FP64 and continue in FP64 with the L-BFGS memory kept; 1 = stop */
    double rel_energy_tol;    /* > 0: also stop once (f_prev - f)/|f| < rel_energy_tol */
    double refine_rel_tol;    /* stall threshold of the FP32 phase; 0 = 1e-7 */
    double target_energy;     /* used when has_target: stop once f <= target_energy */
    int has_target;
    int pair_target_blocks;   /* workgroups per all-pairs launch; 0 = default */
    int progress;             /* batch runs: > 0 prints a progress line to stderr every
                                 `progress` rounds; 0 = silent */
    int reserved[3];
} lddmm_options;

/* Run summary; may be NULL. */
typedef struct lddmm_result {
    double fcn;               /* final objective value */
    double grad_norm;         /* |grad f| at the final iterate */
    double seconds;           /* wall time of the optimisation */
    double pairs_per_eval;    /* kernel pair interactions per objective evaluation */
    int iterations;           /* accepted L-BFGS iterations */
    int evaluations;          /* objective evaluations */
    int status;               /* 0 or line-search code 1..3 */
    int fp32_iterations;      /* iterations done with FP32 pair arithmetic */
    int reached_target;       /* 1 if the run stopped on target_energy */
    int reserved[3];
} lddmm_result;

/* ---------------------------------------------------------------- registration
 * Replaces the LDDMM2D / LDDMM3D MEX entry of the CPU variant
 * (Cpp/src/LDDMM.cpp:11-98, caller Cpp/src/LDDMM.m:16-36): the 22 inputs in
 * their order and meaning (prhs[0..21]), then the sizes MATLAB reads from
 * the arrays (mxGetN of prhs[0], prhs[2], prhs[4]; mxGetM of prhs[3]), then
 * the output plhs[0] = optAlpVec (D*N*(T-1), caller-allocated), which always
 * holds the last accepted iterate.  Host layout. */
int lddmm_run_2d(const double *iniNdeMat, const double *iniAlpVec,
                 const int32_t *dfmElmVtxMat, const int32_t *dfmElmIfoMat,
                 const double *tgtCenPosMat, const double *tgtElmVolVec, const double *tgtUniDirMat,
                 char cenKnlType, double cenKnlWidth, char dirKnlType, double dirKnlWidth, double disWgt,
                 int knlOrder, double knlWidth, double timeStp, int timeNum,
                 int itrMax, double tolVal, double wolfe1, double wolfe2, int memNum, int vbsFlg,
                 int ndeNum, int dfmElmNum, int dfmElmIfoNum, int tgtElmNum,
                 double *optAlpVec);
int lddmm_run_3d(const double *iniNdeMat, const double *iniAlpVec,
                 const int32_t *dfmElmVtxMat, const int32_t *dfmElmIfoMat,
                 const double *tgtCenPosMat, const double *tgtElmVolVec, const double *tgtUniDirMat,
                 char cenKnlType, double cenKnlWidth, char dirKnlType, double dirKnlWidth, double disWgt,
                 int knlOrder, double knlWidth, double timeStp, int timeNum,
                 int itrMax, double tolVal, double wolfe1, double wolfe2, int memNum, int vbsFlg,
                 int ndeNum, int dfmElmNum, int dfmElmIfoNum, int tgtElmNum,
                 double *optAlpVec);
/* Same, with options and a run summary; dimNum is 2 or 3. */
int lddmm_run_ex(int dimNum, const double *iniNdeMat, const double *iniAlpVec,
                 const int32_t *dfmElmVtxMat, const int32_t *dfmElmIfoMat,
                 const double *tgtCenPosMat, const double *tgtElmVolVec, const double *tgtUniDirMat,
                 char cenKnlType, double cenKnlWidth, char dirKnlType, double dirKnlWidth, double disWgt,
                 int knlOrder, double knlWidth, double timeStp, int timeNum,
                 int itrMax, double tolVal, double wolfe1, double wolfe2, int memNum, int vbsFlg,
                 int ndeNum, int dfmElmNum, int dfmElmIfoNum, int tgtElmNum,
                 double *optAlpVec, const lddmm_options *opts, lddmm_result *result);

/* Replaces the LDDMM2D / LDDMM3D MEX entry of the CUDA variant
 * (CUDA/src/LDDMM.cu:13-151, caller CUDA/src/LDDMM.m:21-41): every array is
 * a device pointer in the device layout; d_optAlpVec is caller-allocated on
 * the device.  stream is a hipStream_t (NULL = the default stream); the call
 * returns after the run has finished on it. */
int lddmm_run_dev(int dimNum, const double *d_iniNdeMat, const double *d_iniAlpVec,
                  const int32_t *d_dfmElmVtxMat, const int32_t *d_dfmElmIfoMat,
                  const double *d_tgtCenPosMat, const double *d_tgtElmVolVec, const double *d_tgtUniDirMat,
                  char cenKnlType, double cenKnlWidth, char dirKnlType, double dirKnlWidth, double disWgt,
                  int knlOrder, double knlWidth, double timeStp, int timeNum,
                  int itrMax, double tolVal, double wolfe1, double wolfe2, int memNum, int vbsFlg,
                  int ndeNum, int dfmElmNum, int dfmElmIfoNum, int tgtElmNum,
                  double *d_optAlpVec, const lddmm_options *opts, lddmm_result *result, void *stream);

/* ---------------------------------------------------------------- objective
 * The objective and gradient of the CPU variant, void objgrd(double *fcnVal,
 * double *grdStk, double *alpStk, fcndata &) (Cpp/src/objgrd.cpp:26-90),
 * with fcndata built once by lddmm_objective_create (the role of
 * assignFcnStructMemory, Cpp/src/assignStructMemory.cpp:49-109).  Host
 * layout for alpStk / grdStk (D x N x (T-1)) and for every create input. */
typedef struct lddmm_objective lddmm_objective;
int lddmm_objective_create(lddmm_objective **out, int dimNum, const double *iniNdeMat,
                           const int32_t *dfmElmVtxMat, const int32_t *dfmElmIfoMat,
                           const double *tgtCenPosMat, const double *tgtElmVolVec, const double *tgtUniDirMat,
                           char cenKnlType, double cenKnlWidth, char dirKnlType, double dirKnlWidth,
                           double disWgt, int knlOrder, double knlWidth, double timeStp, int timeNum,
                           int ndeNum, int dfmElmNum, int dfmElmIfoNum, int tgtElmNum);
int lddmm_objgrd(lddmm_objective *obj, double *fcnVal, double *grdStk, const double *alpStk);
/* Device-layout form (the CUDA variant's objgrd, CUDA/src/objgrd.cu:13-76,
 * works on device arrays): d_alpStk / d_grdStk are N x D x (T-1) device
 * arrays, d_fcnVal one device double.  Enqueued on the handle's stream and
 * asynchronous; lddmm_objective_synchronize waits for it. */
int lddmm_objgrd_dev(lddmm_objective *obj, double *d_fcnVal, double *d_grdStk, const double *d_alpStk);
int lddmm_objective_synchronize(lddmm_objective *obj);
/* pair interactions per evaluation: (T-1) N^2 forward + (T-1) N^2 fused
 * backward + Ed^2 + Ed Et varifold */
int lddmm_objective_pairs(lddmm_objective *obj, double *pairsPerEval);
/* D x N x T host-layout node trajectory of the last lddmm_objgrd call (ndeStk). */
int lddmm_objective_trajectory(lddmm_objective *obj, double *ndeStk);
/* Pair arithmetic of later evaluations: LDDMM_PRECISION_FP64 (the default)
 * or LDDMM_PRECISION_FP32. */
int lddmm_objective_set_precision(lddmm_objective *obj, int precision);
/* The constant target-target varifold term, cached at create time. */
int lddmm_objective_target_term(lddmm_objective *obj, double *value);
void lddmm_objective_destroy(lddmm_objective *obj);

/* ---------------------------------------------------------------- pair operators
 * multiplyKernel MEX, 4- and 5-argument forms (Cpp/src/multiplyKernelITF.cpp:16-59;
 * Cpp/src/multiplyKernel.cpp:22,294): v_i = sum_j k(q_i, q_j) alpha_j.  Host layout. */
int lddmm_multiply_kernel(double *vlcMat, const double *lmkMat, const double *alpMat,
                          int knlOrder, double knlWidth, int lmkNum, int dimNum);
int lddmm_multiply_kernel_cross(double *vlcMat, const double *lmkiMat, const double *lmkjMat,
                                const double *alpMat, int knlOrder, double knlWidth,
                                int lmkiNum, int lmkjNum, int dimNum);
/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): lddmm_multiply_kernel_cross on device arrays in the CUDA layout
 * (N x D column-major: x[0..N) y[0..N) [z[0..N)]), on `stream` (a hipStream_t; NULL
 * is the null stream), asynchronous.  The per-split partial sums live in a fixed
 * 16 MiB device array per GPU (2^21 doubles) that the library never frees; sizes
 * whose partials exceed it (splits x dimNum x lmkiNum doubles: at most 7.9e5 below
 * 131,072 targets, dimNum x lmkiNum above, so up to 699,050 targets in 3D) return -1
 * with the need in lddmm_last_error().  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams or threads) race.  Results equal
 * lddmm_multiply_kernel_cross's bit for bit. */
int lddmm_multiply_kernel_cross_noalloc(double *d_vlcMat, const double *d_lmkiMat, const double *d_lmkjMat,
                                        const double *d_alpMat, int knlOrder, double knlWidth, int lmkiNum,
                                        int lmkjNum, int dimNum, void *stream);
/* dqKernel (Cpp/src/dqKernel.cpp:31, 307; MATLAB/src/dqKernelITF.cpp:16-72):
 * gradient of lft^T K(q) rgt with respect to q; cross form gives both sides. */
int lddmm_dq_kernel(double *dqKMat, const double *lmkMat, const double *lftMat, const double *rgtMat,
                    int knlOrder, double knlWidth, int lmkNum, int dimNum);
int lddmm_dq_kernel_cross(double *dqiKMat, double *dqjKMat, const double *lmkiMat, const double *lmkjMat,
                          const double *lftMat, const double *rgtMat, int knlOrder, double knlWidth,
                          int lmkiNum, int lmkjNum, int dimNum);
/* computeKernel (Cpp/src/computeKernel.cpp:21, 246; MATLAB/src/computeKernelITF.cpp:15-59):
 * dense matrix, self form knlMat[i*N + j], cross form knlMat[j*Ni + i]. */
int lddmm_compute_kernel(double *knlMat, const double *lmkMat, int knlOrder, double knlWidth,
                         int lmkNum, int dimNum);
int lddmm_compute_kernel_cross(double *knlMat, const double *lmkiMat, const double *lmkjMat,
                               int knlOrder, double knlWidth, int lmkiNum, int lmkjNum, int dimNum);
/* varifold(...) value and (when dqVfdMat != NULL) vertex gradient
 * (Cpp/src/varifold3D.cpp:393-435, Cpp/src/varifold2D.cpp:242-284; called by
 * varifoldITF, Cpp/src/varifoldITF.cpp:7-19).  The reference's workspace
 * arguments are owned by the library.  Host layout. */
int lddmm_varifold(double *vfdVal, double *dqVfdMat, const double *dfmLmkPosMat,
                   const int32_t *dfmElmVtxMat, const int32_t *dfmElmIfoMat,
                   const double *tgtCenPosMat, const double *tgtUniDirMat, const double *tgtElmVolVec,
                   char cenKnlType, double cenKnlWidth, char dirKnlType, double dirKnlWidth,
                   int dfmLmkNum, int dfmElmNum, int dfmElmIfoNum, int tgtElmNum, int dimNum);

/* ---------------------------------------------------------------- l2 landmark matching
 * The MATLAB variant's alternative matching term l2Dist (MATLAB/src/l2Dist.m:1-5,
 * used as disHdl by MATLAB/src/objgrd.m:29-34): E = (h/2) sum_t <alpha_t, K alpha_t>
 * + (w/2) |q_{T-1} - tgt|^2 for landmark targets with known correspondence.
 * Host layout (D x N). */
int lddmm_l2_dist(double *fcnVal, double *grdMat, const double *ndeMat, const double *tgtMat, double wgtVal,
                  int ndeNum, int dimNum);
int lddmm_objective_create_l2(lddmm_objective **out, int dimNum, const double *iniNdeMat, const double *tgtNdeMat,
                              double disWgt, int knlOrder, double knlWidth, double timeStp, int timeNum, int ndeNum);
int lddmm_run_l2(int dimNum, const double *iniNdeMat, const double *iniAlpVec, const double *tgtNdeMat,
                 double disWgt, int knlOrder, double knlWidth, double timeStp, int timeNum, int itrMax, double tolVal,
                 double wolfe1, double wolfe2, int memNum, int vbsFlg, int ndeNum, double *optAlpVec,
                 const lddmm_options *opts, lddmm_result *result);

/* ---------------------------------------------------------------- batches
 * Many independent registrations of uneven sizes in one set of launches
 * (no counterpart in the reference, which registers one pair per MEX call;
 * every problem follows Cpp/src/LBFGS.cpp and Cpp/src/objgrd.cpp).  Inputs
 * are the host-layout arrays of each problem concatenated problem after
 * problem: iniNdeMats D x N_p blocks, dfmElmVtxMats V x E_p blocks (0-based,
 * local to the problem), targets D x Et_p / Et_p blocks, knlWidths one per
 * problem; kernel and varifold types, weight and time grid are shared.  The
 * adjacency tables are built by the library.  chunk = source points per
 * work item (0 = autotuned on the batch). */
typedef struct lddmm_batch lddmm_batch;
int lddmm_batch_create(lddmm_batch **out, int dimNum, int numProblems, const int32_t *ndeNums,
                       const int32_t *dfmElmNums, const int32_t *tgtElmNums, const double *iniNdeMats,
                       const int32_t *dfmElmVtxMats, const double *tgtCenPosMats, const double *tgtElmVolVecs,
                       const double *tgtUniDirMats, const double *knlWidths, char cenKnlType, double cenKnlWidth,
                       char dirKnlType, double dirKnlWidth, double disWgt, int knlOrder, double timeStp,
                       int timeNum, int precision, int chunk);
/* fcnVals[p] and the concatenated D x N_p x (T-1) gradient blocks at the
 * concatenated momenta (host layout), in the batch's precision. */
int lddmm_batch_objgrd(lddmm_batch *b, double *fcnVals, double *grdStks, const double *alpStks);
/* Runs every registration (L-BFGS per problem, one batched evaluation per
 * round of line-search trials).  iniAlpStks may be NULL (zeros);
 * targetEnergies (one per problem) may be NULL; results (one per problem)
 * and rounds may be NULL. */
int lddmm_batch_run(lddmm_batch *b, const double *iniAlpStks, double *optAlpStks, int itrMax, double tolVal,
                    double wolfe1, double wolfe2, int memNum, const lddmm_options *opts,
                    const double *targetEnergies, lddmm_result *results, int *rounds);
/* Energies at momentum stacks that are constant in time (no counterpart in
 * the reference; for choosing a start of Cpp/src/LDDMM.cpp:87, iniAlpVec, by
 * objective value): fcnVals[k * P + p] is problem p's energy, in the batch's
 * precision, at alpha_t = scales[k] u_p for every t, with u the concatenated
 * D x N_p blocks baseVecs (host layout).  Energies only: no gradient is
 * formed. */
int lddmm_batch_objective_scan(lddmm_batch *b, const double *baseVecs, const double *scales, int nScales,
                               double *fcnVals);
/* lddmm_batch_run from the constant-in-time start alpha_t = startScales[p] u_p
 * (baseVecs as above; a zero multiple is the zero start), built on the
 * device rather than passed as full stacks (no counterpart in the reference;
 * the start is Cpp/src/LDDMM.cpp:87's iniAlpVec). */
int lddmm_batch_run_scaled(lddmm_batch *b, const double *baseVecs, const double *startScales, double *optAlpStks,
                           int itrMax, double tolVal, double wolfe1, double wolfe2, int memNum,
                           const lddmm_options *opts, const double *targetEnergies, lddmm_result *results,
                           int *rounds);
/* chosen source chunk and the pair interactions per evaluation of each problem */
int lddmm_batch_info(lddmm_batch *b, int *chunk, double *pairsPerEval);
/* the launch shape in use: source points per work item, and targets per
 * lane of the FP32 K.alpha (1, 2, 4) and adjoint (1, 2; 4 in 3D Gaussian
 * batches: the matrix-core adjoint, 256 targets per block) pair launches.
 * With chunk = 0 at creation all three are chosen by timing those launches on
 * the batch (cached per batch signature for the life of the process). */
int lddmm_batch_tuning(lddmm_batch *b, int *chunk, int *kvLanes, int *bwdLanes);
/* fixes the launch shape (chunk rounded up to a multiple of 64; an error when
 * the partial-sum buffer sized at creation cannot hold it) */
int lddmm_batch_set_tuning(lddmm_batch *b, int chunk, int kvLanes, int bwdLanes);
/* on != 0 (default): each batched evaluation replays a hipGraph captured for
 * the current active set; 0: launches are issued one by one.  The default is
 * off when the environment holds ROCPROF* variables (a rocprofv3 run) and is
 * set by LDDMM_BATCH_GRAPHS=0/1. */
int lddmm_batch_set_graphs(lddmm_batch *b, int on);
void lddmm_batch_destroy(lddmm_batch *b);

/* ---------------------------------------------------------------- learned warm start
 * (no counterpart in the reference).  A per-vertex multilayer perceptron
 * maps 32 point features of the source to a momentum vector used as the
 * initial L-BFGS iterate.  Trained on the GPU with BF16 MFMA GEMMs (FP32
 * accumulation, FP32 master weights, Adam). */
#define LDDMM_FEATURES 32
/* featMat (N x 32 row-major, FP32) of a 3D triangulated source (host
 * layout ndeMat D x N, elmVtxMat 3 x E 0-based) against a target varifold
 * (host layout) for a deformation kernel of width knlWidth. */
int lddmm_point_features(float *featMat, int dimNum, const double *ndeMat, int ndeNum, const int32_t *elmVtxMat,
                         int elmNum, const double *tgtCenPosMat, const double *tgtUniDirMat,
                         const double *tgtElmVolVec, int tgtElmNum, double knlWidth);
typedef struct lddmm_mlp lddmm_mlp;
/* layers >= 2 dense layers: 32 -> hidden -> ... -> hidden -> 3 (hidden a
 * multiple of 32), ReLU between layers, He-uniform init from seed. */
int lddmm_mlp_create(lddmm_mlp **out, int hidden, int layers, uint64_t seed);
int lddmm_mlp_param_count(lddmm_mlp *m, int *count);
/* parameters in the order W0 (32 x hidden row-major), b0, W1, b1, ...; the
 * last layer is stored with 32 output columns of which the first 3 are used */
int lddmm_mlp_get_params(lddmm_mlp *m, float *params);
int lddmm_mlp_set_params(lddmm_mlp *m, const float *params);   /* also resets Adam */
/* outMat (rows x 3 row-major) from featMat (rows x 32); host arrays */
int lddmm_mlp_forward(lddmm_mlp *m, const float *featMat, int rows, float *outMat);
/* steps Adam steps (lr) on the mean squared error to tgtMat (rows x 3),
 * minibatches of batchRows consecutive rows cycling through the data (the
 * caller shuffles); losses[s] (may be NULL) = loss before step s.  With
 * steps == 1, lr == 0 and batchRows >= rows, grad (may be NULL) receives the
 * full-batch gradient in the parameter order and nothing changes. */
int lddmm_mlp_train(lddmm_mlp *m, const float *featMat, const float *tgtMat, int rows, int steps, int batchRows,
                    float lr, double *losses, float *grad);
void lddmm_mlp_destroy(lddmm_mlp *m);
/* C (M x N row-major FP32) = op(A) op(B), A and B BF16 bit patterns
 * (row-major, transposed when transA / transB), on the MFMA GEMM kernel the
 * MLP uses; M, N multiples of 32, K of 16.  Host arrays. */
int lddmm_gemm_bf16(float *C, const uint16_t *A, const uint16_t *B, int M, int N, int K, int transA, int transB,
                    int splitK);

/* ---------------------------------------------------------------- data formats
 * getElmIfo (Cpp/src/getElmIfo.m:1-59) from a 0-based V x E table; writes
 * the (1 + maxAdj*V) x N table when elmIfoMat != NULL and returns the row
 * count (negative on error). */
int lddmm_get_elm_ifo(int32_t *elmIfoMat, const int32_t *elmVtxMat, int vtxPerElm, int elmNum, int ndeNum);
/* The spatial order in which objectives and batches hold a problem (no
 * counterpart in the reference, which keeps the caller's order): perm[s] =
 * the point in slot s (0-based) of ptsMat (D x N, host layout), such that
 * every aligned block of 64 or 128 slots is one subtree of a median
 * bisection and hence spatially compact.  Host-only; inputs and outputs of
 * every other entry point stay in the caller's order. */
int lddmm_spatial_order(int32_t *perm, const double *ptsMat, int ptsNum, int dimNum);
/* Culling in the FP32 policy (no counterpart in the reference): a range of
 * sources whose Gaussian kernel values (deformation kernel, or the
 * varifold's 'G' position kernel) with a tile of targets are all below
 * 2^-48 is skipped; the terms left out are below 2^-48 times the terms'
 * other factors, 24 binary orders below the FP32 rounding of the sums.  On
 * by default; off computes every pair (results agree to a few FP64 ulps, for
 * checking).
 * Process-wide, read at every launch. */
int lddmm_set_pair_culling(int on);
/* The culling level (no counterpart in the reference): tile pairs whose
 * Gaussian values are all below 2^-bits are skipped, bits in 24..126, 48 by
 * default.  Process-wide, read at every launch. */
int lddmm_set_pair_cull_level(int bits);
/* Culling by group (no counterpart in the reference): the culling kernels
 * test precomputed bounds of aligned 64-source groups and stage only the
 * groups in range, four to a tile (on by default); off tests each staged
 * quarter tile instead.  The FP32 sums are the same per group, so results
 * agree to FP64 rounding.  Process-wide, read at every launch. */
int lddmm_set_pair_groups(int on);
/* Expanded form of the FP32 Gaussian pair terms (no counterpart in the
 * reference): a pair forms r^2 - |x_i|^2 = |x_j|^2 - 2 x_i.x_j with D
 * multiply-adds and the factor exp2(-|x_i|^2) of the Gaussian is applied
 * once per target and range of sources; a target tile whose radius is too
 * large for that factor keeps the direct form.  On by default; off uses the
 * direct form everywhere (results differ by FP32 rounding).  Process-wide,
 * read at every launch. */
int lddmm_set_pair_expansion(int on);
/* Executed-pair count (no counterpart in the reference; a diagnostic for pair
 * rates): between begin and end, the operators that cull (FP32 Gaussian
 * K.alpha, adjoint and 'G' varifold pair kernels) add the pairs they compute
 * to device counters; end synchronises the device and returns the counts in
 * executedPairs[3]: K.alpha, adjoint, varifold.
 * Costs one atomic per computed source range while on; batches launch
 * directly instead of replaying graphs meanwhile.  Process-wide. */
int lddmm_pair_count_begin(void);
/* Measured instruction rates of the device (no counterpart in the reference;
 * the denominators of utilisation figures): rates[0] v_exp_f32, rates[1] FP32
 * FMA, rates[2] FP64 FMA, per second over the whole chip, each from a grid of
 * 4,096 blocks running independent dependency chains. */
int lddmm_measure_rates(double *rates);
int lddmm_pair_count_end(unsigned long long *executedPairs);
/* Element varifold of a mesh (the target preprocessing of the examples,
 * Cpp/examples/circle2square/run_circle2square.m:21-24; the same formulas as
 * landmarksToVarifoldKernel, Cpp/src/varifold3D.cpp:8-39 and
 * Cpp/src/varifold2D.cpp:8-32): centres D x E, unit directions (2D: unit
 * tangent; 3D: unit normal (v1-v0)x(v2-v0)/|.|) D x E, volumes (length or
 * area) E, from ndeMat D x N and a 0-based V x E element table.  Host layout. */
int lddmm_element_varifold(double *cenPosMat, double *uniDirMat, double *elmVolVec, const double *ndeMat,
                           const int32_t *elmVtxMat, int ndeNum, int elmNum, int dimNum);
/* computeNodePosition (Cpp/src/computeNodePosition.m:1-20): ndeStk D x N x T,
 * vlcStk D x N x (T-1) (either may be NULL).  Host layout. */
int lddmm_compute_node_position(double *ndeStk, double *vlcStk, const double *iniNdeMat,
                                const double *alpStk, int knlOrder, double knlWidth, double timeStp,
                                int timeNum, int ndeNum, int dimNum);

/* ---------------------------------------------------------------- utilities
 * (no reference counterpart) */
const char *lddmm_last_error(void);
int lddmm_device_count(int *count);
int lddmm_device_malloc(void **ptr, size_t bytes);
int lddmm_device_free(void *ptr);
int lddmm_memcpy_htod(void *dst, const void *src, size_t bytes);
int lddmm_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int lddmm_device_synchronize(void);
/* Kernel-time accounting: HIP events around every launch group on the
 * stream it runs on.  names[k] / ms[k] / counts[k] for k < return value. */
int lddmm_profile_enable(int on);
int lddmm_profile_reset(void);
int lddmm_profile_read(const char **names, double *ms, long long *counts, int maxEntries);

#ifdef __cplusplus
}
#endif

#endif
