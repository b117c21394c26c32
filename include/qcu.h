/*
 * qcu.h -- C ABI of the MI355X-native qcu hot path (libqcu.so).
 *
 * The first block keeps the six entry points of the reference's include/qcu.h
 * (include/qcu.h:17-30) with identical signatures and meanings, so that the
 * PyQcu Cython binding (python/PyQuda/pyquda/src/qcu.pxd:1-8) and any C caller
 * link against this library unchanged:
 *   - every pointer is a DEVICE pointer in the external PyQuda layouts
 *       gauge   complex128 [4 mu][2 parity][Lt][Lz][Ly][Lx/2][3][3]
 *       fermion complex128 [Lt][Lz][Ly][Lx/2][4 spin][3 colour]   (one parity)
 *       full    complex128 [2 parity][Lt][Lz][Ly][Lx/2][4][3]     (even first)
 *   - parity is the OUTPUT parity; the input fermion lives on the other one;
 *   - boundaries are periodic in all four directions (the reference kernel
 *     wraps periodically, include/define.h:127-134);
 *   - calls are synchronous on return and leave every input bit-identical
 *     (the reference transposes gauge and fermion_in in place and back);
 *   - the converted links and the clover term stay resident while the same
 *     gauge pointer holds the same contents (a 64-bit hash of the field,
 *     computed on the device at every call, decides; the reference rebuilds
 *     them at every call).
 * Errors never abort the process and never throw: the void entry points record
 * the failure, readable with qcuGetLastError()/qcuGetLastErrorString().
 *
 * The second block is new: handle-based entry points with runtime kappa,
 * c_sw, time boundary, tolerance, precision, batching and the learned
 * preconditioner.  They return 0 on success or an error code.
 */
#ifndef QCU_H
#define QCU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct QcuParam_s {
  int lattice_size[4];
} QcuParam;

/* ---- reference entry points (include/qcu.h:21-30) ------------------------ */

/* out = D_{p,1-p} in: Wilson hop on output parity p, no kappa (src/draft.cu:3-28) */
void dslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity);
/* out = A_p^-1 D_{p,1-p} in with A = 1 - sum sigma F at kappa*csw = 1
 * (src/nccl_clover_dslash.cu:3-79) */
void dslashCloverQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity);
/* empty in the reference (src/draft.cu:29); here the same operator as
 * ncclDslashQcu, grid must be {1,1,1,1} */
void mpiDslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity,
                  QcuParam *grid);
/* empty in the reference (src/draft.cu:30-31); kept as a no-op */
void mpiBistabCgQcu(void *gauge, QcuParam *param, QcuParam *grid);
/* Wilson hop, one GPU: grid must be {1,1,1,1} (src/nccl_wilson_dslash.cu:3-19) */
void ncclDslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity,
                   QcuParam *grid);
/* solves (1 - kappa D) x = b on the full lattice, kappa = 0.125, even-odd
 * BiCGSTAB on the odd sites stopped when Re<r,r> < 1e-9 or after 1000
 * iterations (src/nccl_wilson_bistabcg.cu:7-27, include/define.h:67-69);
 * fermion_out = x, fermion_in = b, full-lattice layout */
void ncclBistabCgQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param,
                     QcuParam *grid);

/* ---- errors (new) ------------------------------------------------------- */
enum {
  QCU_OK = 0,
  QCU_ERR_ARG = 1,            /* null / host pointer, bad size or value */
  QCU_ERR_DIMS = 2,           /* lattice extents */
  QCU_ERR_GRID = 3,           /* grid != {1,1,1,1}: this build runs on one GPU */
  QCU_ERR_NOT_CONVERGED = 4,
  QCU_ERR_BREAKDOWN = 5,
  QCU_ERR_ALLOC = 6,          /* device allocation failed */
  QCU_ERR_STATE = 7,          /* gauge not loaded, precision not resident, ... */
  QCU_ERR_HIP_BASE = 100      /* a HIP call failed: code = 100 + hipError_t */
};
int qcuGetLastError(void);
const char *qcuGetLastErrorString(void);
void qcuClearError(void);

/* ---- handle-based API (new) --------------------------------------------- */
typedef struct QcuLattice_s *QcuLatticeHandle;

enum {
  QCU_PREC_FP64 = 0,
  QCU_PREC_FP32 = 1,
  QCU_PREC_FP16 = 2 /* half storage with one FP32 scale per site */
};

/* kappa, csw (0 => Wilson), t_boundary +1 periodic / -1 antiperiodic */
int qcuLatticeCreate(QcuLatticeHandle *h, const int lattice_size[4], double kappa, double csw,
                     int t_boundary);
int qcuLatticeDestroy(QcuLatticeHandle h);
/* gauge: device pointer, external layout; builds the resident links (FP64,
 * FP32, FP16) and, for csw != 0, the packed clover term and its inverse */
int qcuLatticeLoadGauge(QcuLatticeHandle h, const void *gauge);
/* fills a device gauge buffer (external layout) with U = exp(i eps H), seeded */
int qcuGaugeRandom(void *gauge, const int lattice_size[4], double eps, uint64_t seed);
/* fills n complex128 values with Gaussian (re, im ~ N(0,1)) numbers, seeded */
int qcuGaussian(void *field, long n_complex, uint64_t seed);

/* out = [A_p^-1] D_{p,1-p} in on external one-parity fields; clover = 1 adds
 * the inverse clover of the lattice; prec selects the kernel precision */
int qcuDslash(QcuLatticeHandle h, void *out, const void *in, int parity, int clover, int prec);
/* qcuDslash without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * result as qcuDslash.  The two staged half spinors live in a 32 MiB device
 * array of the library (one per GPU, held until the process ends), so the
 * lattice must satisfy 2 x Vh x (192 bytes FP64, 96 FP32, 52 FP16, each
 * rounded up to 256 bytes) <= 32 MiB (e.g. 16^3x32 in FP64); larger ones are
 * refused with QCU_ERR_DIMS.  Calls that overlap on one GPU share that array
 * and race; the call is synchronous, like qcuDslash. */
int qcuDslashNoAlloc(QcuLatticeHandle h, void *out, const void *in, int parity, int clover, int prec);
/* out = M in = (A - kappa D) in on the full lattice, FP64 */
int qcuMatFull(QcuLatticeHandle h, void *out, const void *in);
/* |b - M x| / |b| in FP64 on the full lattice */
int qcuTrueResidual(QcuLatticeHandle h, const void *x, const void *b, double *res);

typedef struct QcuSolveParam_s {
  double tol;            /* target |b - M x| / |b| (full lattice, FP64) */
  int max_iter;          /* BiCGSTAB iterations (sloppy iterations count) */
  int precision;         /* QCU_PREC_*: inner (sloppy) precision */
  double reliable_delta; /* reliable update when |r| < delta |r_last_update| */
  int chunk;             /* iterations between host checks (graph length); 0: 8, or 4 with precond */
  int use_graph;         /* capture the chunk as a hipGraph and replay it */
  void *precond;         /* QcuNetHandle or NULL */
  int legacy_stop;       /* 1: stop on Re<r,r> < tol absolute (reference rule) */
  int method;            /* QCU_METHOD_*: 0 (a zeroed struct) is BiCGSTAB */
} QcuSolveParam;

enum {
  QCU_METHOD_BICGSTAB = 0, /* even-odd BiCGSTAB (include/lattice_bistabcg.h:221-340) */
  QCU_METHOD_CGNR = 1      /* CG on the normal equations of the odd Schur system, FP64
                              (completes include/lattice_bistabcg.h:341-450, which lacks D^+) */
};

typedef struct QcuSolveInfo_s {
  int iters;
  int reliable_updates;
  int converged;
  int fallback;          /* 1: the preconditioned solve stalled, plain solve used */
  double true_res;
  double seconds;
  int restarts;
  int pad;
} QcuSolveInfo;

/* x, b: full-lattice device fields in the external layout */
int qcuSolve(QcuLatticeHandle h, void *x, const void *b, const QcuSolveParam *p, QcuSolveInfo *info);

/* Times `iters` launches of the internal hop kernel on resident fields
 * (xpay = 0: out = [A^-1] D in; 1: the second half of the Schur operator) with
 * HIP events on the library's stream; block / remap override the tuned launch
 * shape when > 0 / >= 0.  ms = average milliseconds per launch. */
int qcuBenchHop(QcuLatticeHandle h, int prec, int clover, int xpay, int iters, int block, int remap,
                double *ms);

/* ---- batches (new) -------------------------------------------------------
 * Many systems in one launch per kernel: nrhs[k] right-hand sides on lattice
 * lats[k] (interleaved so that they share every link load), for nlat
 * lattices of any shapes.  System order: lattice 0 rhs 0..nrhs[0]-1, then
 * lattice 1, ...; x[i], b[i] are the full-lattice external fields of system i. */
typedef struct QcuBatch_s *QcuBatchHandle;
int qcuBatchCreate(QcuBatchHandle *h, const QcuLatticeHandle *lats, const int *nrhs, int nlat);
int qcuBatchDestroy(QcuBatchHandle h);
int qcuBatchSystems(QcuBatchHandle h, int *nsys);
/* info: one record per system */
int qcuBatchSolve(QcuBatchHandle h, void *const *x, const void *const *b, const QcuSolveParam *p,
                  QcuSolveInfo *info);
/* ms per launch of the batched hop kernel (all systems, one parity) */
int qcuBatchBenchHop(QcuBatchHandle h, int prec, int xpay, int iters, double *ms);
int qcuBatchTune(QcuBatchHandle h, int prec, double *best_ms);

/* launch-shape autotuner: sweeps block size, XCD-aware block order and the
 * instruction-scheduling grouping of the hop kernel for this lattice and
 * precision and keeps the fastest; the table persists in a text file */
int qcuTune(QcuLatticeHandle h, int prec, double *best_ms);
int qcuTuneGet(QcuLatticeHandle h, int prec, int *block, int *remap, int *group);
/* pins the launch shape of the hop kernel for this lattice (or batch) and
 * precision instead of tuning it: block 64, 128 or 256 threads; remap 0/1
 * (XCD-aware block order off/on); group 0, 1 or 2 (scheduling fences after
 * none / every / every second direction).  Solver graphs captured with an
 * earlier shape are re-captured. */
int qcuTuneSet(QcuLatticeHandle h, int prec, int block, int remap, int group);
int qcuBatchTuneSet(QcuBatchHandle h, int prec, int block, int remap, int group);
int qcuTuneSave(const char *path);
int qcuTuneLoad(const char *path);

/* ---- learned preconditioner (new) ----------------------------------------
 * A gauge-covariant linear network K on the lattice h: nlayers layers mapping
 * chan[l] -> chan[l+1] spinor channels (chan[0] = chan[nlayers] = 1, at most 8),
 *   out_o(x) = sum_i sum_mu [A_mu U_mu(x) in_i(x+mu) + B_mu U_mu(x-mu)^dag in_i(x-mu)] + C in_i(x)
 * with learned 4x4 complex spin matrices per (layer, term, o, i), FP32.
 * Weight order: layer, term t (2 mu: forward hop A_mu, 2 mu + 1: backward
 * hop B_mu, 8: C), output channel, input channel, spin row, spin column.
 * Initialised to the identity plus seeded noise of size init_scale.
 * Passed as QcuSolveParam.precond it is applied as a fixed right
 * preconditioner of the odd-site Schur system (odd input, even sites zero,
 * odd part of the output kept). */
typedef struct QcuNet_s *QcuNetHandle;
int qcuNetCreate(QcuNetHandle *h, QcuLatticeHandle lat, int nlayers, const int *chan, uint64_t seed,
                 double init_scale);
int qcuNetDestroy(QcuNetHandle h);
long qcuNetWeightCount(QcuNetHandle h);
/* complex128 weights / gradient on the HOST, n = qcuNetWeightCount */
int qcuNetGetWeights(QcuNetHandle h, double *host, long n);
int qcuNetSetWeights(QcuNetHandle h, const double *host, long n);
int qcuNetGetGradient(QcuNetHandle h, double *host, long n);
/* full-lattice fields [2][Vh][4][3] complex128 (device): out = K in */
int qcuNetForward(QcuNetHandle h, void *out, const void *in);
/* gradient of a real loss L with respect to the weights (read back with
 * qcuNetGetGradient) and, when grad_in is not NULL, to the input, given
 * grad_out = dL/dRe(out) + i dL/dIm(out) at the input `in` */
int qcuNetBackward(QcuNetHandle h, void *grad_in, const void *grad_out, const void *in);
/* Adam training on odd-site sources srcs[k] (one-parity external fields,
 * device) of the loss |Mhat K b - b|^2 / |b|^2, Mhat the FP32 odd Schur
 * operator; losses[e] = mean loss seen during epoch e (may be NULL) */
int qcuNetTrain(QcuNetHandle h, const void *const *srcs, int n, int epochs, int batch, double lr,
                double *losses, double *seconds);
/* loss of one source and its weight gradient (read with qcuNetGetGradient) */
int qcuNetLossGrad(QcuNetHandle h, const void *src, double *loss);
/* mean loss over the sources, no update */
int qcuNetLoss(QcuNetHandle h, const void *const *srcs, int n, double *loss);
/* ms per launch of one layer's forward kernel (both parities) and its
 * algorithmic flops per launch, timed with HIP events on the lattice stream;
 * the flops count 8 per complex multiply-add, 6 in the matrix-core channel
 * mixing (three real products per complex product) */
int qcuNetBenchLayer(QcuNetHandle h, int layer, int iters, double *ms, double *flops);
/* average ms of one preconditioner application as the solver runs it (odd
 * field embedded, forward pass keeping only the output, odd part extracted;
 * no counterpart in the reference) */
int qcuNetBenchApply(QcuNetHandle h, int iters, double *ms);
/* average ms of one layer's kernels as the preconditioner runs them (after
 * one application, on its activations; the last layer of a network whose
 * last layer is fused: its gather pass); no counterpart in the reference */
int qcuNetBenchPrecondLayer(QcuNetHandle h, int layer, int iters, double *ms);
/* out = K in on one-parity (odd) external fields, as applied inside the solver */
int qcuNetApplyOdd(QcuNetHandle h, void *out, const void *in);
/* weights file: one line "QCUNET1 {json metadata}" (channels, weight count,
 * dtype, and the lattice, kappa, c_sw and boundary the network was trained
 * on), then the weights as little-endian FP32 (re, im) pairs in the order of
 * qcuNetGetWeights.  qcuNetLoad requires the same channel list and resets the
 * optimiser state; the weights act site by site, so any lattice can load them. */
int qcuNetSave(QcuNetHandle h, const char *path);
int qcuNetLoad(QcuNetHandle h, const char *path);

#ifdef __cplusplus
}
#endif
#endif
