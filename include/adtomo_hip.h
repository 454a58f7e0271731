/*
 * adtomo_hip.h -- C ABI of the MI355X (gfx950) Eikonal forward/adjoint.
 *
 * Drop-in replacement for the TensorFlow custom ops that ADTomo.jl loads
 * from deps/CustomOps (src/eikonal_op.jl:3-32):
 *
 *   Eikonal / EikonalGrad              deps/CustomOps/Eikonal/Eikonal.cpp:20-63
 *   EikonalThreeD / EikonalThreeDGrad  deps/CustomOps/Eikonal3D/EikonalThreeD.cpp:14-65
 *
 * Conventions
 *   - Plain C types only; every buffer is caller-owned device memory (HBM)
 *     unless the ADT_HOST_POINTERS flag is passed where a `flags` argument
 *     exists.  Inputs are never written.
 *   - Layout is the reference's: 3D fields are row-major (m, n, l) with
 *     flat index i*n*l + j*l + k (Eikonal3D.cpp:6-8); 2D fields have
 *     (n+1) rows (depth j) of (m+1) columns (width i), flat j*(m+1) + i
 *     (Eikonal.h:22-52).  The _strided twins of the single-event ops take
 *     any per-axis strides instead (a Julia column-major Array as it is).
 *   - Sources are 1-based (srcx = column i+1, srcy = row j+1), like the
 *     reference op's inputs (Eikonal.cpp:125 subtracts 1).
 *   - Return value: ADT_OK (0) or a negative ADT_E* code; adt_last_error()
 *     gives the message (thread-local).  Where the reference printed
 *     (non-convergence, verbose iterations) this library fills adt_info.
 *   - `stream` is a hipStream_t (NULL = default stream).  Calls that return
 *     adt_info synchronise the stream; otherwise calls are asynchronous.
 *   - FP64 (_f64) follows the reference bit for bit in the forward; the
 *     _f32 twins evaluate the same expressions in binary32.
 */
#ifndef ADTOMO_HIP_H
#define ADTOMO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ADT_OK 0
#define ADT_EINVAL (-1)   /* bad dimensions, source out of range, NULL buffer */
#define ADT_EHIP (-2)     /* HIP runtime error */
#define ADT_ENOMEM (-3)   /* device allocation failed */
#define ADT_ENODEV (-4)   /* no HIP device */

/* Pointer-kind flag: all data pointers of the call are host memory; the
 * library stages them through device memory. */
#define ADT_HOST_POINTERS 0x1
/* Batched calls: run the forward, the adjoint, or both. */
#define ADT_FORWARD 0x2
#define ADT_ADJOINT 0x4
/* _f32 calls: FP32 storage with every forward local solve computed in FP64
 * and rounded once to FP32 (the pure binary32 forward loses accuracy to
 * cancellation in the 3D quadratic when travel times are large against
 * f*h).  No effect on _f64 calls. */
#define ADT_F32_COMPUTE_F64 0x8

#define ADT_MAX_LOGGED_ITERS 100

/* Per-call (per-event) diagnostics.  Replaces the reference's printf:
 * "ERROR: Eikonal does not converge!" (Eikonal.h:88-90) becomes
 * converged == 0; "Iteration %d, Error = %0.6e" (Eikonal3D.cpp:82-84)
 * becomes err_log[0..iterations-1]. */
typedef struct adt_info {
  int32_t iterations;    /* forward iterations run (K) */
  int32_t converged;     /* 1 if the stop rule fired within the cap */
  double err;            /* last stop-rule value (3D max-norm, 2D rel-L2); with NaN
                            differences, the reference's memory-order value (the
                            maximum after the last one) */
  int32_t ambiguous;     /* 2D: err within 1e-12 (rel) of the 1e-8 threshold */
  int32_t adj_sweeps;    /* adjoint Gauss-Seidel sweeps run */
  int32_t adj_converged; /* 1 if the adjoint reached its exact fixed point */
  int32_t fault;         /* nonzero: a kernel's bounded wait for another wave
                            timed out, so this event's results are invalid */
  int64_t n_singular;    /* 2D adjoint: non-source nodes with no upwind
                            neighbour (singular rows in the reference) */
  double err_log[ADT_MAX_LOGGED_ITERS];
} adt_info;

/* One event of a batched 3D call (one eikonal3d op of scripts/inversion.jl:61).
 * Boxes may differ in size per event. */
typedef struct adt_event3d {
  int64_t m, n, l;      /* box dims, each >= 2 */
  double h;             /* grid spacing */
  double tol;           /* forward stop tolerance (max-norm) */
  const void *u0;       /* [m*n*l] initial travel times */
  const void *f;        /* [m*n*l] slowness */
  const void *grad_u;   /* [m*n*l] adjoint source dL/du (ADT_ADJOINT) */
  void *u;              /* [m*n*l] out: travel times (in for ADT_ADJOINT only) */
  void *grad_u0;        /* [m*n*l] out: dL/du0 (may be NULL) */
  void *grad_f;         /* [m*n*l] out: dL/df (ADT_ADJOINT) */
} adt_event3d;

const char *adt_version(void);
const char *adt_last_error(void);

/* ---- 2D: op Eikonal (Eikonal.cpp:20-46, EikonalOp::Compute :74-127) ---- */
int adt_eikonal2d_fwd_f64(const double *f, int64_t m, int64_t n, int64_t srcx, int64_t srcy,
                          double h, double *u, adt_info *info, int32_t flags, void *stream);
int adt_eikonal2d_fwd_f32(const float *f, int64_t m, int64_t n, int64_t srcx, int64_t srcy,
                          float h, float *u, adt_info *info, int32_t flags, void *stream);

/* ---- 2D: op EikonalGrad (Eikonal.cpp:48-63, EikonalGradOp::Compute :141-221).
 * Only grad_f is produced (the reference leaves the other outputs unset). */
int adt_eikonal2d_bwd_f64(const double *grad_u, const double *u, const double *f, int64_t m,
                          int64_t n, int64_t srcx, int64_t srcy, double h, double *grad_f,
                          adt_info *info, int32_t flags, void *stream);
int adt_eikonal2d_bwd_f32(const float *grad_u, const float *u, const float *f, int64_t m,
                          int64_t n, int64_t srcx, int64_t srcy, float h, float *grad_f,
                          adt_info *info, int32_t flags, void *stream);

/* ---- 3D: op EikonalThreeD (EikonalThreeD.cpp:14-45, Compute :77-140).
 * `verbose` prints the reference's per-iteration line after the solve. */
int adt_eikonal3d_fwd_f64(const double *u0, const double *f, double h, int64_t m, int64_t n,
                          int64_t l, double tol, int32_t verbose, double *u, adt_info *info,
                          int32_t flags, void *stream);
int adt_eikonal3d_fwd_f32(const float *u0, const float *f, float h, int64_t m, int64_t n,
                          int64_t l, float tol, int32_t verbose, float *u, adt_info *info,
                          int32_t flags, void *stream);

/* ---- 3D forward without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as adt_eikonal3d_fwd_f64 (bit for bit), device pointers only; the
 * iteration snapshot, descriptor and diagnostics live in a fixed device
 * array of 16 MiB + 4 KiB that each device holds until the process ends, so
 * events are limited to 2^21 nodes (ADT_EINVAL beyond) and calls that
 * overlap on one device race (serialise them, e.g. on one stream). */
int adt_eikonal3d_fwd_noalloc_f64(const double *u0, const double *f, double h, int64_t m,
                                  int64_t n, int64_t l, double tol, int32_t verbose, double *u,
                                  adt_info *info, void *stream);

/* ---- 3D: op EikonalThreeDGrad (EikonalThreeD.cpp:47-65, Compute :154-250).
 * Produces grad_u0 and grad_f (the only meaningful outputs of the reference). */
int adt_eikonal3d_bwd_f64(const double *grad_u, const double *u, const double *u0,
                          const double *f, double h, int64_t m, int64_t n, int64_t l,
                          double *grad_u0, double *grad_f, adt_info *info, int32_t flags,
                          void *stream);
int adt_eikonal3d_bwd_f32(const float *grad_u, const float *u, const float *u0, const float *f,
                          float h, int64_t m, int64_t n, int64_t l, float *grad_u0,
                          float *grad_f, adt_info *info, int32_t flags, void *stream);

/* ---- Strided layouts of the four single-event ops above (the per-axis
 * strides of the boundary that src/eikonal_op.jl:3-32's Julia callers need
 * to pass arrays without a transposed copy).  Every array of a call holds
 * element (i, j, k) at offset i*strides[0] + j*strides[1] + k*strides[2]
 * elements; 2D arrays hold (i, j) (column i, row j) at i*strides[0] +
 * j*strides[1].  A Julia column-major (m, n, l) Array passes
 * strides {1, m, m*n}; a Julia (n+1, m+1) Matrix passes {n+1, 1}; the
 * row-major layout of the calls above is {n*l, l, 1} / {1, m+1}.  Strides
 * of axes longer than one must be positive and must not map two elements
 * to one address (ADT_EINVAL otherwise).  Positions between elements (a
 * padded layout) are neither read nor written.  The library gathers the
 * arrays to the row-major layout on the device and scatters the results
 * back, so the results equal those of the row-major calls bit for bit. */
int adt_eikonal2d_fwd_strided_f64(const double *f, int64_t m, int64_t n, int64_t srcx,
                                  int64_t srcy, double h, const int64_t strides[2], double *u,
                                  adt_info *info, int32_t flags, void *stream);
int adt_eikonal2d_fwd_strided_f32(const float *f, int64_t m, int64_t n, int64_t srcx,
                                  int64_t srcy, float h, const int64_t strides[2], float *u,
                                  adt_info *info, int32_t flags, void *stream);
int adt_eikonal2d_bwd_strided_f64(const double *grad_u, const double *u, const double *f,
                                  int64_t m, int64_t n, int64_t srcx, int64_t srcy, double h,
                                  const int64_t strides[2], double *grad_f, adt_info *info,
                                  int32_t flags, void *stream);
int adt_eikonal2d_bwd_strided_f32(const float *grad_u, const float *u, const float *f,
                                  int64_t m, int64_t n, int64_t srcx, int64_t srcy, float h,
                                  const int64_t strides[2], float *grad_f, adt_info *info,
                                  int32_t flags, void *stream);
int adt_eikonal3d_fwd_strided_f64(const double *u0, const double *f, double h, int64_t m,
                                  int64_t n, int64_t l, double tol, int32_t verbose,
                                  const int64_t strides[3], double *u, adt_info *info,
                                  int32_t flags, void *stream);
int adt_eikonal3d_fwd_strided_f32(const float *u0, const float *f, float h, int64_t m,
                                  int64_t n, int64_t l, float tol, int32_t verbose,
                                  const int64_t strides[3], float *u, adt_info *info,
                                  int32_t flags, void *stream);
int adt_eikonal3d_bwd_strided_f64(const double *grad_u, const double *u, const double *u0,
                                  const double *f, double h, int64_t m, int64_t n, int64_t l,
                                  const int64_t strides[3], double *grad_u0, double *grad_f,
                                  adt_info *info, int32_t flags, void *stream);
int adt_eikonal3d_bwd_strided_f32(const float *grad_u, const float *u, const float *u0,
                                  const float *f, float h, int64_t m, int64_t n, int64_t l,
                                  const int64_t strides[3], float *grad_u0, float *grad_f,
                                  adt_info *info, int32_t flags, void *stream);

/* ---- Batched 3D: many independent events (the per-station op loop of
 * scripts/inversion.jl:47-62, sharded over MPI ranks by
 * src/mpi_optimize.jl:58-70) in one persistent launch.  flags: ADT_FORWARD
 * and/or ADT_ADJOINT.  `infos` (host, nev entries) may be NULL; when given,
 * the call synchronises the stream. */
int adt_eikonal3d_batched_f64(const adt_event3d *events, int32_t nev, int32_t flags,
                              adt_info *infos, void *stream);
int adt_eikonal3d_batched_f32(const adt_event3d *events, int32_t nev, int32_t flags,
                              adt_info *infos, void *stream);

/* ---- Plain per-event launch loop (baseline): one kernel launch per
 * hyperplane, host-synchronised convergence check per iteration. */
int adt_eikonal3d_fwd_plain_f64(const double *u0, const double *f, double h, int64_t m,
                                int64_t n, int64_t l, double tol, double *u, adt_info *info,
                                void *stream);
int adt_eikonal3d_bwd_plain_f64(const double *grad_u, const double *u, const double *u0,
                                const double *f, double h, int64_t m, int64_t n, int64_t l,
                                double *grad_u0, double *grad_f, adt_info *info, void *stream);

/* ---- hipGraph-captured batched forward+adjoint.  The graph owns its
 * descriptor table and workspace; the events' buffers must outlive it.
 * dtype: 0 = FP64, 1 = FP32. */
typedef struct adt_graph adt_graph;
int adt_graph_create(const adt_event3d *events, int32_t nev, int32_t flags, int32_t dtype,
                     adt_graph **out, void *stream);
int adt_graph_launch(adt_graph *g, void *stream);
int adt_graph_infos(adt_graph *g, adt_info *infos, void *stream);
int adt_graph_destroy(adt_graph *g);

/* ---- Launch tuning of the 3D kernels (one global setting per precision;
 * dtype 0 = FP64, 1 = FP32, and optionally one per box-shape class, see
 * adt_tune_class_set).  Read when a call runs or a graph is created (one
 * graph may hold events of several classes, each with its class's
 * setting); every setting gives the same travel times (bit for bit) and
 * gradients (to rounding).  The reference has no counterpart (its op is
 * single-threaded CPU code). */
typedef struct adt_tune {
  int32_t diag_max_l; /* forward: events with l <= diag_max_l use the
                         diagonal-streaming kernel (any l, in k chunks of
                         64; default 2^30 = all), the others k-streaming */
  int32_t adj_wg;     /* adjoint workgroup size: 256, 512 or 1024 */
  int32_t order;      /* 0: events in the given order; 1: largest first */
  int32_t fuse_max_n; /* a call or graph with both ADT_FORWARD and
                          ADT_ADJOINT whose every event takes the diagonal
                          forward and has at most fuse_max_n nodes runs one
                          kernel that does each event's forward and then its
                          adjoint (an event's adjoint starts as soon as its
                          forward converges; adj_wg does not apply); larger
                          events, or 0, use separate forward and adjoint
                          kernels.  Default 2^21: the fused adjoint runs on
                          the forward's 8 waves, which pays off for events
                          of about 10^6 nodes and costs on much larger ones */
  float fwd_ms;       /* adt_autotune: measured forward / adjoint time per */
  float adj_ms;       /* launch of the chosen setting (ms) */
  int32_t adj_algo;   /* adjoint traversal: 0 = over 8x8x8 tiles, each
                         resolved in LDS by one wave (default; events of at
                         most 65536 tiles, larger ones take 1), 1 = node by
                         node over the whole event (the only one that fuses
                         with the forward, see fuse_max_n) */
  int32_t fwd_segs;   /* diagonal forward: the k axis is cut into fwd_segs
                         segments of up to 8 nodes, one per wave (1..8; 0 =
                         8), and swept in chunks of 8 * fwd_segs nodes */
} adt_tune;
int adt_get_tune(int32_t dtype, adt_tune *out);
int adt_set_tune(int32_t dtype, const adt_tune *t);

/* Box-shape classes: events whose k extent l needs the same number of
 * 8-node blocks (l_blocks = ceil(l / 8)), whose shorter horizontal side
 * min(m, n) the same number of 64-row bands (bands = ceil(min(m, n) / 64))
 * and whose node count the same binary order (log2n = floor(log2(m n l)))
 * share a class.  A class setting replaces the global one's per-event
 * fields (diag_max_l, fwd_segs, adj_algo) for the events of that class;
 * adj_wg, order and fuse_max_n stay global. */
typedef struct adt_tune_class {
  int32_t l_blocks, bands, log2n;
  adt_tune tune;
} adt_tune_class;
/* Sets (t != NULL) or removes (t == NULL) the setting of the class of box
 * (m, n, l). */
int adt_tune_class_set(int32_t dtype, int64_t m, int64_t n, int64_t l, const adt_tune *t);
/* Number of class settings of dtype; adt_tune_class_get returns the i-th
 * (in class order). */
int adt_tune_class_count(int32_t dtype);
int adt_tune_class_get(int32_t dtype, int32_t i, adt_tune_class *out);
int adt_tune_class_clear(int32_t dtype);
/* Writes every setting (global and class, both precisions) to a JSON file,
 * or reads such a file back (replacing the settings it holds: the global
 * ones it names and all class ones). */
int adt_tune_save(const char *path);
int adt_tune_load(const char *path);
/* Times the settings on the given events (device pointers, as for
 * adt_eikonal3d_batched_*; `reps` timed launches each): for each box-shape
 * class among them, on that class's events, the forward kernel and its
 * segment count and the adjoint traversal, stored as the class's setting;
 * then, on all events, the event order, the node-level adjoint's workgroup
 * size and the fusion threshold, stored as the global setting for `dtype`
 * and returned in `best` (may be NULL).  The events' outputs hold a valid
 * solve afterwards.  With ADT_ADJOINT but not ADT_FORWARD, u must already
 * hold the forward travel times. */
int adt_autotune(const adt_event3d *events, int32_t nev, int32_t flags, int32_t dtype,
                 int32_t reps, adt_tune *best, void *stream);

/* ---- Objective around the path (scripts/inversion.jl), all on the device.
 * Grid coordinates are 1-based and real-valued, as in the reference's
 * station/event tables (x along i, y along j, z along k). */

/* One station's initial travel times (scripts/inversion.jl:48-60):
 * u0 = 1000, and at the 8 nodes around (x, y, z)
 * u0 = |(x,y,z) - node| * h / vel0[node] * scale
 * (scale = 1 for P; vp/vs for S, scripts/inversion_joint.jl:65-73). */
typedef struct adt_source3d {
  double x, y, z;  /* 1-based coordinates, 1 <= floor, ceil <= dim */
  double scale;
  void *u0;        /* out: [m*n*l] device */
} adt_source3d;
int adt_source_u0_3d_batched_f64(const adt_source3d *src, int32_t nsrc, const double *vel0,
                                 int64_t m, int64_t n, int64_t l, double h, void *stream);
int adt_source_u0_3d_batched_f32(const adt_source3d *src, int32_t nsrc, const float *vel0,
                                 int64_t m, int64_t n, int64_t l, float h, void *stream);

/* Receiver gather + misfit of one solve (station) against nrec receivers
 * (events) at host coordinates (rx, ry, rz) (scripts/inversion.jl:64-105):
 * tcal = trilinear interpolation of u (a dimension whose floor and ceil
 * coincide is not interpolated), loss = sum over tobs != -1 of
 * weight * rho(tobs - tcal) with rho(r) = r^2 (the reference) or, when
 * huber_delta > 0, the Huber form rho(r) = r^2 for |r| <= delta and
 * 2 delta |r| - delta^2 beyond; grad_u = d loss / d u (dense, zero except
 * at the interpolation nodes; contributions of receivers sharing a node are
 * added atomically, so their order may vary between runs).  tcal and
 * grad_u may be NULL; loss is a device scalar. */
typedef struct adt_misfit3d {
  int64_t m, n, l;
  const void *u;       /* [m*n*l] travel times */
  const void *tobs;    /* [nrec] observed times, -1 = missing */
  const void *weight;  /* [nrec] quality weights */
  void *tcal;          /* out: [nrec] */
  void *grad_u;        /* out: [m*n*l] */
  void *loss;          /* out: scalar */
  double huber_delta;  /* <= 0: squared residuals (scripts/inversion.jl:102) */
} adt_misfit3d;
int adt_misfit3d_batched_f64(const adt_misfit3d *items, int32_t nitem, const double *rx,
                             const double *ry, const double *rz, int64_t nrec, void *stream);
int adt_misfit3d_batched_f32(const adt_misfit3d *items, int32_t nitem, const double *rx,
                             const double *ry, const double *rz, int64_t nrec, void *stream);

/* ---- One inversion step per launch: for each event its forward, then the
 * receiver misfit of its travel times (items[e], as adt_misfit3d_batched_*;
 * items[e] must name events[e]'s box, u and grad_u), then the adjoint with
 * that misfit gradient as the adjoint source (scripts/inversion.jl:47-105,
 * the forward, loss and gradient of each station).  flags must hold
 * ADT_FORWARD and ADT_ADJOINT (plus ADT_F32_COMPUTE_F64 for _f32).  With
 * the fused launch setting and every event on the diagonal forward, one
 * kernel runs the three phases per event; otherwise three launches run in
 * stream order.  The receiver coordinates are host arrays; the call waits
 * for their copy.  The reference has no single counterpart (its op, the
 * loss and the gradient op run as separate graph nodes). */
int adt_eikonal3d_misfit_batched_f64(const adt_event3d *events, const adt_misfit3d *items,
                                     int32_t nev, const double *rx, const double *ry,
                                     const double *rz, int64_t nrec, int32_t flags,
                                     adt_info *infos, void *stream);
int adt_eikonal3d_misfit_batched_f32(const adt_event3d *events, const adt_misfit3d *items,
                                     int32_t nev, const double *rx, const double *ry,
                                     const double *rz, int64_t nrec, int32_t flags,
                                     adt_info *infos, void *stream);
/* hipGraph of the same step (as adt_graph_create); the items and the
 * receiver coordinates are copied into the graph's workspace at creation. */
int adt_graph_create_misfit(const adt_event3d *events, const adt_misfit3d *items, int32_t nev,
                            const double *rx, const double *ry, const double *rz, int64_t nrec,
                            int32_t flags, int32_t dtype, adt_graph **out, void *stream);

/* Model parametrisation (scripts/inversion.jl:42-43,61): fvar = 2 sigmoid(x)
 * - 1 + vel0, f = 1 / fvar (f may be NULL). */
int adt_model3d_f64(const double *x, const double *vel0, int64_t N, double *fvar, double *f,
                    void *stream);
int adt_model3d_f32(const float *x, const float *vel0, int64_t N, float *fvar, float *f,
                    void *stream);
/* Its adjoint: grad_x = (grad_fvar - grad_f / fvar^2) * 2 sigmoid(x)
 * (1 - sigmoid(x)); grad_f or grad_fvar may be NULL. */
int adt_model3d_grad_f64(const double *x, const double *fvar, const double *grad_f,
                         const double *grad_fvar, int64_t N, double *grad_x, void *stream);
int adt_model3d_grad_f32(const float *x, const float *fvar, const float *grad_f,
                         const float *grad_fvar, int64_t N, float *grad_x, void *stream);

/* S-wave slowness of the joint inversion (scripts/inversion_joint.jl:79):
 * f_s = scale / fvar (scale = the vp/vs ratio). */
int adt_slowness3d_f64(const double *fvar, int64_t N, double scale, double *f_s, void *stream);
int adt_slowness3d_f32(const float *fvar, int64_t N, float scale, float *f_s, void *stream);
/* Its adjoint: grad_fvar += -scale * grad_fs / fvar^2 (in place), and
 * grad_scale (device scalar, may be NULL) = sum of grad_fs / fvar. */
int adt_slowness3d_grad_f64(const double *fvar, int64_t N, double scale, const double *grad_fs,
                            double *grad_fvar, double *grad_scale, void *stream);
int adt_slowness3d_grad_f32(const float *fvar, int64_t N, float scale, const float *grad_fs,
                            float *grad_fvar, float *grad_scale, void *stream);

/* Smoothing regulariser (scripts/inversion.jl:107-121): value = lambda *
 * sum |fvar - box(fvar)| with box = the mean over an sh x sh x sv window
 * (odd sizes) with periodic wrap; grad_fvar = d value / d fvar (may be
 * NULL).  value is a device scalar. */
int adt_smooth3d_l1_f64(const double *fvar, int64_t m, int64_t n, int64_t l, int32_t sh,
                        int32_t sv, double lambda, double *grad_fvar, double *value,
                        void *stream);
int adt_smooth3d_l1_f32(const float *fvar, int64_t m, int64_t n, int64_t l, int32_t sh,
                        int32_t sv, float lambda, float *grad_fvar, float *value, void *stream);

/* out = sum over k < narr of arrays[k] (host array of device pointers),
 * in that order (e.g. dL/df over the stations of one rank). */
int adt_sum_arrays_f64(const double *const *arrays, int32_t narr, int64_t N, double *out,
                       void *stream);
int adt_sum_arrays_f32(const float *const *arrays, int32_t narr, int64_t N, float *out,
                       void *stream);

/* ---- Diagnostics: device arithmetic of the path, for bit-exact checks
 * against host IEEE results.  op 0: sqrt(a); 1: a/3; 2: a/2; 3: a/b;
 * 4: 3D local solve (a = 3 values, b = f,h per element);
 * 5: 2D local solve (a = 2 values, b = f,h per element);
 * 6: plain copy out = a (a streaming kernel of known traffic, n elements
 *    read and written once, for calibrating memory counters);
 * 7: sqrt(a) as the forward's step loops take it (the fast refinement, the
 *    IEEE routine where it flags the operand);
 * 8: op 4 as the forward's step loops take it (the values sorted by
 *    minima and maxima, fast square roots; the update redone as op 4 where
 *    an input is NaN or -0 or a discriminant is flagged). */
int adt_selftest_arith_f64(int32_t op, const double *a, const double *b, double *out,
                           int64_t n, void *stream);
int adt_selftest_arith_f32(int32_t op, const float *a, const float *b, float *out, int64_t n,
                           void *stream);

#ifdef __cplusplus
}
#endif
#endif /* ADTOMO_HIP_H */
