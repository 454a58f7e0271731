/*
 * CPU oracle for the Eikonal forward/adjoint path (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load this library, and only as the checker or the reported CPU baseline;
 * the product (adtomo.jl_amd) never links or calls it.
 *
 * Built by oracle/Makefile with `-O3 -DNDEBUG -ffp-contract=off`, mirroring
 * deps/CustomOps/CMakeLists.txt:47 of the reference (no FMA contraction, no
 * fast-math), so FP64 results equal the reference CPU op's bit for bit.
 *
 * Function map (reference file:line):
 *   adt_oracle_eikonal2d_fwd_f64  deps/CustomOps/Eikonal/Eikonal.h:14-93
 *   adt_oracle_eikonal2d_bwd_f64  deps/CustomOps/Eikonal/Eikonal.h:95-199
 *   adt_oracle_eikonal3d_fwd_f64  deps/CustomOps/Eikonal3D/Eikonal3D.cpp:11-94
 *   adt_oracle_eikonal3d_bwd_f64  deps/CustomOps/Eikonal3D/Eikonal3D.cpp:96-198
 * The _f32 twins evaluate the same expressions in binary32;
 * adt_oracle_eikonal3d_fwd_m32 keeps binary32 storage and evaluates each
 * local solve in binary64, rounded once (the library's ADT_F32_COMPUTE_F64
 * variant; pinned against an independent numpy formulation in the tests).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define REAL double
#define SFX _f64
#define SQRT sqrt
#define FABS fabs
#define C(x) (x)
#include "eikonal_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS
#undef C

#define REAL float
#define SFX _f32
#define SQRT sqrtf
#define FABS fabsf
#define C(x) ((float)(x))
#include "eikonal_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS
#undef C

/* FP32 storage with FP64 local solves (the library's ADT_F32_COMPUTE_F64
 * variant): the neighbour minima are binary32 values, the local solve of
 * Eikonal3D.cpp:11-28 runs in binary64 on them and its result is rounded
 * once to binary32 before u = min(u_new, u); the stop rule is the binary32
 * twin's.  Same loops as sweep3d / adt_oracle_eikonal3d_fwd. */
#define U3(i, j, k) u[(i) * n * l + (j) * l + (k)]
static void sweep3d_m32(float *u, const float *f, double h, int64_t m, int64_t n, int64_t l,
                        int dI, int dJ, int dK)
{
    int64_t i0 = dI == 1 ? 0 : m - 1, i1 = dI == 1 ? m : -1;
    int64_t j0 = dJ == 1 ? 0 : n - 1, j1 = dJ == 1 ? n : -1;
    int64_t k0 = dK == 1 ? 0 : l - 1, k1 = dK == 1 ? l : -1;
    for (int64_t i = i0; i != i1; i += dI)
        for (int64_t j = j0; j != j1; j += dJ)
            for (int64_t k = k0; k != k1; k += dK) {
                float uxmin = i == 0 ? U3(i + 1, j, k)
                              : (i == m - 1 ? U3(i - 1, j, k)
                                            : smin_f32(U3(i + 1, j, k), U3(i - 1, j, k)));
                float uymin = j == 0 ? U3(i, j + 1, k)
                              : (j == n - 1 ? U3(i, j - 1, k)
                                            : smin_f32(U3(i, j + 1, k), U3(i, j - 1, k)));
                float uzmin = k == 0 ? U3(i, j, k + 1)
                              : (k == l - 1 ? U3(i, j, k - 1)
                                            : smin_f32(U3(i, j, k + 1), U3(i, j, k - 1)));
                float u_new = (float)solution3d_f64((double)uxmin, (double)uymin, (double)uzmin,
                                                    (double)f[i * n * l + j * l + k], h);
                U3(i, j, k) = smin_f32(u_new, U3(i, j, k));
            }
}
#undef U3

int adt_oracle_eikonal3d_fwd_m32(const float *u0, const float *f, double h, int64_t m,
                                 int64_t n, int64_t l, float tol, int max_iter, int verbose,
                                 float *u, float *uold, int *iters_out, double *err_out)
{
    const int64_t N = m * n * l;
    memcpy(u, u0, sizeof(float) * (size_t)N);
    int it, ran = 0;
    float err = 0;
    for (it = 0; it < max_iter; it++) {
        memcpy(uold, u, sizeof(float) * (size_t)N);
        for (int s = 0; s < 8; s++)
            sweep3d_m32(u, f, h, m, n, l, kDirs_f32[s][0], kDirs_f32[s][1], kDirs_f32[s][2]);
        err = 0.0f;
        for (int64_t p = 0; p < N; p++) err = smax_f32(fabsf(u[p] - uold[p]), err);
        ran = it + 1;
        if (verbose) printf("Iteration %d, Error = %0.6e\n", it, (double)err);
        if (err < tol) break;
    }
    if (iters_out) *iters_out = ran;
    if (err_out) *err_out = (double)err;
    return 0;
}

/* Batched CPU baseline: `nev` independent 3D events with per-event boxes,
 * each solved serially (forward then adjoint) exactly as the reference does
 * per MPI rank.  Events are spread over `nthreads` POSIX threads, one event
 * per thread at a time (mirrors `mpirun --bind-to core -n N`, one
 * single-threaded op per rank, docs/README.md:73). */
#include <pthread.h>

typedef struct {
    const double *const *u0;
    const double *const *f;
    const double *const *g;
    double *const *u;
    double *const *grad_u0;
    double *const *grad_f;
    const int64_t *dims; /* 3 per event */
    const double *h;
    double tol;
    int nev;
    int next;
    int *iters;
    pthread_mutex_t mu;
} batch_ctx;

static void *batch_worker(void *arg)
{
    batch_ctx *c = (batch_ctx *)arg;
    for (;;) {
        pthread_mutex_lock(&c->mu);
        int e = c->next++;
        pthread_mutex_unlock(&c->mu);
        if (e >= c->nev) break;
        int64_t m = c->dims[3 * e], n = c->dims[3 * e + 1], l = c->dims[3 * e + 2];
        int64_t N = m * n * l;
        double *scratch = (double *)malloc(sizeof(double) * (size_t)N);
        int it = 0;
        double err = 0;
        adt_oracle_eikonal3d_fwd_f64(c->u0[e], c->f[e], c->h[e], m, n, l, c->tol,
                                     20, 0, c->u[e], scratch, &it, &err);
        if (c->iters) c->iters[e] = it;
        if (c->g)
            adt_oracle_eikonal3d_bwd_f64(c->g[e], c->u[e], c->u0[e], c->f[e], c->h[e],
                                         m, n, l, c->grad_u0[e], c->grad_f[e], scratch);
        free(scratch);
    }
    return NULL;
}

int adt_oracle_eikonal3d_batch_f64(int nev, const double *const *u0,
                                   const double *const *f, const double *const *g,
                                   const int64_t *dims, const double *h, double tol,
                                   double *const *u, double *const *grad_u0,
                                   double *const *grad_f, int *iters, int nthreads)
{
    batch_ctx c;
    c.u0 = u0;
    c.f = f;
    c.g = g;
    c.u = u;
    c.grad_u0 = grad_u0;
    c.grad_f = grad_f;
    c.dims = dims;
    c.h = h;
    c.tol = tol;
    c.nev = nev;
    c.next = 0;
    c.iters = iters;
    pthread_mutex_init(&c.mu, NULL);
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, batch_worker, &c);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(th);
    pthread_mutex_destroy(&c.mu);
    return 0;
}
