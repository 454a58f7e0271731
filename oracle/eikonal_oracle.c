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
 * The _f32 twins evaluate the same expressions in binary32.
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
