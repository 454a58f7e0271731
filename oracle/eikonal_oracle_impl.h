/*
 * Serial CPU restatement of the ADTomo.jl Eikonal op (TEST INFRASTRUCTURE).
 *
 * This file is included twice by eikonal_oracle.c, once with REAL=double
 * (suffix _f64, the semantics of the reference op) and once with REAL=float
 * (suffix _f32, the semantics this project defines for its FP32 variant: the
 * same expressions evaluated in binary32).  Every expression keeps the
 * operand order of the reference so that, built with -ffp-contract=off, the
 * FP64 results are the reference's results bit for bit.
 *
 * std::min(x, y) in the reference is `(y < x) ? y : x`; std::max(x, y) is
 * `(x < y) ? y : x`.  The argument order is kept at every call site because
 * it decides which operand survives a NaN.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

static inline REAL FN(smin)(REAL x, REAL y) { return (y < x) ? y : x; }
static inline REAL FN(smax)(REAL x, REAL y) { return (x < y) ? y : x; }

/* ---------------------------------------------------------------- 2D ---- */

/* deps/CustomOps/Eikonal/Eikonal.h:14-20 (`solution`). */
static REAL FN(solution2d)(REAL a, REAL b, REAL f, REAL h)
{
    REAL d = FABS(a - b);
    if (d >= f * h)
        return FN(smin)(a, b) + f * h;
    else
        return (a + b + SQRT(C(2.0) * f * f * h * h - (a - b) * (a - b))) / C(2.0);
}

/* deps/CustomOps/Eikonal/Eikonal.h:22-52 (`sweep`): one Gauss-Seidel sweep,
 * i outer, j inner, in the given directions; the source node is skipped.
 * Flat index j*(m+1)+i. */
static void FN(sweep2d)(REAL *u, const REAL *f, int64_t m, int64_t n, REAL h,
                        int64_t ix, int64_t jx, int di, int dj)
{
    const int64_t W = m + 1;
    for (int64_t ii = 0; ii <= m; ii++) {
        int64_t i = di > 0 ? ii : m - ii;
        for (int64_t jj = 0; jj <= n; jj++) {
            int64_t j = dj > 0 ? jj : n - jj;
            if (i == ix && j == jx) continue;
            REAL a, b;
            if (i == 0)
                a = u[j * W + 1];
            else if (i == m)
                a = u[j * W + m - 1];
            else
                a = FN(smin)(u[j * W + i + 1], u[j * W + i - 1]);
            if (j == 0)
                b = u[W + i];
            else if (j == n)
                b = u[(n - 1) * W + i];
            else
                b = FN(smin)(u[(j - 1) * W + i], u[(j + 1) * W + i]);
            REAL u_new = FN(solution2d)(a, b, f[j * W + i], h);
            u[j * W + i] = FN(smin)(u[j * W + i], u_new);
        }
    }
}

/* deps/CustomOps/Eikonal/Eikonal.h:54-93 (`forward`).  Returns 0 when the
 * relative-L2 stop rule fired, 1 when the 100-iteration cap was reached (the
 * reference prints "ERROR: Eikonal does not converge!").  The two Eigen norms
 * are sequential sums of squares here; `*err_out` is the last err computed. */
int FN(adt_oracle_eikonal2d_fwd)(const REAL *f, int64_t m, int64_t n, REAL h,
                                 int64_t ix, int64_t jx, int max_iter, REAL *u,
                                 REAL *uold, int *iters_out, double *err_out)
{
    const int64_t W = m + 1, N = (m + 1) * (n + 1);
    for (int64_t i = 0; i < m + 1; i++)
        for (int64_t j = 0; j < n + 1; j++) {
            u[j * W + i] = C(100000.0);
            if (i == ix && j == jx) u[j * W + i] = C(0.0);
        }
    for (int64_t p = 0; p < N; p++) uold[p] = u[p];
    int converged = 0, it;
    double err = 0.0;
    for (it = 0; it < max_iter; it++) {
        FN(sweep2d)(u, f, m, n, h, ix, jx, +1, +1);
        FN(sweep2d)(u, f, m, n, h, ix, jx, -1, +1);
        FN(sweep2d)(u, f, m, n, h, ix, jx, -1, -1);
        FN(sweep2d)(u, f, m, n, h, ix, jx, +1, -1);
        REAL sd = 0, so = 0;
        for (int64_t p = 0; p < N; p++) {
            REAL d = u[p] - uold[p];
            sd = sd + d * d;
            so = so + uold[p] * uold[p];
        }
        REAL e = SQRT(sd) / SQRT(so);
        err = (double)e;
        if (e < C(1e-8)) {
            converged = 1;
            break;
        }
        for (int64_t p = 0; p < N; p++) uold[p] = u[p];
    }
    if (iters_out) *iters_out = converged ? it + 1 : it;
    if (err_out) *err_out = err;
    return converged ? 0 : 1;
}

/* Comparison used to order nodes by decreasing travel time (ties by index). */
static const REAL *FN(g_sort_u);
static int FN(cmp_desc)(const void *pa, const void *pb)
{
    int64_t a = *(const int64_t *)pa, b = *(const int64_t *)pb;
    REAL ua = FN(g_sort_u)[a], ub = FN(g_sort_u)[b];
    if (ua > ub) return -1;
    if (ua < ub) return 1;
    return (a < b) ? -1 : (a > b);
}

static int64_t *FN(order_desc)(const REAL *u, int64_t N)
{
    int64_t *ord = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N > 0 ? N : 1));
    for (int64_t p = 0; p < N; p++) ord[p] = p;
    FN(g_sort_u) = u;
    qsort(ord, (size_t)N, sizeof(int64_t), FN(cmp_desc));
    return ord;
}

/* deps/CustomOps/Eikonal/Eikonal.h:95-199 (`backward`).
 *
 * The reference assembles the upwind Jacobian A (row r: diagonal
 * sum_d 2(u_r - a_d), off-diagonal 2(a_d - u_r) at the argmin neighbour of
 * each active dimension d; the source row is the identity) and solves
 * A^T lambda = grad_u with SparseLU.  Every off-diagonal entry of row r
 * points to a node with strictly smaller u, so A is triangular in the order
 * of decreasing u and the solve is an exact back-substitution: each lambda_r
 * is pulled from its (already final) downstream neighbours c in the fixed
 * order (i-1, i+1, j-1, j+1):
 *     lambda_r = (g_r + sum_c 2(u_c - u_r) lambda_c) / D_r.
 * The LU solve of the reference differs from this only by rounding.
 * A non-source node without an active dimension makes the reference matrix
 * singular; such a node gets lambda = 0 and is counted in *n_singular. */
int FN(adt_oracle_eikonal2d_bwd)(const REAL *g, const REAL *u, const REAL *f,
                                 int64_t m, int64_t n, REAL h, int64_t ix,
                                 int64_t jx, REAL *grad_f, REAL *lambda,
                                 int64_t *n_singular)
{
    const int64_t W = m + 1, N = (m + 1) * (n + 1);
    const int64_t src = jx * W + ix;
    int64_t *upx = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
    int64_t *upy = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
    REAL *D = (REAL *)malloc(sizeof(REAL) * (size_t)N);
    int64_t nsing = 0;
    for (int64_t j = 0; j < n + 1; j++)
        for (int64_t i = 0; i < m + 1; i++) {
            int64_t idx = j * W + i;
            upx[idx] = -1;
            upy[idx] = -1;
            if (idx == src) {
                D[idx] = C(1.0);
                continue;
            }
            int have = 0;
            REAL diag = 0;
            int64_t nb;
            REAL a;
            if (i == 0) {
                nb = j * W + 1;
                a = u[nb];
            } else if (i == m) {
                nb = j * W + m - 1;
                a = u[nb];
            } else {
                int left = u[j * W + i + 1] > u[j * W + i - 1];
                a = left ? u[j * W + i - 1] : u[j * W + i + 1];
                nb = left ? j * W + i - 1 : j * W + i + 1;
            }
            if (u[idx] > a) {
                diag = C(2.0) * (u[idx] - a);
                have = 1;
                upx[idx] = nb;
            }
            REAL b;
            if (j == 0) {
                nb = W + i;
                b = u[nb];
            } else if (j == n) {
                nb = (n - 1) * W + i;
                b = u[nb];
            } else {
                int up = u[(j + 1) * W + i] > u[(j - 1) * W + i];
                b = up ? u[(j - 1) * W + i] : u[(j + 1) * W + i];
                nb = up ? (j - 1) * W + i : (j + 1) * W + i;
            }
            if (u[idx] > b) {
                REAL t = C(2.0) * (u[idx] - b);
                diag = have ? diag + t : t;
                have = 1;
                upy[idx] = nb;
            }
            D[idx] = diag;
            if (!have) nsing++;
        }
    int64_t *ord = FN(order_desc)(u, N);
    for (int64_t p = 0; p < N; p++) lambda[p] = 0;
    for (int64_t q = 0; q < N; q++) {
        int64_t r = ord[q];
        int64_t i = r % W, j = r / W;
        if (r != src && upx[r] < 0 && upy[r] < 0) {
            lambda[r] = 0;
            continue;
        }
        REAL num = g[r];
        int64_t cs[4] = {i > 0 ? r - 1 : -1, i < m ? r + 1 : -1,
                         j > 0 ? r - W : -1, j < n ? r + W : -1};
        for (int t = 0; t < 4; t++) {
            int64_t c = cs[t];
            if (c < 0 || c == src) continue;
            int64_t upc = t < 2 ? upx[c] : upy[c];
            if (upc != r) continue;
            num = num + (C(2.0) * (u[c] - u[r])) * lambda[c];
        }
        lambda[r] = num / D[r];
    }
    for (int64_t p = 0; p < N; p++) {
        REAL dFdf = C(-2.0) * f[p] * h * h;
        if (p == src) dFdf = C(0.0);
        grad_f[p] = -lambda[p] * dFdf;
    }
    free(ord);
    free(upx);
    free(upy);
    free(D);
    if (n_singular) *n_singular = nsing;
    return 0;
}

/* ---------------------------------------------------------------- 3D ---- */

/* deps/CustomOps/Eikonal3D/Eikonal3D.cpp:11-28 (`calculate_unique_solution`). */
static REAL FN(solution3d)(REAL a1_, REAL a2_, REAL a3_, REAL f, REAL h)
{
    REAL a1 = a1_, a2 = a2_, a3 = a3_, temp;
    if (a1 > a2) { temp = a1; a1 = a2; a2 = temp; }
    if (a1 > a3) { temp = a1; a1 = a3; a3 = temp; }
    if (a2 > a3) { temp = a2; a2 = a3; a3 = temp; }

    REAL x = a1 + f * h;
    if (x <= a2) return x;
    REAL B = -(a1 + a2);
    REAL Cc = (a1 * a1 + a2 * a2 - f * f * h * h) / C(2.0);
    x = (-B + SQRT(B * B - C(4.0) * Cc)) / C(2.0);
    if (x <= a3) return x;
    B = C(-2.0) * (a1 + a2 + a3) / C(3.0);
    Cc = (a1 * a1 + a2 * a2 + a3 * a3 - f * f * h * h) / C(3.0);
    x = (-B + SQRT(B * B - C(4.0) * Cc)) / C(2.0);
    return x;
}

#define U3(i, j, k) u[(i) * n * l + (j) * l + (k)]

/* deps/CustomOps/Eikonal3D/Eikonal3D.cpp:35-57 (`sweeping_over_I_J_K`):
 * lexicographic Gauss-Seidel sweep (i outer, k inner), every node updated. */
static void FN(sweep3d)(REAL *u, const REAL *f, REAL h, int64_t m, int64_t n,
                        int64_t l, int dI, int dJ, int dK)
{
    int64_t i0 = dI == 1 ? 0 : m - 1, i1 = dI == 1 ? m : -1;
    int64_t j0 = dJ == 1 ? 0 : n - 1, j1 = dJ == 1 ? n : -1;
    int64_t k0 = dK == 1 ? 0 : l - 1, k1 = dK == 1 ? l : -1;
    for (int64_t i = i0; i != i1; i += dI)
        for (int64_t j = j0; j != j1; j += dJ)
            for (int64_t k = k0; k != k1; k += dK) {
                REAL uxmin = i == 0 ? U3(i + 1, j, k)
                             : (i == m - 1 ? U3(i - 1, j, k)
                                           : FN(smin)(U3(i + 1, j, k), U3(i - 1, j, k)));
                REAL uymin = j == 0 ? U3(i, j + 1, k)
                             : (j == n - 1 ? U3(i, j - 1, k)
                                           : FN(smin)(U3(i, j + 1, k), U3(i, j - 1, k)));
                REAL uzmin = k == 0 ? U3(i, j, k + 1)
                             : (k == l - 1 ? U3(i, j, k - 1)
                                           : FN(smin)(U3(i, j, k + 1), U3(i, j, k - 1)));
                REAL u_new = FN(solution3d)(uxmin, uymin, uzmin,
                                            f[i * n * l + j * l + k], h);
                U3(i, j, k) = FN(smin)(u_new, U3(i, j, k));
            }
}

/* The 8 sweep orderings of Eikonal3D.cpp:59-68, in their order. */
static const int FN(kDirs)[8][3] = {{1, 1, 1},  {-1, 1, 1},  {-1, -1, 1}, {1, -1, 1},
                                    {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, -1}};

/* Eikonal3D.cpp:71-94 (`solve`/`forward`): u = u0; up to max_iter (20 in the
 * reference) iterations of {snapshot, 8 sweeps, err = max|u - u_old|}; stop
 * when err < tol.  Returns the number of iterations run in *iters_out and
 * the last err in *err_out.  `uold` is caller scratch of N elements. */
int FN(adt_oracle_eikonal3d_fwd)(const REAL *u0, const REAL *f, REAL h,
                                 int64_t m, int64_t n, int64_t l, REAL tol,
                                 int max_iter, int verbose, REAL *u, REAL *uold,
                                 int *iters_out, double *err_out)
{
    const int64_t N = m * n * l;
    memcpy(u, u0, sizeof(REAL) * (size_t)N);
    int it, ran = 0;
    REAL err = 0;
    for (it = 0; it < max_iter; it++) {
        memcpy(uold, u, sizeof(REAL) * (size_t)N);
        for (int s = 0; s < 8; s++)
            FN(sweep3d)(u, f, h, m, n, l, FN(kDirs)[s][0], FN(kDirs)[s][1],
                        FN(kDirs)[s][2]);
        err = C(0.0);
        for (int64_t p = 0; p < N; p++) err = FN(smax)(FABS(u[p] - uold[p]), err);
        ran = it + 1;
        if (verbose) printf("Iteration %d, Error = %0.6e\n", it, (double)err);
        if (err < tol) break;
    }
    if (iters_out) *iters_out = ran;
    if (err_out) *err_out = (double)err;
    return 0;
}

/* Eikonal3D.cpp:96-198 (`backward`).
 *
 * Z = {r : u_r == u0_r} U {r : no dimension with u_r > umin_d(r)}; the
 * reference zeroes the rows and columns of Z, puts 1 on their diagonal and
 * zeroes g on Z, so lambda_Z = 0.  For r outside Z the transposed system is
 * back-substituted in decreasing u, pulling from the neighbours c (order
 * i-1, i+1, j-1, j+1, k-1, k+1) whose argmin neighbour along the connecting
 * dimension is r:
 *     lambda_r = (g_r + sum_c 2(u_c - u_r) lambda_c) / D_r,
 *     D_r = sum over active d (in x, y, z order) of 2(u_r - umin_d(r)).
 * grad_f = -lambda * (-2 f h h); grad_u0 = g where u == u0, else 0. */
int FN(adt_oracle_eikonal3d_bwd)(const REAL *g, const REAL *u, const REAL *u0,
                                 const REAL *f, REAL h, int64_t m, int64_t n,
                                 int64_t l, REAL *grad_u0, REAL *grad_f,
                                 REAL *lambda)
{
    const int64_t N = m * n * l;
    int64_t *up = (int64_t *)malloc(sizeof(int64_t) * 3 * (size_t)N);
    REAL *D = (REAL *)malloc(sizeof(REAL) * (size_t)N);
    unsigned char *Z = (unsigned char *)malloc((size_t)N);
    for (int64_t p = 0; p < N; p++) grad_u0[p] = (u[p] == u0[p]) ? g[p] : C(0.0);
    for (int64_t i = 0; i < m; i++)
        for (int64_t j = 0; j < n; j++)
            for (int64_t k = 0; k < l; k++) {
                int64_t r = i * n * l + j * l + k;
                up[3 * r] = up[3 * r + 1] = up[3 * r + 2] = -1;
                D[r] = 0;
                if (u[r] == u0[r]) {
                    Z[r] = 1;
                    continue;
                }
                REAL mins[3];
                int64_t ids[3];
                mins[0] = i == 0 ? U3(i + 1, j, k)
                          : (i == m - 1 ? U3(i - 1, j, k)
                                        : FN(smin)(U3(i + 1, j, k), U3(i - 1, j, k)));
                mins[1] = j == 0 ? U3(i, j + 1, k)
                          : (j == n - 1 ? U3(i, j - 1, k)
                                        : FN(smin)(U3(i, j + 1, k), U3(i, j - 1, k)));
                mins[2] = k == 0 ? U3(i, j, k + 1)
                          : (k == l - 1 ? U3(i, j, k - 1)
                                        : FN(smin)(U3(i, j, k + 1), U3(i, j, k - 1)));
                ids[0] = i == 0 ? r + n * l
                         : (i == m - 1 ? r - n * l
                                       : (U3(i + 1, j, k) > U3(i - 1, j, k) ? r - n * l : r + n * l));
                ids[1] = j == 0 ? r + l
                         : (j == n - 1 ? r - l
                                       : (U3(i, j + 1, k) > U3(i, j - 1, k) ? r - l : r + l));
                ids[2] = k == 0 ? r + 1
                         : (k == l - 1 ? r - 1
                                       : (U3(i, j, k + 1) > U3(i, j, k - 1) ? r - 1 : r + 1));
                int have = 0;
                REAL diag = 0;
                for (int d = 0; d < 3; d++) {
                    if (u[r] > mins[d]) {
                        REAL t = C(2.0) * (u[r] - mins[d]);
                        diag = have ? diag + t : t;
                        have = 1;
                        up[3 * r + d] = ids[d];
                    }
                }
                Z[r] = have ? 0 : 1;
                D[r] = diag;
            }
    int64_t *ord = FN(order_desc)(u, N);
    for (int64_t p = 0; p < N; p++) lambda[p] = 0;
    const int64_t strides[3] = {n * l, l, 1};
    const int64_t dims[3] = {m, n, l};
    for (int64_t q = 0; q < N; q++) {
        int64_t r = ord[q];
        if (Z[r]) {
            lambda[r] = 0;
            continue;
        }
        int64_t coord[3] = {r / (n * l), (r / l) % n, r % l};
        REAL num = g[r];
        for (int d = 0; d < 3; d++) {
            for (int side = -1; side <= 1; side += 2) {
                int64_t cc = coord[d] + side;
                if (cc < 0 || cc >= dims[d]) continue;
                int64_t c = r + side * strides[d];
                if (Z[c] || up[3 * c + d] != r) continue;
                num = num + (C(2.0) * (u[c] - u[r])) * lambda[c];
            }
        }
        lambda[r] = num / D[r];
    }
    for (int64_t p = 0; p < N; p++) {
        REAL rhs = C(-2.0) * f[p] * h * h;
        grad_f[p] = -lambda[p] * rhs;
    }
    free(ord);
    free(up);
    free(D);
    free(Z);
    return 0;
}

#undef U3
#undef FN
#undef CAT
#undef CAT2
