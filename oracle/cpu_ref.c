/*
 * cpu_ref.c - C/OpenMP restatement of the reference CPU propagator core.
 *
 * Test oracle and timed CPU baseline only; the product never links it.
 *
 * It performs, for one simulate() call, the numeric loop of
 * superspinsim/generate_simulator.py (CPU branches, numba prange loops):
 *
 *   for r in 0..F-1                                  (:1954)
 *     A[i*J+j] = h sum_g (sum_s w[j,s] c[r,i,s,g]) G_g   (:278-448)
 *       or, rotating, sum_s sum_g (Gt[s,g] c) (h w[j,s])  (:450-532)
 *     A /= 4^R; (+I if not residual)                  (:534-578)
 *     2R squarings X <- 2X + X X  (or X X)            (:706-724, 882-897)
 *     TE[i] <- E_0 (E_1 (... E_{J-1} TE[i]))          (:1259-1294)
 *     rotating: TE[i] <- exp(L0 h) TE[i]              (:945-967, GPU bounds)
 *   prefix: TE[i+1] <- TE[i+1] TE[i]                  (:1308-1321)
 *
 * Matrices are n x n, row-major.  With the residual flag every matrix holds
 * (M - I) and products are L + R + L R.  The coefficient table is sampled by
 * the caller at the reference's node times.  Parallelism is over the
 * N*J independent matrices, like the reference's outer prange.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

enum { CPUREF_RESIDUAL = 1, CPUREF_ROTATING = 2 };

/* out = L R (+ L + R when residual); out must not alias L or R. */
static void mul(int n, const double *L, const double *Rm, double *out,
                int residual)
{
    if (residual) {
        for (int i = 0; i < n * n; ++i)
            out[i] = L[i] + Rm[i];
    } else {
        memset(out, 0, sizeof(double) * (size_t)n * n);
    }
    for (int y = 0; y < n; ++y) {
        double *orow = out + (size_t)y * n;
        const double *lrow = L + (size_t)y * n;
        for (int k = 0; k < n; ++k) {
            const double a = lrow[k];
            const double *rrow = Rm + (size_t)k * n;
            for (int x = 0; x < n; ++x)
                orow[x] += a * rrow[x];
        }
    }
}

/* out = 2X + X X (residual) or X X. */
static void square(int n, const double *X, double *out, int residual)
{
    if (residual) {
        for (int i = 0; i < n * n; ++i)
            out[i] = 2.0 * X[i];
    } else {
        memset(out, 0, sizeof(double) * (size_t)n * n);
    }
    for (int y = 0; y < n; ++y) {
        double *orow = out + (size_t)y * n;
        const double *xrow = X + (size_t)y * n;
        for (int k = 0; k < n; ++k) {
            const double a = xrow[k];
            const double *krow = X + (size_t)k * n;
            for (int x = 0; x < n; ++x)
                orow[x] += a * krow[x];
        }
    }
}

static void eig_rows(int n, double *te, double *tmp, int nd,
                     const double *dstep, int ns, const double *sstep,
                     int residual)
{
    memcpy(tmp, te, sizeof(double) * (size_t)n * n);
    for (int y = 0; y < nd; ++y) {
        const double d0 = dstep[2 * y], d1 = dstep[2 * y + 1];
        const double *r = tmp + (size_t)(2 * y) * n;
        const double *im = tmp + (size_t)(2 * y + 1) * n;
        double *outr = te + (size_t)(2 * y) * n;
        double *outi = te + (size_t)(2 * y + 1) * n;
        for (int x = 0; x < n; ++x) {
            double sr, si;
            if (residual) {
                sr = (1 + d0) * r[x] - d1 * im[x];
                si = d1 * r[x] + (1 + d0) * im[x];
                if (x == 2 * y) { sr += d0; si += d1; }
                if (x == 2 * y + 1) { sr -= d1; si += d0; }
            } else {
                sr = d0 * r[x] - d1 * im[x];
                si = d1 * r[x] + d0 * im[x];
            }
            outr[x] = sr;
            outi[x] = si;
        }
    }
    for (int y = 0; y < ns; ++y) {
        const int row = 2 * nd + y;
        const double s = sstep[y];
        double *o = te + (size_t)row * n;
        const double *in = tmp + (size_t)row * n;
        for (int x = 0; x < n; ++x) {
            double v = residual ? (1 + s) * in[x] : s * in[x];
            if (residual && x == row)
                v += s;
            o[x] = v;
        }
    }
}

/*
 * gens:    non-rotating (G, n, n); rotating (S, G, n, n)
 * weights: (J, S)
 * coef:    (F, N, S, G) samples at the reference node times
 * te:      out (N, n, n)
 * Returns 0, or -1 if scratch allocation fails.
 */
int cpuref_propagate(int n, int G, int J, int S, int F, int N, int R,
                     int flags, const double *gens, const double *weights,
                     const double *coef, double h, int nd,
                     const double *dstep, int ns, const double *sstep,
                     double *te, int threads)
{
    const int residual = flags & CPUREF_RESIDUAL;
    const int rotating = flags & CPUREF_ROTATING;
    const size_t nn = (size_t)n * n;
    const size_t batch = (size_t)N * J;
    const double scale = pow(4.0, R);
    double *A = malloc(sizeof(double) * nn * batch);
    double *B = malloc(sizeof(double) * nn * batch);
    double *T = malloc(sizeof(double) * nn * (size_t)N);
    if (!A || !B || !T) {
        free(A); free(B); free(T);
        return -1;
    }
#ifdef _OPENMP
    if (threads > 0)
        omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    for (int i = 0; i < N; ++i) {
        double *t = te + nn * i;
        if (residual) {
            memset(t, 0, sizeof(double) * nn);
        } else {
            memset(t, 0, sizeof(double) * nn);
            for (int d = 0; d < n; ++d)
                t[(size_t)d * n + d] = 1.0;
        }
    }

    for (int r = 0; r < F; ++r) {
        const double *cr = coef + (size_t)r * N * S * G;
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)batch; ++k) {
            const int i = (int)(k / J), j = (int)(k % J);
            double *a = A + nn * k;
            const double *c = cr + (size_t)i * S * G;
            if (rotating) {
                memset(a, 0, sizeof(double) * nn);
                for (int s = 0; s < S; ++s) {
                    const double hw = h * weights[j * S + s];
                    for (int g = 0; g < G; ++g) {
                        const double cg = c[s * G + g];
                        const double *gm = gens + ((size_t)s * G + g) * nn;
                        for (size_t e = 0; e < nn; ++e)
                            a[e] += (gm[e] * cg) * hw;
                    }
                }
            } else {
                double cw[64];
                for (int g = 0; g < G; ++g) {
                    double acc = 0.0;
                    for (int s = 0; s < S; ++s)
                        acc += weights[j * S + s] * c[s * G + g];
                    cw[g] = acc;
                }
                memset(a, 0, sizeof(double) * nn);
                for (int g = 0; g < G; ++g) {
                    const double *gm = gens + (size_t)g * nn;
                    for (size_t e = 0; e < nn; ++e)
                        a[e] += cw[g] * gm[e];
                }
                for (size_t e = 0; e < nn; ++e)
                    a[e] *= h;
            }
            for (size_t e = 0; e < nn; ++e)
                a[e] /= scale;
            if (!residual)
                for (int d = 0; d < n; ++d)
                    a[(size_t)d * n + d] += 1.0;
            double *x = a, *y = B + nn * k;
            for (int q = 0; q < 2 * R; ++q) {
                square(n, x, y, residual);
                double *tmp = x; x = y; y = tmp;
            }
            if (x != a)
                memcpy(a, x, sizeof(double) * nn);
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            double *t = te + nn * i;
            double *w = T + nn * i;
            for (int j = J - 1; j >= 0; --j) {
                mul(n, A + nn * ((size_t)i * J + j), t, w, residual);
                memcpy(t, w, sizeof(double) * nn);
            }
            if (rotating)
                eig_rows(n, t, w, nd, dstep, ns, sstep, residual);
        }
    }
    for (int i = 0; i + 1 < N; ++i) {
        mul(n, te + nn * (i + 1), te + nn * i, T, residual);
        memcpy(te + nn * (i + 1), T, sizeof(double) * nn);
    }
    free(A);
    free(B);
    free(T);
    return 0;
}
