/*
 * C restatement of minirecast's FDK/FBP path in float64 (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
 * libraries built from this file, and only as the checker or as the timed CPU
 * baseline.  Two builds (oracle/Makefile):
 *   liboracle_exact.so  -O2 -ffp-contract=off: float64 oracle for large workloads,
 *                       pinned against oracle/oracle.py (tests/test_oracle_c.py);
 *   liboracle_fast.so   -O3 -ffast-math: the CPU baseline, a restatement of the
 *                       reference's Numba kernels (njit fastmath, prange over points,
 *                       serial over views, float64 accumulation over float32 q).
 *
 * Functions and the reference code they restate:
 *   orc_ramp_response     recon.py:34-55
 *   orc_filter_rows       recon.py:58-64 (float64 throughout)
 *   orc_fdk_filter_view   recon.py:85-104 (one view of a [R][k][C] block)
 *   orc_bp_cone           recon.py:193-219 (bp_cone_numba, with _bil_nb :156-173)
 *   orc_bp_parallel       recon.py:175-191 (bp_parallel_numba)
 *   orc_bp_cone_acc_view  one view of recon.py:205-217 over float64 q, accumulated
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int fft_length(int n) {
    int N = 1;
    while (N < 2 * n) N <<= 1;
    return N;
}

/* In-place radix-2 complex FFT (re, im interleaved), forward sign. */
static void fft64(double* a, int N, const double* tw /* N/2 (cos, sin) pairs */) {
    for (int i = 1, j = 0; i < N; ++i) {
        int bit = N >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = a[2 * i]; a[2 * i] = a[2 * j]; a[2 * j] = t;
            t = a[2 * i + 1]; a[2 * i + 1] = a[2 * j + 1]; a[2 * j + 1] = t;
        }
    }
    for (int len = 2; len <= N; len <<= 1) {
        const int half = len / 2, stride = N / len;
        for (int i = 0; i < N; i += len) {
            for (int k = 0; k < half; ++k) {
                const double wr = tw[2 * (k * stride)], wi = tw[2 * (k * stride) + 1];
                double* u = a + 2 * (i + k);
                double* v = a + 2 * (i + k + half);
                const double vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
                v[0] = u[0] - vr; v[1] = u[1] - vi;
                u[0] += vr; u[1] += vi;
            }
        }
    }
}

static double* twiddles(int N) {
    double* tw = (double*)malloc(sizeof(double) * N);
    for (int k = 0; k < N / 2; ++k) {
        const double ang = -2.0 * M_PI * (double)k / (double)N;
        tw[2 * k] = cos(ang);
        tw[2 * k + 1] = sin(ang);
    }
    return tw;
}

static double window_at(int window, double f) {
    switch (window) {
    case 0: return 1.0;
    case 1: { const double y = M_PI * (f == 0.0 ? 1e-20 : f); return sin(y) / y; }
    case 2: return 0.5 * (1.0 + cos(2.0 * M_PI * f));
    case 3: return cos(M_PI * f);
    default: return 0.0;
    }
}

/* H[N] of recon.py:34-55; returns N (or -1 for a bad window). */
int orc_ramp_response(int n, double tau, int window, double* H) {
    if (window < 0 || window > 3 || n < 1) return -1;
    const int N = fft_length(n);
    double* a = (double*)calloc(2 * (size_t)N, sizeof(double));
    a[0] = 1.0 / (4.0 * tau * tau);
    for (int i = 1; i < N; ++i) {
        const int idx = i < N / 2 ? i : i - N;
        if (idx % 2 != 0) {
            const double t = M_PI * idx * tau;
            a[2 * i] = -1.0 / (t * t);
        }
    }
    double* tw = twiddles(N);
    fft64(a, N, tw);
    for (int k = 0; k < N; ++k) {
        const int kk = k < N / 2 ? k : k - N;
        H[k] = a[2 * k] * tau * window_at(window, kk * (1.0 / N));
    }
    free(tw);
    free(a);
    return N;
}

/* Filters one row x[n] (float64 in) with real H[N] into y[n]; work[2N]. */
static void filter_one(const double* x, int n, const double* H, int N, const double* tw,
                       double* work, double* y) {
    memset(work, 0, sizeof(double) * 2 * (size_t)N);
    for (int i = 0; i < n; ++i) work[2 * i] = x[i];
    fft64(work, N, tw);
    for (int k = 0; k < N; ++k) {           /* conj(X) H: inverse through a forward FFT */
        work[2 * k] *= H[k];
        work[2 * k + 1] *= -H[k];
    }
    fft64(work, N, tw);
    for (int i = 0; i < n; ++i) y[i] = work[2 * i] / N;
}

/* rows x n float32 rows -> float64, ramp-filtered with pitch tau (recon.py:58-64). */
int orc_filter_rows(const float* in, int rows, int n, double tau, int window, double* out) {
    const int Nq = fft_length(n);
    double* H = (double*)malloc(sizeof(double) * Nq);
    const int N = orc_ramp_response(n, tau, window, H);
    if (N < 0) { free(H); return -1; }
    double* tw = twiddles(N);
#pragma omp parallel
    {
        double* work = (double*)malloc(sizeof(double) * 2 * (size_t)N);
        double* x = (double*)malloc(sizeof(double) * n);
#pragma omp for schedule(static)
        for (int r = 0; r < rows; ++r) {
            for (int i = 0; i < n; ++i) x[i] = in[(size_t)r * n + i];
            filter_one(x, n, H, N, tw, work, out + (size_t)r * n);
        }
        free(work);
        free(x);
    }
    free(tw);
    free(H);
    return 0;
}

static void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* cone constants of one cone_vec row (recon.py:70-82) */
typedef struct { double s[3], d[3], u[3], v[3], uu[3], vv[3], n[3], sdd, sod; } cone_k;

static void cone_const(const double* V, cone_k* K) {
    for (int i = 0; i < 3; ++i) { K->s[i] = V[i]; K->d[i] = V[3 + i]; K->u[i] = V[6 + i]; K->v[i] = V[9 + i]; }
    cross3(K->u, K->v, K->n);
    const double nn = sqrt(dot3(K->n, K->n));
    for (int i = 0; i < 3; ++i) K->n[i] /= nn;
    double ds[3] = {K->d[0] - K->s[0], K->d[1] - K->s[1], K->d[2] - K->s[2]};
    const double sg = dot3(ds, K->n);
    const double sign = sg > 0 ? 1.0 : (sg < 0 ? -1.0 : 0.0);
    for (int i = 0; i < 3; ++i) K->n[i] *= sign;
    K->sdd = dot3(ds, K->n);
    double ms[3] = {-K->s[0], -K->s[1], -K->s[2]};
    K->sod = dot3(ms, K->n);
    const double u2 = dot3(K->u, K->u), v2 = dot3(K->v, K->v);
    for (int i = 0; i < 3; ++i) { K->uu[i] = K->u[i] / u2; K->vv[i] = K->v[i] / v2; }
}

/* FDK weight + ramp filter of view v of block [R][kb][C] with cone_vec row V
 * -> out[R][C] float64 (recon.py:85-104). */
int orc_fdk_filter_view(const float* block, int R, int kb, int C, int v, const double* V,
                        int window, double* out) {
    cone_k K;
    cone_const(V, &K);
    double p0[3], off[3];
    for (int i = 0; i < 3; ++i) { p0[i] = K.s[i] + K.sdd * K.n[i]; off[i] = K.d[i] - p0[i]; }
    const double un = sqrt(dot3(K.u, K.u)), vn = sqrt(dot3(K.v, K.v));
    const double ou = dot3(off, K.u) / un, ov = dot3(off, K.v) / vn;
    const double tau = un * K.sod / K.sdd;
    double* H = (double*)malloc(sizeof(double) * fft_length(C));
    const int N = orc_ramp_response(C, tau, window, H);
    if (N < 0) { free(H); return -1; }
    double* tw = twiddles(N);
#pragma omp parallel
    {
        double* work = (double*)malloc(sizeof(double) * 2 * (size_t)N);
        double* x = (double*)malloc(sizeof(double) * C);
#pragma omp for schedule(static)
        for (int r = 0; r < R; ++r) {
            const double b = ov + (r - (R - 1) / 2.0) * vn;
            const float* src = block + ((size_t)r * kb + v) * C;
            for (int c = 0; c < C; ++c) {
                const double a = ou + (c - (C - 1) / 2.0) * un;
                const double w = K.sdd / sqrt(K.sdd * K.sdd + a * a + b * b);
                x[c] = (double)src[c] * w;
            }
            filter_one(x, C, H, N, tw, work, out + (size_t)r * C);
        }
        free(work);
        free(x);
    }
    free(tw);
    free(H);
    return 0;
}

/* _bil_nb (recon.py:156-173) over float32 or float64 images */
#define BILINEAR(T)                                                                    \
    static inline double bil_##T(const T* img, double r, double c, int R, int C) {    \
        const double rf = floor(r), cf = floor(c);                                     \
        const int r0 = (int)rf, c0 = (int)cf;                                          \
        const double fr = r - r0, fc = c - c0;                                         \
        double v = 0.0;                                                                \
        if (0 <= r0 && r0 < R) {                                                       \
            if (0 <= c0 && c0 < C) v += (1.0 - fr) * (1.0 - fc) * img[(size_t)r0 * C + c0]; \
            if (0 <= c0 + 1 && c0 + 1 < C) v += (1.0 - fr) * fc * img[(size_t)r0 * C + c0 + 1]; \
        }                                                                              \
        if (0 <= r0 + 1 && r0 + 1 < R) {                                               \
            if (0 <= c0 && c0 < C) v += fr * (1.0 - fc) * img[(size_t)(r0 + 1) * C + c0]; \
            if (0 <= c0 + 1 && c0 + 1 < C) v += fr * fc * img[(size_t)(r0 + 1) * C + c0 + 1]; \
        }                                                                              \
        return v;                                                                      \
    }
BILINEAR(float)
BILINEAR(double)

/* bp_cone_numba (recon.py:193-219): q float32 [A][R][C]; S, D, UU, VV, N [A][3];
 * SDD, SOD [A]; pts [M][3]; out [M] float64.  Points outer (parallel), views inner. */
void orc_bp_cone(const float* q, int A, int R, int C, const double* S, const double* D,
                 const double* UU, const double* VV, const double* Nv, const double* SDD,
                 const double* SOD, const double* pts, int64_t M, double scale, double* out) {
    const double cr = (R - 1) / 2.0, cc = (C - 1) / 2.0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < M; ++i) {
        const double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
        double acc = 0.0;
        for (int k = 0; k < A; ++k) {
            const double rx = px - S[3 * k], ry = py - S[3 * k + 1], rz = pz - S[3 * k + 2];
            const double L = rx * Nv[3 * k] + ry * Nv[3 * k + 1] + rz * Nv[3 * k + 2];
            const double t = SDD[k] / L;
            const double X0 = S[3 * k] + t * rx - D[3 * k];
            const double X1 = S[3 * k + 1] + t * ry - D[3 * k + 1];
            const double X2 = S[3 * k + 2] + t * rz - D[3 * k + 2];
            const double col = X0 * UU[3 * k] + X1 * UU[3 * k + 1] + X2 * UU[3 * k + 2] + cc;
            const double row = X0 * VV[3 * k] + X1 * VV[3 * k + 1] + X2 * VV[3 * k + 2] + cr;
            const double w = SOD[k] / L;
            acc += w * w * bil_float(q + (size_t)k * R * C, row, col, R, C);
        }
        out[i] = acc * scale;
    }
}

/* bp_parallel_numba (recon.py:175-191) */
void orc_bp_parallel(const float* q, int A, int R, int C, const double* cos_t,
                     const double* sin_t, double pu, double pv, const double* pts, int64_t M,
                     double scale, double col_shift, double* out) {
    const double cr = (R - 1) / 2.0, cc = (C - 1) / 2.0 + col_shift;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < M; ++i) {
        const double x = pts[3 * i], y = pts[3 * i + 1];
        const double row = pts[3 * i + 2] / pv + cr;
        double acc = 0.0;
        for (int k = 0; k < A; ++k) {
            const double col = (-x * sin_t[k] + y * cos_t[k]) / pu + cc;
            acc += bil_float(q + (size_t)k * R * C, row, col, R, C);
        }
        out[i] = acc * scale;
    }
}

/* acc[i] += (SOD/L)^2 * bilinear(q64, row, col) for one cone view (float64 q), with the
 * detector moved to D - shift*u (recon.py:420). */
void orc_bp_cone_acc_view(const double* q, int R, int C, const double* V, double shift,
                          const double* pts, int64_t M, double* acc) {
    cone_k K;
    cone_const(V, &K);
    double D[3];
    for (int i = 0; i < 3; ++i) D[i] = K.d[i] - shift * K.u[i];
    const double cr = (R - 1) / 2.0, cc = (C - 1) / 2.0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < M; ++i) {
        const double rx = pts[3 * i] - K.s[0], ry = pts[3 * i + 1] - K.s[1], rz = pts[3 * i + 2] - K.s[2];
        const double L = rx * K.n[0] + ry * K.n[1] + rz * K.n[2];
        const double t = K.sdd / L;
        const double X0 = K.s[0] + t * rx - D[0];
        const double X1 = K.s[1] + t * ry - D[1];
        const double X2 = K.s[2] + t * rz - D[2];
        const double col = X0 * K.uu[0] + X1 * K.uu[1] + X2 * K.uu[2] + cc;
        const double row = X0 * K.vv[0] + X1 * K.vv[1] + X2 * K.vv[2] + cr;
        const double w = K.sod / L;
        acc[i] += w * w * bil_double(q, row, col, R, C);
    }
}

/* bp_parallel_numba (recon.py:175-191) on float64 filtered views q, element (view k, row
 * r, col c) at q[k * s_view + r * s_row + c]; acc[i] += the sum over the A views (no
 * scale), with the column shift col_shift (recon.py:181). */
void orc_bp_parallel_acc(const double* q, int A, int R, int C, int64_t s_view, int64_t s_row,
                         const double* cos_t, const double* sin_t, double pu, double pv,
                         double col_shift, const double* pts, int64_t M, double* acc) {
    const double cr = (R - 1) / 2.0, cc = (C - 1) / 2.0 + col_shift;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < M; ++i) {
        const double x = pts[3 * i], y = pts[3 * i + 1];
        const double row = pts[3 * i + 2] / pv + cr;
        const int r0 = (int)floor(row);
        const double fr = row - r0;
        double sum = 0.0;
        for (int k = 0; k < A; ++k) {
            const double col = (-x * sin_t[k] + y * cos_t[k]) / pu + cc;
            const int c0 = (int)floor(col);
            const double fc = col - c0;
            const double* qk = q + (int64_t)k * s_view;
            double v = 0.0;
            for (int dr = 0; dr < 2; ++dr) {
                const int ri = r0 + dr;
                if (ri < 0 || ri >= R) continue;
                const double wr = dr ? fr : 1.0 - fr;
                for (int dc = 0; dc < 2; ++dc) {
                    const int ci = c0 + dc;
                    if (ci < 0 || ci >= C) continue;
                    v += wr * (dc ? fc : 1.0 - fc) * qk[(int64_t)ri * s_row + ci];
                }
            }
            sum += v;
        }
        acc[i] += sum;
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
