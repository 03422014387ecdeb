/* Runs the C oracle (oracle/oracle_c.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer on small inputs, including points whose taps leave the
 * detector on every side, single-column rows and a short scan, and checks a few
 * identities between its entry points.  Exit status 0: no sanitizer report, every
 * check held. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int orc_ramp_response(int n, double tau, int window, double* H);
int orc_filter_rows(const float* in, int rows, int n, double tau, int window, double* out);
int orc_fdk_filter_view(const float* block, int R, int kb, int C, int v, const double* V,
                        int window, double* out);
void orc_bp_cone(const float* q, int A, int R, int C, const double* S, const double* D,
                 const double* UU, const double* VV, const double* Nv, const double* SDD,
                 const double* SOD, const double* pts, int64_t M, double scale, double* out);
void orc_bp_parallel(const float* q, int A, int R, int C, const double* cos_t,
                     const double* sin_t, double pu, double pv, const double* pts, int64_t M,
                     double scale, double col_shift, double* out);
void orc_bp_cone_acc_view(const double* q, int R, int C, const double* V, double shift,
                          const double* pts, int64_t M, double* acc);
void orc_bp_parallel_acc(const double* q, int A, int R, int C, int64_t s_view, int64_t s_row,
                         const double* cos_t, const double* sin_t, double pu, double pv,
                         double col_shift, const double* pts, int64_t M, double* acc);

static double urand(unsigned* s) { *s = *s * 1664525u + 1013904223u; return (*s >> 8) / 16777216.0; }

int main(void) {
    unsigned seed = 7;
    int failures = 0;
    /* ramp responses, every window, odd and tiny lengths */
    for (int w = 0; w < 4; ++w)
        for (int n = 1; n <= 33; n += 8) {
            double H[128];
            const int N = orc_ramp_response(n, 0.7, w, H);
            if (N < 2 * n || N > 128) { ++failures; fprintf(stderr, "FAIL ramp N=%d n=%d\n", N, n); }
        }
    if (orc_ramp_response(5, 1.0, 9, NULL) >= 0) { ++failures; fprintf(stderr, "FAIL bad window\n"); }
    /* filtering: one column, and a block of rows */
    enum { R = 5, K = 3, C = 7, A = K, M = 200 };
    float blk[R * K * C];
    for (int i = 0; i < R * K * C; ++i) blk[i] = (float)(urand(&seed) * 2.0 - 0.5);
    double f1[R];
    float one[R];
    for (int i = 0; i < R; ++i) one[i] = blk[i];
    if (orc_filter_rows(one, R, 1, 0.5, 1, f1)) { ++failures; fprintf(stderr, "FAIL n=1\n"); }
    double frows[R * K * C];
    if (orc_filter_rows(blk, R * K, C, 0.5, 2, frows)) { ++failures; fprintf(stderr, "FAIL rows\n"); }
    /* a short cone scan, points inside and far outside the field of view */
    double V[A * 12], S[A * 3], D[A * 3], UU[A * 3], VV[A * 3], Nv[A * 3], SDD[A], SOD[A];
    for (int k = 0; k < A; ++k) {
        const double t = 2.0 * M_PI * k / A, ct = cos(t), st = sin(t);
        double* v = V + 12 * k;
        v[0] = 40.0 * st; v[1] = -40.0 * ct; v[2] = 0.0;             /* source */
        v[3] = -30.0 * st; v[4] = 30.0 * ct; v[5] = 0.0;              /* detector centre */
        v[6] = 1.1 * ct; v[7] = 1.1 * st; v[8] = 0.0;                 /* u */
        v[9] = 0.0; v[10] = 0.0; v[11] = 1.1;                         /* v */
        for (int d = 0; d < 3; ++d) {
            S[3 * k + d] = v[d];
            D[3 * k + d] = v[3 + d];
            UU[3 * k + d] = v[6 + d] / (1.1 * 1.1);
            VV[3 * k + d] = v[9 + d] / (1.1 * 1.1);
        }
        Nv[3 * k] = -st; Nv[3 * k + 1] = ct; Nv[3 * k + 2] = 0.0;
        SDD[k] = 70.0; SOD[k] = 40.0;
    }
    double q64[A * R * C];
    float q32[A * R * C];
    for (int k = 0; k < A; ++k) {
        if (orc_fdk_filter_view(blk, R, K, C, k, V + 12 * k, 0, q64 + k * R * C)) {
            ++failures; fprintf(stderr, "FAIL fdk view %d\n", k);
        }
    }
    for (int i = 0; i < A * R * C; ++i) q32[i] = (float)q64[i];
    double pts[M * 3], out[M], acc[M], out2[M], acc2[M];
    for (int i = 0; i < M * 3; ++i) pts[i] = (urand(&seed) * 2.0 - 1.0) * (i % 3 == 2 ? 12.0 : 30.0);
    orc_bp_cone(q32, A, R, C, S, D, UU, VV, Nv, SDD, SOD, pts, M, 1.0, out);
    for (int i = 0; i < M; ++i) acc[i] = 0.0;
    for (int k = 0; k < A; ++k) {
        double qk[R * C];
        for (int j = 0; j < R * C; ++j) qk[j] = q32[k * R * C + j];
        orc_bp_cone_acc_view(qk, R, C, V + 12 * k, 0.0, pts, M, acc);
    }
    for (int i = 0; i < M; ++i)
        if (fabs(acc[i] - out[i]) > 1e-9 * (1.0 + fabs(out[i]))) {
            ++failures; fprintf(stderr, "FAIL cone acc vs bp at %d\n", i); break;
        }
    /* parallel beam: float32 entry point against the strided float64 accumulation */
    double ct[A], st[A], qd[A * R * C];
    for (int k = 0; k < A; ++k) { ct[k] = cos(0.9 * k); st[k] = sin(0.9 * k); }
    for (int i = 0; i < A * R * C; ++i) qd[i] = q32[i];
    orc_bp_parallel(q32, A, R, C, ct, st, 1.3, 2.1, pts, M, 1.0, 0.75, out2);
    for (int i = 0; i < M; ++i) acc2[i] = 0.0;
    orc_bp_parallel_acc(qd, A, R, C, (int64_t)R * C, C, ct, st, 1.3, 2.1, 0.75, pts, M, acc2);
    for (int i = 0; i < M; ++i)
        if (fabs(acc2[i] - out2[i]) > 1e-9 * (1.0 + fabs(out2[i]))) {
            ++failures; fprintf(stderr, "FAIL parallel acc vs bp at %d\n", i); break;
        }
    if (failures) fprintf(stderr, "%d check(s) failed\n", failures);
    else printf("oracle_check: all checks passed\n");
    return failures ? 1 : 0;
}
