// Synthetic acquisition: exact line integrals of constant-density ellipsoids
// (phantom.py:56-98), one ray per detector pixel, float64 per ray, stored float32
// in the reference's [rows][angles][cols] projection layout.
#include "mr_internal.h"

namespace mr {

namespace {

// ellipsoid record: value, centre[3], M[9] (row-major, x - centre -> unit sphere)
constexpr int ELL = 13;

__global__ __launch_bounds__(256) void phantom_kernel(const double* __restrict__ ell, int K,
                                                      int kind, const double* __restrict__ geo,
                                                      int A, int R, int C, double pu, double pv,
                                                      double far, float* __restrict__ out) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    const int k = blockIdx.z;
    if (c >= C) return;
    const double cc = c - 0.5 * (C - 1), rr = r - 0.5 * (R - 1);
    double s[3], d[3];
    if (kind == MR_KIND_CONE) {
        const double* g = geo + 12 * k;
        double nrm = 0.0;
        for (int i = 0; i < 3; ++i) {
            s[i] = g[i];
            d[i] = g[3 + i] + cc * g[6 + i] + rr * g[9 + i] - s[i];
            nrm += d[i] * d[i];
        }
        nrm = sqrt(nrm);
        for (int i = 0; i < 3; ++i) d[i] /= nrm;
    } else {
        const double th = geo[k];
        const double ct = cos(th), st = sin(th);
        const double u = cc * pu, v = rr * pv;
        d[0] = ct; d[1] = st; d[2] = 0.0;
        s[0] = u * -st - far * ct;
        s[1] = u * ct - far * st;
        s[2] = v;
    }
    double acc = 0.0;
    for (int e = 0; e < K; ++e) {
        const double* E = ell + ELL * e;
        const double o0 = s[0] - E[1], o1 = s[1] - E[2], o2 = s[2] - E[3];
        const double* M = E + 4;
        const double p0 = M[0] * o0 + M[1] * o1 + M[2] * o2;
        const double p1 = M[3] * o0 + M[4] * o1 + M[5] * o2;
        const double p2 = M[6] * o0 + M[7] * o1 + M[8] * o2;
        const double q0 = M[0] * d[0] + M[1] * d[1] + M[2] * d[2];
        const double q1 = M[3] * d[0] + M[4] * d[1] + M[5] * d[2];
        const double q2 = M[6] * d[0] + M[7] * d[1] + M[8] * d[2];
        const double qq = q0 * q0 + q1 * q1 + q2 * q2;
        const double pq = p0 * q0 + p1 * q1 + p2 * q2;
        const double pp = p0 * p0 + p1 * p1 + p2 * p2;
        const double disc = pq * pq - qq * (pp - 1.0);
        if (disc > 0.0) acc += E[0] * 2.0 * sqrt(disc) / qq;
    }
    out[(static_cast<int64_t>(r) * A + k) * C + c] = static_cast<float>(acc);
}

}  // namespace

int launch_phantom(const double* d_ell, int K, int kind, const double* d_geo, int A, int R, int C,
                   double pu, double pv, double far, float* out, hipStream_t s) {
    dim3 grid((C + 255) / 256, R, A);
    hipLaunchKernelGGL(phantom_kernel, grid, dim3(256), 0, s, d_ell, K, kind, d_geo, A, R, C, pu,
                       pv, far, out);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

}  // namespace mr
