// Wavefront-wide reductions and the wave-parallel Kabsch shared by the
// chain kernel and the learned-ranking feature kernel.
#pragma once
#include "ia_common.h"

namespace ia {

IA_D double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

IA_D double group8_sum(double v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    return v;
}

IA_D int group8_isum(int v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    return v;
}

// Wave-parallel two-pass Kabsch on n point pairs produced by get(q, ...).
// Every lane ends with the same R, t (xor butterflies are order-symmetric).
template <bool FAST, class Get>
IA_D void wave_kabsch(int n, Get get, double *R, double *t) {
    const int lane = threadIdx.x & 63;
    double c[6] = {0, 0, 0, 0, 0, 0};
    for (int q = lane; q < n; q += 64) {
        double x, y, z, a, b, cc;
        get(q, x, y, z, a, b, cc);
        c[0] += x; c[1] += y; c[2] += z; c[3] += a; c[4] += b; c[5] += cc;
    }
#pragma unroll
    for (int d = 0; d < 6; d++) c[d] = wave_sum(c[d]);
    const double cnt = (double)n;
    double cp[3] = {c[0] / cnt, c[1] / cnt, c[2] / cnt};
    double cq[3] = {c[3] / cnt, c[4] / cnt, c[5] / cnt};
    double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = lane; q < n; q += 64) {
        double x, y, z, a, b, cc;
        get(q, x, y, z, a, b, cc);
        const double d0 = x - cp[0], d1 = y - cp[1], d2 = z - cp[2];
        const double e0 = a - cq[0], e1 = b - cq[1], e2 = cc - cq[2];
        h[0] += d0 * e0; h[1] += d0 * e1; h[2] += d0 * e2;
        h[3] += d1 * e0; h[4] += d1 * e1; h[5] += d1 * e2;
        h[6] += d2 * e0; h[7] += d2 * e1; h[8] += d2 * e2;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) h[k] = wave_sum(h[k]);
    solve_rotation<FAST>(h, R);
    finish_t(cp, cq, R, t);
}

}  // namespace ia
