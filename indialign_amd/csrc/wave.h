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

// Lane-xor partners inside groups of 8 lanes by DPP (VALU moves, no LDS
// round trip): xor 1 and xor 2 are quad permutes; xor 4 takes lane + 4 in
// banks 0 and 2 of each 16-lane row (row_shl:4) and lane - 4 in banks 1 and
// 3 (row_shr:4).
template <int M>
IA_D int xor_lane(int v) {
    static_assert(M == 1 || M == 2 || M == 4, "group-of-8 partner");
    if constexpr (M == 1) return __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xf, 0xf, false);
    else if constexpr (M == 2) return __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xf, 0xf, false);
    else {
        const int t = __builtin_amdgcn_update_dpp(v, v, 0x104, 0xf, 0x5, false);
        return __builtin_amdgcn_update_dpp(t, v, 0x114, 0xf, 0xA, false);
    }
}
template <int M>
IA_D double xor_lane(double v) {
    return __hiloint2double(xor_lane<M>(__double2hiint(v)), xor_lane<M>(__double2loint(v)));
}

// Sum over each group of 8 lanes, left in every lane of the group (the
// xor-1, 2, 4 butterfly; a + b and b + a round alike, so the 8 lanes agree).
IA_D double group8_sum(double v) {
    v += xor_lane<1>(v);
    v += xor_lane<2>(v);
    v += xor_lane<4>(v);
    return v;
}

IA_D int group8_isum(int v) {
    v += xor_lane<1>(v);
    v += xor_lane<2>(v);
    v += xor_lane<4>(v);
    return v;
}

// Lane k's value of each quad (DPP quad_perm broadcast, a VALU move).
template <int K>
IA_D double quad_bcast(double v) {
    constexpr int ctrl = K * 0x55;
    return __hiloint2double(__builtin_amdgcn_update_dpp(0, __double2hiint(v), ctrl, 0xf, 0xf, false),
                            __builtin_amdgcn_update_dpp(0, __double2loint(v), ctrl, 0xf, 0xf, false));
}

// horn_rotation for a correlation h held identically by the four lanes of
// each quad.  The Jacobi sweep is the reference's (same rotations, same
// roundings); each lane keeps only row (lane & 3) of V, so a rotation
// updates two entries of V per lane instead of eight, and the top
// eigenvector's four components are gathered from the quad at the end.
// Lanes of a quad must take this call together.
template <int P, int Q>
IA_D void jacobi_rot_quad(double a[4][4], double vr[4]) {
    double c, s;
    if (!jacobi_rot_a<P, Q>(a, c, s)) return;
    jacobi_rot_vrow<P, Q>(c, s, vr);
}

IA_D void horn_rotation_quad(const double h[9], double *R) {
    const int row = threadIdx.x & 3;
    double f[4][4];
    horn_key(h, f);
    double vr[4];
#pragma unroll
    for (int c = 0; c < 4; c++) vr[c] = c == row ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 8; sweep++) {
        jacobi_rot_quad<0, 1>(f, vr);
        jacobi_rot_quad<0, 2>(f, vr);
        jacobi_rot_quad<0, 3>(f, vr);
        jacobi_rot_quad<1, 2>(f, vr);
        jacobi_rot_quad<1, 3>(f, vr);
        jacobi_rot_quad<2, 3>(f, vr);
    }
    const int top = top_eigen(f);
    const double mine = top == 0 ? vr[0] : top == 1 ? vr[1] : top == 2 ? vr[2] : vr[3];
    quat_to_R(quad_bcast<0>(mine), quad_bcast<1>(mine), quad_bcast<2>(mine), quad_bcast<3>(mine), R);
}

// solve_rotation for a correlation held identically by each quad.
template <bool FAST>
IA_D void solve_rotation_quad(const double h[9], double *R) {
    if (FAST && qcp_rotation(h, R)) return;
    horn_rotation_quad(h, R);
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
    solve_rotation_quad<FAST>(h, R);
    finish_t(cp, cq, R, t);
}

// wave_kabsch over the entries q < n for which get(q, ...) returns true
// (a masked point list; at least 3 such entries).  Each entry enters every
// sum through a 0/1 weight: fma(1, v, s) = s + v, fma(0, v, s) = s.
template <bool FAST, class Get>
IA_D void wave_kabsch_masked(int n, Get get, double *R, double *t) {
    const int lane = threadIdx.x & 63;
    double c[6] = {0, 0, 0, 0, 0, 0}, cw = 0.0;
    for (int q = lane; q < n; q += 64) {
        double x, y, z, a, b, cc;
        const double w = get(q, x, y, z, a, b, cc) ? 1.0 : 0.0;
        c[0] = fma(w, x, c[0]); c[1] = fma(w, y, c[1]); c[2] = fma(w, z, c[2]);
        c[3] = fma(w, a, c[3]); c[4] = fma(w, b, c[4]); c[5] = fma(w, cc, c[5]);
        cw += w;
    }
#pragma unroll
    for (int d = 0; d < 6; d++) c[d] = wave_sum(c[d]);
    const double cnt = wave_sum(cw);
    double cp[3] = {c[0] / cnt, c[1] / cnt, c[2] / cnt};
    double cq[3] = {c[3] / cnt, c[4] / cnt, c[5] / cnt};
    double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = lane; q < n; q += 64) {
        double x, y, z, a, b, cc;
        const double w = get(q, x, y, z, a, b, cc) ? 1.0 : 0.0;
        const double d0 = x - cp[0], d1 = y - cp[1], d2 = z - cp[2];
        const double e0 = a - cq[0], e1 = b - cq[1], e2 = cc - cq[2];
        h[0] = fma(d0 * e0, w, h[0]); h[1] = fma(d0 * e1, w, h[1]); h[2] = fma(d0 * e2, w, h[2]);
        h[3] = fma(d1 * e0, w, h[3]); h[4] = fma(d1 * e1, w, h[4]); h[5] = fma(d1 * e2, w, h[5]);
        h[6] = fma(d2 * e0, w, h[6]); h[7] = fma(d2 * e1, w, h[7]); h[8] = fma(d2 * e2, w, h[8]);
    }
#pragma unroll
    for (int k = 0; k < 9; k++) h[k] = wave_sum(h[k]);
    solve_rotation_quad<FAST>(h, R);
    finish_t(cp, cq, R, t);
}

}  // namespace ia
