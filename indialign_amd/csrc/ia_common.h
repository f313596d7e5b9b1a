// Shared host/device definitions for the gfx950 TM-score superposition search.
//
// All FP64 expressions below are written in the evaluation order of the
// reference (indialign_c/core.h) and the library is compiled with
// -ffp-contract=off, so a lane that sums in index order reproduces the
// reference's sequential FP64 result bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define IA_HD __host__ __device__ __forceinline__
#define IA_D __device__ __forceinline__

namespace ia {

constexpr double kNeg = -1e30;          // "no result yet" (search.h:67, rescue.h:56)
constexpr int kMaxCut = 11;             // 6 multipliers + 5 offsets (indialign.cpp:54-67)
constexpr int kWave = 64;

// One candidate transform + score: R[9] (out_d = sum_k P_k R[3k+d]), t[3], score.
struct Hit {
    double R[9];
    double t[3];
    double score;
};

// Per-pair constants, computed on the host once per call.
struct PairInfo {
    int N;          // residues
    int off;        // element offset of this pair in the coordinate / mask arenas
    int plen, nlen; // pred-valid / native-valid counts (compact DP sizes)
    int V;          // valid count
    int C;          // d0 search candidates
    double Lnorm, d0, d0s, sd8;
    double cut[kMaxCut];
};

// Coordinate arena: SoA FP64 for pred (p*) and native (q*), masks, compact
// pred/native coordinates (pc*, qc*) and compact->full index maps.
// One point, 32 bytes: a single wide (scalar or vector) load.  w is the
// valid flag (1.0 / 0.0) in the full-index pred copy, 0 elsewhere.
struct alignas(32) P4 {
    double x, y, z, w;
};

// A point pair, 64 bytes: one wide load for both sides.
struct alignas(64) P8 {
    P4 p, q;
};

struct Arena {
    const double *px, *py, *pz, *qx, *qy, *qz;
    const uint8_t *v, *pv, *nv;
    const double *pcx, *pcy, *pcz, *qcx, *qcy, *qcz;  // compact by pv / nv
    const int *pmap, *nmap;                            // compact -> full index
    const float *qfx, *qfy, *qfz;                      // compact native, FP32 (fast method)
    const char *ssp, *ssn;                             // secondary structure (full index)
    const PairInfo *info;
    const int *isel;                                   // selected work items (learned method)
    const P4 *p4, *q4;                                 // full-index AoS copies (p4.w = valid)
    const P4 *pc4, *qc4;                               // compact AoS copies
    const P8 *pq8;                                     // full-index pairs (pred, native)
    // seed-bank work counters: [0] algorithmic FP64 operations of the refine
    // loop and TM-scores (per pass 26 n + 31 n_sel + 3400, score 29 n),
    // [1] refinement passes
    unsigned long long *bank_work;
    // chain-kernel work counter: algorithmic FP64 operations of the chains'
    // superpositions and scores outside the NW cells (Kabsch 31 n + 3400,
    // TM-score 29 n, lite-search passes 26 n), as the seed banks count them
    unsigned long long *chain_work;
};

IA_HD void set_identity(double *R, double *t) {
    R[0] = 1; R[1] = 0; R[2] = 0;
    R[3] = 0; R[4] = 1; R[5] = 0;
    R[6] = 0; R[7] = 0; R[8] = 1;
    t[0] = t[1] = t[2] = 0;
}

// Rotated + translated point, reference order (core.h:127-131).
IA_HD void xform(const double *R, const double *t, double x, double y, double z, double &mx,
                 double &my, double &mz) {
    mx = x * R[0] + y * R[3] + z * R[6] + t[0];
    my = x * R[1] + y * R[4] + z * R[7] + t[1];
    mz = x * R[2] + y * R[5] + z * R[8] + t[2];
}

IA_HD double dist2_moved(const double *R, const double *t, double x, double y, double z,
                         double qx, double qy, double qz) {
    double mx, my, mz;
    xform(R, t, x, y, z, mx, my, mz);
    const double dx = mx - qx, dy = my - qy, dz = mz - qz;
    return dx * dx + dy * dy + dz * dz;
}

// Same distance with fused multiply-adds (fast method only: not the
// reference's rounding).
IA_HD double dist2_moved_fma(const double *R, const double *t, double x, double y, double z,
                             double qx, double qy, double qz) {
    const double dx = fma(x, R[0], fma(y, R[3], fma(z, R[6], t[0]))) - qx;
    const double dy = fma(x, R[1], fma(y, R[4], fma(z, R[7], t[1]))) - qy;
    const double dz = fma(x, R[2], fma(y, R[5], fma(z, R[8], t[2]))) - qz;
    return fma(dx, dx, fma(dy, dy, dz * dz));
}

// Cyclic Jacobi, 8 sweeps, pair order (0,1)(0,2)(0,3)(1,2)(1,3)(2,3),
// rotation skipped when |a_pq| < 1e-15 (core.h:29-60).  Fully unrolled so
// every index is a compile-time constant (the matrices stay in registers).
//
// On the device the square roots and the 1 / x divides run without the
// range scaling and special-case fix-up of the IEEE sequences wherever the
// argument is known to be in range (the remaining operations are the same,
// so the results are bit-identical): 1 + tn^2 is in [1, 2], and
// 1 + theta^2 is finite below 2^1022 for |theta| < 2^511.

// sqrt(x), correctly rounded, for 2^-767 <= x < 2^1022: reciprocal square
// root seed, one coupled step for sqrt and 1/(2 sqrt), two Newton corrections.
IA_D double sqrt_in_range(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    g = fma(fma(-g, g, x), h, g);
    return fma(fma(-g, g, x), h, g);
}
// 1 / y, correctly rounded, for 1 <= y < 2^1022: reciprocal seed, two Newton
// steps, one remainder correction.
IA_D double recip_in_range(double y) {
    double r = __builtin_amdgcn_rcp(y);
    r = fma(r, fma(-y, r, 1.0), r);
    r = fma(r, fma(-y, r, 1.0), r);
    return fma(fma(-y, r, 1.0), r, r);
}

// One rotation of the cyclic Jacobi on a: computes (c, s), updates a, and
// returns false when |a_pq| < 1e-15 skips it (core.h:36-57).
template <int P, int Q>
IA_HD bool jacobi_rot_a(double a[4][4], double &c, double &s) {
    const double apq = a[P][Q];
    if (fabs(apq) < 1e-15) return false;
    const double theta = (a[Q][Q] - a[P][P]) / (2.0 * apq);
#ifdef __HIP_DEVICE_COMPILE__
    double tn;
    const double at = fabs(theta);
    if (at >= 0x1p27 && at < 0x1p511) {
        // A nearly converged rotation.  theta^2 rounds to at least 2^54, so
        // 1 + theta^2 rounds back to it, and the correctly rounded square
        // root of a rounded square is |theta| (radix 2): the denominator is
        // exactly 2|theta|.  Then |tn| <= 2^-28, 1 + tn^2 rounds to 1 and
        // c = 1 / sqrt(1) = 1.  The same values as the general path below.
        const double r = recip_in_range(at + at);
        tn = theta >= 0 ? r : -r;
        c = 1.0;
    } else {
        if (at < 0x1p511) {
            const double r = recip_in_range(at + sqrt_in_range(1.0 + theta * theta));
            tn = theta >= 0 ? r : -r;
        } else {
            tn = (theta >= 0 ? 1.0 : -1.0) / (at + sqrt(1.0 + theta * theta));
        }
        c = recip_in_range(sqrt_in_range(1.0 + tn * tn));
    }
#else
    const double tn = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
    c = 1.0 / sqrt(1.0 + tn * tn);
#endif
    s = tn * c;
    a[P][P] -= tn * apq;
    a[Q][Q] += tn * apq;
    a[P][Q] = 0.0;
    a[Q][P] = 0.0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        if (r == P || r == Q) continue;
        const double xp = a[r][P], xq = a[r][Q];
        a[r][P] = c * xp - s * xq;
        a[P][r] = a[r][P];
        a[r][Q] = s * xp + c * xq;
        a[Q][r] = a[r][Q];
    }
    return true;
}

// The same rotation applied to columns P, Q of one row of V.
template <int P, int Q>
IA_HD void jacobi_rot_vrow(double c, double s, double vr[4]) {
    const double xp = vr[P], xq = vr[Q];
    vr[P] = c * xp - s * xq;
    vr[Q] = s * xp + c * xq;
}

template <int P, int Q>
IA_HD void jacobi_rot(double a[4][4], double v[4][4]) {
    double c, s;
    if (!jacobi_rot_a<P, Q>(a, c, s)) return;
#pragma unroll
    for (int r = 0; r < 4; r++) jacobi_rot_vrow<P, Q>(c, s, v[r]);
}

IA_HD void jacobi4(double a[4][4], double v[4][4]) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) v[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 8; sweep++) {
        jacobi_rot<0, 1>(a, v);
        jacobi_rot<0, 2>(a, v);
        jacobi_rot<0, 3>(a, v);
        jacobi_rot<1, 2>(a, v);
        jacobi_rot<1, 3>(a, v);
        jacobi_rot<2, 3>(a, v);
    }
}

// Horn key matrix of the 3x3 correlation h (row r = P axis, col c = Q axis),
// core.h:65-75.
IA_HD void horn_key(const double h[9], double f[4][4]) {
    f[0][0] = h[0] + h[4] + h[8];
    f[0][1] = f[1][0] = h[5] - h[7];
    f[0][2] = f[2][0] = h[6] - h[2];
    f[0][3] = f[3][0] = h[1] - h[3];
    f[1][1] = h[0] - h[4] - h[8];
    f[1][2] = f[2][1] = h[1] + h[3];
    f[1][3] = f[3][1] = h[6] + h[2];
    f[2][2] = -h[0] + h[4] - h[8];
    f[2][3] = f[3][2] = h[5] + h[7];
    f[3][3] = -h[0] - h[4] + h[8];
}

// Index of the largest diagonal entry, first maximum wins (core.h:78-80).
IA_HD int top_eigen(const double f[4][4]) {
    int top = 0;
    double best = f[0][0];
    if (f[1][1] > best) { top = 1; best = f[1][1]; }
    if (f[2][2] > best) { top = 2; best = f[2][2]; }
    if (f[3][3] > best) { top = 3; best = f[3][3]; }
    return top;
}

// Normalised quaternion -> R, identity below norm 1e-12 (core.h:81-89).
IA_HD void quat_to_R(double q0, double q1, double q2, double q3, double *R) {
    const double nrm = sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    if (nrm < 1e-12) {
        q0 = 1.0; q1 = q2 = q3 = 0.0;
    } else {
        q0 /= nrm; q1 /= nrm; q2 /= nrm; q3 /= nrm;
    }
    R[0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    R[1] = 2 * (q1 * q2 + q0 * q3);
    R[2] = 2 * (q1 * q3 - q0 * q2);
    R[3] = 2 * (q1 * q2 - q0 * q3);
    R[4] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
    R[5] = 2 * (q2 * q3 + q0 * q1);
    R[6] = 2 * (q1 * q3 + q0 * q2);
    R[7] = 2 * (q2 * q3 - q0 * q1);
    R[8] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
}

// Horn key matrix from the 3x3 correlation h, dominant eigenvector (first
// maximum wins), quaternion -> R (core.h:64-90).
IA_HD void horn_rotation(const double h[9], double *R) {
    double f[4][4];
    horn_key(h, f);
    double v[4][4];
    jacobi4(f, v);
    const int top = top_eigen(f);
    double q0, q1, q2, q3;
    if (top == 0) { q0 = v[0][0]; q1 = v[1][0]; q2 = v[2][0]; q3 = v[3][0]; }
    else if (top == 1) { q0 = v[0][1]; q1 = v[1][1]; q2 = v[2][1]; q3 = v[3][1]; }
    else if (top == 2) { q0 = v[0][2]; q1 = v[1][2]; q2 = v[2][2]; q3 = v[3][2]; }
    else { q0 = v[0][3]; q1 = v[1][3]; q2 = v[2][3]; q3 = v[3][3]; }
    quat_to_R(q0, q1, q2, q3, R);
}

// 3x3 minor of a 4x4 matrix without row I and column J.
template <int I, int J>
IA_HD double minor3(const double m[4][4]) {
    constexpr int r0 = I == 0 ? 1 : 0, r1 = I <= 1 ? 2 : 1, r2 = I <= 2 ? 3 : 2;
    constexpr int c0 = J == 0 ? 1 : 0, c1 = J <= 1 ? 2 : 1, c2 = J <= 2 ? 3 : 2;
    return m[r0][c0] * (m[r1][c1] * m[r2][c2] - m[r1][c2] * m[r2][c1]) -
           m[r0][c1] * (m[r1][c0] * m[r2][c2] - m[r1][c2] * m[r2][c0]) +
           m[r0][c2] * (m[r1][c0] * m[r2][c1] - m[r1][c1] * m[r2][c0]);
}

// Row I of adj(M) for symmetric M; keeps it in q if its norm is the largest.
template <int I>
IA_HD void adj_row_best(const double m[4][4], double &bn, double *q) {
    constexpr double s0 = (I % 2 == 0) ? 1.0 : -1.0;
    const double a = s0 * minor3<I, 0>(m), b = -s0 * minor3<I, 1>(m);
    const double c = s0 * minor3<I, 2>(m), d = -s0 * minor3<I, 3>(m);
    const double nn = a * a + b * b + c * c + d * d;
    if (nn > bn) {
        bn = nn;
        q[0] = a; q[1] = b; q[2] = c; q[3] = d;
    }
}

// Closed-form alternative to horn_rotation for the fast method (FP64):
// the largest eigenvalue of the traceless Horn key matrix F is the largest
// root of its characteristic quartic x^4 + c2 x^2 + c1 x + c0 (Theobald,
// Acta Cryst. A61:478, 2005), found by Newton's method from the upper bound
// sqrt(3)*|H|_F >= s1 + s2 + s3; the eigenvector is the largest row of
// adj(F - x I).  Returns false when the spectrum is too degenerate for the
// adjugate (the caller falls back to horn_rotation).
IA_HD bool qcp_rotation(const double h[9], double *R) {
    double m[4][4];
    m[0][0] = h[0] + h[4] + h[8];
    m[0][1] = m[1][0] = h[5] - h[7];
    m[0][2] = m[2][0] = h[6] - h[2];
    m[0][3] = m[3][0] = h[1] - h[3];
    m[1][1] = h[0] - h[4] - h[8];
    m[1][2] = m[2][1] = h[1] + h[3];
    m[1][3] = m[3][1] = h[6] + h[2];
    m[2][2] = -h[0] + h[4] - h[8];
    m[2][3] = m[3][2] = h[5] + h[7];
    m[3][3] = -h[0] - h[4] + h[8];
    double n2 = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) n2 += h[k] * h[k];
    if (!(n2 > 1e-300)) return false;
    const double det3 = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                        h[2] * (h[3] * h[7] - h[4] * h[6]);
    const double detF = m[0][0] * minor3<0, 0>(m) - m[0][1] * minor3<0, 1>(m) +
                        m[0][2] * minor3<0, 2>(m) - m[0][3] * minor3<0, 3>(m);
    const double c2 = -2.0 * n2, c1 = -8.0 * det3, c0 = detF;
    double x = sqrt(3.0 * n2);
    for (int it = 0; it < 60; it++) {
        const double x2 = x * x, b = (x2 + c2) * x, a = b + c1;
        const double d = (a * x + c0) / (2.0 * x2 * x + b + a);
        x -= d;
        if (fabs(d) <= 1e-14 * fabs(x)) break;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) m[i][i] -= x;
    double bn = 0.0, q[4] = {1.0, 0.0, 0.0, 0.0};
    adj_row_best<0>(m, bn, q);
    adj_row_best<1>(m, bn, q);
    adj_row_best<2>(m, bn, q);
    adj_row_best<3>(m, bn, q);
    // a tiny adjugate means a (near) repeated top eigenvalue
    if (!(bn > 1e-20 * n2 * n2 * n2)) return false;
    const double inv = 1.0 / sqrt(bn);
    const double q0 = q[0] * inv, q1 = q[1] * inv, q2 = q[2] * inv, q3 = q[3] * inv;
    R[0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    R[1] = 2 * (q1 * q2 + q0 * q3);
    R[2] = 2 * (q1 * q3 - q0 * q2);
    R[3] = 2 * (q1 * q2 - q0 * q3);
    R[4] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
    R[5] = 2 * (q2 * q3 + q0 * q1);
    R[6] = 2 * (q1 * q3 + q0 * q2);
    R[7] = 2 * (q2 * q3 - q0 * q1);
    R[8] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
    return true;
}

// Rotation from a correlation: exact Jacobi path, or the closed form with a
// Jacobi fallback.
template <bool FAST>
IA_HD void solve_rotation(const double h[9], double *R) {
    if (FAST && qcp_rotation(h, R)) return;
    horn_rotation(h, R);
}

// t = cQ - cP.R (core.h:121-122).
IA_HD void finish_t(const double *cp, const double *cq, const double *R, double *t) {
    t[0] = cq[0] - (cp[0] * R[0] + cp[1] * R[3] + cp[2] * R[6]);
    t[1] = cq[1] - (cp[0] * R[1] + cp[1] * R[4] + cp[2] * R[7]);
    t[2] = cq[2] - (cp[0] * R[2] + cp[1] * R[5] + cp[2] * R[8]);
}

// Sequential two-pass Kabsch accumulators (core.h:94-123): feed the selected
// points in index order to add_centroid, call centre(), feed them again to
// add_cov, then solve().
struct KabschSeq {
    double cp[3], cq[3], cnt, h[9];
    IA_HD void reset() {
        cp[0] = cp[1] = cp[2] = cq[0] = cq[1] = cq[2] = 0.0;
        cnt = 0.0;
#pragma unroll
        for (int k = 0; k < 9; k++) h[k] = 0.0;
    }
    IA_HD void add_centroid(double x, double y, double z, double a, double b, double c) {
        cp[0] += x; cp[1] += y; cp[2] += z;
        cq[0] += a; cq[1] += b; cq[2] += c;
        cnt += 1.0;
    }
    IA_HD void centre() {
        cp[0] /= cnt; cp[1] /= cnt; cp[2] /= cnt;
        cq[0] /= cnt; cq[1] /= cnt; cq[2] /= cnt;
    }
    IA_HD void add_cov(double x, double y, double z, double a, double b, double c) {
        const double d0 = x - cp[0], d1 = y - cp[1], d2 = z - cp[2];
        const double e0 = a - cq[0], e1 = b - cq[1], e2 = c - cq[2];
        h[0] += d0 * e0; h[1] += d0 * e1; h[2] += d0 * e2;
        h[3] += d1 * e0; h[4] += d1 * e1; h[5] += d1 * e2;
        h[6] += d2 * e0; h[7] += d2 * e1; h[8] += d2 * e2;
    }
    template <bool FAST = false>
    IA_HD void solve(double *R, double *t) const {
        solve_rotation<FAST>(h, R);
        finish_t(cp, cq, R, t);
    }
    // Single-pass alternative (fast method): raw sums and second moments,
    // H = sum x q^T - cnt cp cq^T after finish_moments().  In FP64 the
    // cancellation costs ~|x||q| / |H| ulps, about 1e-15 relative for
    // coordinates within 100 A.
    IA_HD void add_moment(double x, double y, double z, double a, double b, double c) {
        cp[0] += x; cp[1] += y; cp[2] += z;
        cq[0] += a; cq[1] += b; cq[2] += c;
        cnt += 1.0;
        h[0] = fma(x, a, h[0]); h[1] = fma(x, b, h[1]); h[2] = fma(x, c, h[2]);
        h[3] = fma(y, a, h[3]); h[4] = fma(y, b, h[4]); h[5] = fma(y, c, h[5]);
        h[6] = fma(z, a, h[6]); h[7] = fma(z, b, h[7]); h[8] = fma(z, c, h[8]);
    }
    IA_HD void finish_moments() {
        centre();
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) h[3 * i + j] -= cnt * cp[i] * cq[j];
    }
};

// TM term 1/(1 + d2/d0sq) (core.h:147,163).
// a / b rounded to nearest from inv = RN(1 / b): a faithful quotient a * inv
// corrected once with an exact FMA remainder is the correctly rounded
// quotient (Markstein's theorem; no underflow or overflow in this range).
// Bit-identical to a / b, at a multiply and two FMAs when b is loop-invariant.
IA_HD double div_by(double a, double b, double inv) {
    const double q = a * inv;
    return fma(fma(-b, q, a), inv, q);
}

// TM term 1/(1 + d2/d0sq) (core.h:147,163), rounded exactly as written; the
// reciprocal of d0sq is loop-invariant wherever this runs in a loop.
IA_HD double tm_term(double d2, double d0sq) {
    return 1.0 / (1.0 + div_by(d2, d0sq, 1.0 / d0sq));
}

// ---- window seeds over n points (seeds.h:15-37), computed on the fly ----
// Returns the number of distinct window lengths (descending) in lens.
IA_HD int window_lengths(int V, int *lens) {
    if (V < 4) return 0;
    int c[9] = {V, V / 2, V / 4, V / 8, 64, 32, 16, 8, 4};
    int n = 0;
    // insertion into a descending unique list
    for (int k = 0; k < 9; k++) {
        const int x = c[k];
        if (x < 4 || x > V) continue;
        bool dup = false;
        for (int j = 0; j < n; j++) dup |= lens[j] == x;
        if (dup) continue;
        int pos = n;
        while (pos > 0 && lens[pos - 1] < x) {
            lens[pos] = lens[pos - 1];
            pos--;
        }
        lens[pos] = x;
        n++;
    }
    return n;
}

IA_HD int window_start_count(int n, int len, int ms) {
    const int total = n - len + 1;
    if (total <= 0) return 0;
    if (total <= ms) return total;
    int step = total / ms;
    if (step < 1) step = 1;
    const int cnt = (total - 1) / step + 1;           // 0, step, ... < total
    return cnt + (((cnt - 1) * step != total - 1) ? 1 : 0);
}

IA_HD int window_start(int n, int len, int ms, int idx) {
    const int total = n - len + 1;
    if (total <= ms) return idx;
    int step = total / ms;
    if (step < 1) step = 1;
    const int cnt = (total - 1) / step + 1;
    return idx < cnt ? idx * step : total - 1;
}

// Seeds over n all-valid points: k = 0 is the whole set, then the windows.
// (All windows hold >= 4 points, so none is dropped by the >= 3 filter.)
IA_HD int implicit_seed_count(int n, int ms) {
    if (n < 3) return 0;
    int lens[9];
    const int nl = window_lengths(n, lens);
    int K = 1;
    for (int j = 0; j < nl; j++) K += window_start_count(n, lens[j], ms);
    return K;
}

IA_HD void implicit_seed(int n, int ms, int k, int &start, int &len) {
    if (k == 0) { start = 0; len = n; return; }
    int lens[9];
    const int nl = window_lengths(n, lens);
    int r = k - 1;
    for (int j = 0; j < nl; j++) {
        const int c = window_start_count(n, lens[j], ms);
        if (r < c) { start = window_start(n, lens[j], ms, r); len = lens[j]; return; }
        r -= c;
    }
    start = 0; len = 0;
}

// Window seeds of a PyTorch-engine DP list (search.py:47-77 over the list,
// dp.py:158-159): k = 0 is the whole list, then for each length of
// window_lengths(vl) (vl = the list's valid count) the start positions over
// the padded list length nl; seed masks are ANDed with the entries' flags.
IA_HD int torch_list_seed_count(int nl, int vl, int ms) {
    int lens[9];
    const int n_lens = window_lengths(vl, lens);
    int K = 1;
    for (int j = 0; j < n_lens; j++) K += window_start_count(nl, lens[j], ms);
    return K;
}

IA_HD void torch_list_seed(int nl, int vl, int ms, int k, int &start, int &len) {
    if (k == 0) { start = 0; len = nl; return; }
    int lens[9];
    const int n_lens = window_lengths(vl, lens);
    int r = k - 1;
    for (int j = 0; j < n_lens; j++) {
        const int c = window_start_count(nl, lens[j], ms);
        if (r < c) { start = window_start(nl, lens[j], ms, r); len = lens[j]; return; }
        r -= c;
    }
    start = 0; len = 0;
}

}  // namespace ia
