// CPU FP64 oracle for the TM-score superposition search.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h): a restatement of the reference
// algorithm written for this repository, used as the checker by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Built with IEEE
// semantics (-fno-fast-math -ffp-contract=off) so that every expression is
// evaluated in the order written; each function cites the reference lines
// whose behaviour it restates.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <random>
#include <vector>

namespace {

struct Xf {
    double R[9];
    double t[3];
};
struct Hit {
    Xf x;
    double score;
};

constexpr double kNeg = -1e30;

Xf identity_xf() {
    Xf x{};
    x.R[0] = x.R[4] = x.R[8] = 1.0;
    return x;
}
Hit hit_of(double s) { return Hit{identity_xf(), s}; }

void pack(const Hit &h, double *out13) {
    std::memcpy(out13, h.x.R, 72);
    std::memcpy(out13 + 9, h.x.t, 24);
    out13[12] = h.score;
}

// ---- d0 normalisation: core.h:12-25 -------------------------------------
void d0_params(double L, double &d0, double &d0s, double &sd8) {
    const double Lc = L < 1.0 ? 1.0 : L;
    d0 = Lc <= 11.0 ? 0.3 : Lc <= 15.0 ? 0.4 : Lc <= 19.0 ? 0.5 : Lc <= 23.0 ? 0.6 : 0.7;
    if (Lc >= 30.0) d0 = 0.6 * std::sqrt(Lc - 0.5) - 2.5;
    if (d0 < 0.3) d0 = 0.3;
    const double v = 1.24 * std::cbrt(Lc - 15.0 > 0.0 ? Lc - 15.0 : 0.0) - 1.8;
    d0s = v < 4.5 ? 4.5 : (v > 8.0 ? 8.0 : v);
    sd8 = 1.5 * std::pow(Lc, 0.3) + 3.5;
}

// ---- cyclic Jacobi on the 4x4 key matrix: core.h:29-60 ------------------
// Eight sweeps over the pairs (0,1) (0,2) (0,3) (1,2) (1,3) (2,3); a rotation
// is skipped when |a_pq| < 1e-15.
void jacobi_sym4(double a[4][4], double v[4][4]) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) v[i][j] = i == j ? 1.0 : 0.0;
    static const int P[6] = {0, 0, 0, 1, 1, 2};
    static const int Qi[6] = {1, 2, 3, 2, 3, 3};
    for (int sweep = 0; sweep < 8; sweep++) {
        for (int k = 0; k < 6; k++) {
            const int p = P[k], q = Qi[k];
            const double apq = a[p][q];
            if (std::fabs(apq) < 1e-15) continue;
            const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
            const double tn = (theta >= 0 ? 1.0 : -1.0) /
                              (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
            const double c = 1.0 / std::sqrt(1.0 + tn * tn);
            const double s = tn * c;
            a[p][p] -= tn * apq;
            a[q][q] += tn * apq;
            a[p][q] = 0.0;
            a[q][p] = 0.0;
            for (int r = 0; r < 4; r++) {
                if (r == p || r == q) continue;
                const double xp = a[r][p], xq = a[r][q];
                a[r][p] = c * xp - s * xq;
                a[p][r] = a[r][p];
                a[r][q] = s * xp + c * xq;
                a[q][r] = a[r][q];
            }
            for (int r = 0; r < 4; r++) {
                const double xp = v[r][p], xq = v[r][q];
                v[r][p] = c * xp - s * xq;
                v[r][q] = s * xp + c * xq;
            }
        }
    }
}

// ---- Horn quaternion from a 3x3 correlation: core.h:64-90 ---------------
void horn_rotation(const double h[3][3], double *R) {
    double f[4][4];
    f[0][0] = h[0][0] + h[1][1] + h[2][2];
    f[0][1] = f[1][0] = h[1][2] - h[2][1];
    f[0][2] = f[2][0] = h[2][0] - h[0][2];
    f[0][3] = f[3][0] = h[0][1] - h[1][0];
    f[1][1] = h[0][0] - h[1][1] - h[2][2];
    f[1][2] = f[2][1] = h[0][1] + h[1][0];
    f[1][3] = f[3][1] = h[2][0] + h[0][2];
    f[2][2] = -h[0][0] + h[1][1] - h[2][2];
    f[2][3] = f[3][2] = h[1][2] + h[2][1];
    f[3][3] = -h[0][0] - h[1][1] + h[2][2];
    double vec[4][4];
    jacobi_sym4(f, vec);
    int top = 0;
    for (int i = 1; i < 4; i++)
        if (f[i][i] > f[top][top]) top = i;  // first maximum wins
    double q0 = vec[0][top], q1 = vec[1][top], q2 = vec[2][top], q3 = vec[3][top];
    const double nrm = std::sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    if (nrm < 1e-12) {
        q0 = 1.0;
        q1 = q2 = q3 = 0.0;
    } else {
        q0 /= nrm;
        q1 /= nrm;
        q2 /= nrm;
        q3 /= nrm;
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

// Translation that carries the P centroid onto the Q centroid under R
// (core.h:121-122).
void finish_translation(const double cp[3], const double cq[3], Xf &x) {
    for (int d = 0; d < 3; d++)
        x.t[d] = cq[d] - (cp[0] * x.R[d] + cp[1] * x.R[3 + d] + cp[2] * x.R[6 + d]);
}

// ---- masked two-pass Kabsch: core.h:94-123 ------------------------------
// Selected rows are summed in index order; fewer than 3 rows -> identity.
Xf superpose(const double *P, const double *Q, const uint8_t *m, int n) {
    double cp[3] = {0, 0, 0}, cq[3] = {0, 0, 0}, cnt = 0.0;
    for (int i = 0; i < n; i++) {
        if (!m[i]) continue;
        for (int d = 0; d < 3; d++) {
            cp[d] += P[3 * i + d];
            cq[d] += Q[3 * i + d];
        }
        cnt += 1.0;
    }
    if (cnt < 3.0) return identity_xf();
    for (int d = 0; d < 3; d++) {
        cp[d] /= cnt;
        cq[d] /= cnt;
    }
    double h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < n; i++) {
        if (!m[i]) continue;
        double a[3], b[3];
        for (int d = 0; d < 3; d++) {
            a[d] = P[3 * i + d] - cp[d];
            b[d] = Q[3 * i + d] - cq[d];
        }
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) h[r][c] += a[r] * b[c];
    }
    Xf x;
    horn_rotation(h, x.R);
    finish_translation(cp, cq, x);
    return x;
}

// ---- weighted Kabsch: core.h:183-214 ------------------------------------
Xf superpose_weighted(const double *P, const double *Q, const uint8_t *v, const double *w,
                      int n) {
    double cp[3] = {0, 0, 0}, cq[3] = {0, 0, 0}, ws = 0.0;
    for (int i = 0; i < n; i++) {
        if (!v[i] || w[i] <= 0) continue;
        for (int d = 0; d < 3; d++) {
            cp[d] += w[i] * P[3 * i + d];
            cq[d] += w[i] * Q[3 * i + d];
        }
        ws += w[i];
    }
    if (ws < 1e-12) return identity_xf();
    for (int d = 0; d < 3; d++) {
        cp[d] /= ws;
        cq[d] /= ws;
    }
    double h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < n; i++) {
        if (!v[i] || w[i] <= 0) continue;
        double a[3], b[3];
        for (int d = 0; d < 3; d++) {
            a[d] = P[3 * i + d] - cp[d];
            b[d] = Q[3 * i + d] - cq[d];
        }
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) h[r][c] += w[i] * a[r] * b[c];
    }
    Xf x;
    horn_rotation(h, x.R);
    finish_translation(cp, cq, x);
    return x;
}

// Squared distance between x.R-rotated, translated P[i] and Q[j]
// (expression order of core.h:141-145).
inline double moved_d2(const Xf &x, const double *P, int i, const double *Q, int j) {
    const double *p = P + 3 * i, *q = Q + 3 * j;
    const double mx = p[0] * x.R[0] + p[1] * x.R[3] + p[2] * x.R[6] + x.t[0];
    const double my = p[0] * x.R[1] + p[1] * x.R[4] + p[2] * x.R[7] + x.t[1];
    const double mz = p[0] * x.R[2] + p[1] * x.R[5] + p[2] * x.R[8] + x.t[2];
    const double dx = mx - q[0], dy = my - q[1], dz = mz - q[2];
    return dx * dx + dy * dy + dz * dz;
}

// ---- TM-score: core.h:134-166 (use_d8 selects tm_score vs tm_score_no_d8)
double tm_eval(const double *P, const double *Q, const uint8_t *v, const Xf &x, double d0,
               double Lnorm, double sd8, int n, bool use_d8) {
    const double d0sq = std::max(d0 * d0, 1e-12), cut = sd8 * sd8;
    double acc = 0.0;
    for (int i = 0; i < n; i++) {
        if (!v[i]) continue;
        const double d2 = moved_d2(x, P, i, Q, i);
        if (use_d8 && !(d2 <= cut)) continue;
        acc += 1.0 / (1.0 + d2 / d0sq);
    }
    return acc / std::max(Lnorm, 1.0);
}

// ---- weighted TM refinement: core.h:218-237 -----------------------------
Hit weighted_refine(const double *P, const double *Q, const uint8_t *v, double d0, double Lnorm,
                    int n, int iters, Xf x) {
    const double d0sq = std::max(d0 * d0, 1e-12);
    std::vector<double> w(n);
    for (int it = 0; it < iters; it++) {
        for (int i = 0; i < n; i++)
            w[i] = v[i] ? 1.0 / (1.0 + moved_d2(x, P, i, Q, i) / d0sq) : 0.0;
        x = superpose_weighted(P, Q, v, w.data(), n);
    }
    return Hit{x, tm_eval(P, Q, v, x, d0, Lnorm, 0.0, n, false)};
}

// ---- seed windows: seeds.h:15-82 ----------------------------------------
std::vector<int> window_lengths(int V) {
    std::vector<int> out;
    if (V < 4) return out;
    const int cand[9] = {V, V / 2, V / 4, V / 8, 64, 32, 16, 8, 4};
    for (int c : cand)
        if (c >= 4 && c <= V) out.push_back(c);
    std::sort(out.begin(), out.end(), [](int a, int b) { return a > b; });
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

std::vector<int> window_starts(int n, int len, int max_starts) {
    std::vector<int> out;
    const int total = n - len + 1;
    if (total <= 0) return out;
    if (total <= max_starts) {
        for (int s = 0; s < total; s++) out.push_back(s);
        return out;
    }
    const int step = std::max(1, total / max_starts);
    for (int s = 0; s < total; s += step) out.push_back(s);
    if (out.back() != total - 1) out.push_back(total - 1);
    return out;
}

// Masks are stored back to back, N bytes each.  Global seed first, then the
// windows by length (descending) and start; masks with < 3 set bytes dropped.
void seed_masks(int N, const uint8_t *valid, bool use_frag, int max_starts,
                std::vector<uint8_t> &out) {
    out.clear();
    std::vector<uint8_t> cand(valid, valid + N);
    if (use_frag) {
        int V = 0;
        for (int i = 0; i < N; i++) V += valid[i];
        for (int len : window_lengths(V)) {
            for (int s : window_starts(N, len, max_starts)) {
                int c = 0;
                for (int i = s; i < s + len && i < N; i++) c += valid[i];
                if (c < 3) continue;
                const size_t o = cand.size();
                cand.resize(o + N, 0);
                for (int i = s; i < s + len && i < N; i++) cand[o + i] = valid[i];
            }
        }
    }
    const int K = (int)(cand.size() / (size_t)N);
    for (int k = 0; k < K; k++) {
        int c = 0;
        for (int i = 0; i < N; i++) c += cand[(size_t)k * N + i];
        if (c >= 3) out.insert(out.end(), cand.begin() + (size_t)k * N, cand.begin() + (size_t)(k + 1) * N);
    }
}

// ---- FNV-1a mask hash: seeds.h:102-109 ----------------------------------
uint64_t fnv1a(const uint8_t *m, int n) {
    uint64_t h = 1469598103934665603ULL;
    for (int i = 0; i < n; i++) {
        h ^= m[i];
        h *= 1099511628211ULL;
    }
    return h;
}

// ---- anchor contact seeds: seeds.h:164-283 ------------------------------
// The reference scans a hash grid of cell size r; a pair (i, j) is examined
// only when their cells are adjacent on every axis, which is restated here as
// an explicit cell-adjacency test in front of the distance tests.
void anchor_seeds(const double *P, const double *Q, const uint8_t *valid, int N, double d0s,
                  double d0, int max_anchors, std::vector<uint8_t> &masks) {
    double radii[6];
    int nr = 0;
    radii[nr++] = d0s;
    radii[nr++] = d0s * 1.5;
    radii[nr++] = d0s * 2.0;
    const double extra[3] = {d0 + 1.0, d0 + 2.0, d0 + 3.0};
    for (double e : extra) {
        bool dup = false;
        for (int j = 0; j < nr && !dup; j++) dup = std::fabs(radii[j] - e) < 0.01;
        if (!dup) radii[nr++] = e;
    }
    std::vector<int> vi;
    for (int i = 0; i < N; i++)
        if (valid[i]) vi.push_back(i);
    const int V = (int)vi.size();
    if (V < 3) return;
    const int K0 = (int)(masks.size() / (size_t)N);
    std::unordered_set<uint64_t> seen;
    for (int k = 0; k < K0; k++) seen.insert(fnv1a(&masks[(size_t)k * N], N));

    std::vector<int64_t> cell(3 * (size_t)N);
    for (int ri = 0; ri < nr; ri++) {
        const double r = radii[ri];
        if (r <= 0) continue;
        const double r2 = r * r, tol = std::max(0.5, r * 0.25);
        for (int i : vi)
            for (int d = 0; d < 3; d++) cell[3 * i + d] = (int64_t)std::floor(P[3 * i + d] / r);
        auto consistent = [&](int i, int j) {
            for (int d = 0; d < 3; d++) {
                const int64_t dc = cell[3 * i + d] - cell[3 * j + d];
                if (dc < -1 || dc > 1) return false;
            }
            const double pdx = P[3 * i] - P[3 * j], pdy = P[3 * i + 1] - P[3 * j + 1],
                         pdz = P[3 * i + 2] - P[3 * j + 2];
            const double pd2 = pdx * pdx + pdy * pdy + pdz * pdz;
            if (pd2 > r2) return false;
            const double ndx = Q[3 * i] - Q[3 * j], ndy = Q[3 * i + 1] - Q[3 * j + 1],
                         ndz = Q[3 * i + 2] - Q[3 * j + 2];
            const double nd2 = ndx * ndx + ndy * ndy + ndz * ndz;
            if (nd2 > r2) return false;
            return std::fabs(std::sqrt(pd2) - std::sqrt(nd2)) <= tol;
        };
        std::vector<std::pair<int, int>> ranked;
        for (int a = 0; a < V; a++) {
            int c = 1;
            for (int j : vi)
                if (j != vi[a] && consistent(vi[a], j)) c++;
            if (c >= 3) ranked.push_back({c, a});
        }
        // Same comparator and input order as seeds.h:237-238, so libstdc++'s
        // std::sort yields the same order among equal counts.
        std::sort(ranked.begin(), ranked.end(),
                  [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
                      return x.first > y.first;
                  });
        int added = 0;
        std::vector<uint8_t> m(N);
        for (const auto &e : ranked) {
            if (added >= max_anchors) break;
            const int i = vi[e.second];
            std::fill(m.begin(), m.end(), 0);
            m[i] = 1;
            int mc = 1;
            for (int j : vi)
                if (j != i && consistent(i, j)) {
                    m[j] = 1;
                    mc++;
                }
            if (mc < 3) continue;
            const uint64_t h = fnv1a(m.data(), N);
            if (!seen.insert(h).second) continue;
            masks.insert(masks.end(), m.begin(), m.end());
            added++;
        }
    }
}

// ---- top-k distance seed: seeds.h:85-98 ---------------------------------
void topk_mask(const double *d2, const uint8_t *valid, int N, double frac, uint8_t *mask) {
    int V = 0;
    for (int i = 0; i < N; i++) V += valid[i];
    const int k = std::min(std::max(3, (int)(V * frac)), N);
    std::vector<std::pair<double, int>> order;
    for (int i = 0; i < N; i++)
        if (valid[i]) order.push_back({d2[i], i});
    std::sort(order.begin(), order.end());
    std::memset(mask, 0, N);
    for (int i = 0; i < std::min(k, (int)order.size()); i++) mask[order[i].second] = 1;
}

// ---- d0s-cutoff refinement loop: search.h:28-56 / threading.h:15-37 -----
// Select valid rows within d0s; re-superpose on >= 3 rows; stop when the
// selected count repeats.
void cutoff_refine(const double *P, const double *Q, const uint8_t *v, double d0s, int n,
                   int iters, Xf &x, std::vector<uint8_t> &sel, int *passes = nullptr) {
    const double lim = d0s * d0s;
    int last = -1;
    for (int it = 0; it < iters; it++) {
        int cnt = 0;
        for (int i = 0; i < n; i++) {
            sel[i] = v[i] && moved_d2(x, P, i, Q, i) <= lim;
            cnt += sel[i];
        }
        if (cnt >= 3) {
            x = superpose(P, Q, sel.data(), n);
            if (passes) ++*passes;
        }
        if (cnt == last) break;
        last = cnt;
    }
}

Xf refine_seed(const double *P, const double *Q, const uint8_t *v, const uint8_t *seed,
               double d0s, int n, int iters, std::vector<uint8_t> &ws, std::vector<uint8_t> &sel,
               int *passes = nullptr) {
    for (int i = 0; i < n; i++) ws[i] = seed[i] & v[i];
    Xf x = superpose(P, Q, ws.data(), n);
    if (passes) *passes = 1;
    cutoff_refine(P, Q, v, d0s, n, iters, x, sel, passes);
    return x;
}

// ---- seed bank: search.h:62-138, serial order ---------------------------
// Candidates are visited seed-major; strict '>' keeps the first maximum;
// before each seed the bank stops once the best reaches early_term.
Hit seed_bank(const double *P, const double *Q, const uint8_t *v, const uint8_t *seeds, int K,
              const double *cut, int C, double d0, double sd8, double Lnorm, int n, int iters,
              bool use_d8, double early_term) {
    Hit best = hit_of(kNeg);
    std::vector<uint8_t> ws(n), sel(n);
    for (int k = 0; k < K; k++) {
        if (best.score >= early_term) break;
        for (int c = 0; c < C; c++) {
            const Xf x = refine_seed(P, Q, v, seeds + (size_t)k * n, cut[c], n, iters, ws, sel);
            const double s = tm_eval(P, Q, v, x, d0, Lnorm, sd8, n, use_d8);
            if (s > best.score) best = Hit{x, s};
        }
    }
    return best;
}

// ---- Needleman-Wunsch: nw.h:10-68 ---------------------------------------
// H and trace are (Np+1) x (Nn+1); trace -1 = boundary, 0 = diagonal,
// 1 = up, 2 = left.  A gap move pays gap_open only when the cell it leaves
// was reached diagonally.  Ties prefer diagonal, then left, then up.  Cells
// depend only on their up/left/diagonal neighbours, so a row-major sweep
// produces the same values as the reference's anti-diagonal sweep.
void nw_fill(const double *S, int Np, int Nn, double gap, std::vector<double> &H,
             std::vector<int8_t> &T) {
    const int W = Nn + 1;
    H.assign((size_t)(Np + 1) * W, 0.0);
    T.assign((size_t)(Np + 1) * W, -1);
    for (int i = 1; i <= Np; i++) {
        for (int j = 1; j <= Nn; j++) {
            const size_t c = (size_t)i * W + j, u = c - W, l = c - 1, dg = u - 1;
            const double diag = H[dg] + S[(size_t)(i - 1) * Nn + (j - 1)];
            double up = H[u];
            if (T[u] == 0) up += gap;
            double left = H[l];
            if (T[l] == 0) left += gap;
            if (diag >= up && diag >= left) {
                H[c] = diag;
                T[c] = 0;
            } else if (left >= up) {
                H[c] = left;
                T[c] = 2;
            } else {
                H[c] = up;
                T[c] = 1;
            }
        }
    }
}

// Walk from (Np, Nn) back to a border; returns aligned (p, n) index pairs in
// increasing order.
int nw_walk(const std::vector<int8_t> &T, int Np, int Nn, int *ap, int *an) {
    const int W = Nn + 1;
    int i = Np, j = Nn, k = 0;
    std::vector<std::pair<int, int>> rev;
    while (i > 0 && j > 0) {
        const int8_t d = T[(size_t)i * W + j];
        if (d == 0) {
            rev.push_back({i - 1, j - 1});
            i--;
            j--;
        } else if (d == 1) {
            i--;
        } else {
            j--;
        }
    }
    for (auto it = rev.rbegin(); it != rev.rend(); ++it, ++k) {
        ap[k] = it->first;
        an[k] = it->second;
    }
    return k;
}

// ---- the per-pair search context ----------------------------------------
struct Pair {
    int N;
    const double *P, *Q;
    const uint8_t *v, *pv, *nv;
    double Lnorm, d0, d0s, sd8;
    std::vector<int> pmap, nmap;  // compact indices of pred-/native-valid rows

    Pair(int n, const double *p, const double *q, const uint8_t *va, const uint8_t *pva,
         const uint8_t *nva, double L)
        : N(n), P(p), Q(q), v(va), pv(pva), nv(nva), Lnorm(L) {
        d0_params(L, d0, d0s, sd8);
        for (int i = 0; i < N; i++)
            if (pv[i]) pmap.push_back(i);
        for (int i = 0; i < N; i++)
            if (nv[i]) nmap.push_back(i);
    }
    int plen() const { return (int)pmap.size(); }
    int nlen() const { return (int)nmap.size(); }
    double same_index(const Xf &x) const { return tm_eval(P, Q, v, x, d0, Lnorm, 0.0, N, false); }
};

// ---- detailed search on an alignment: align.h:15-50 ---------------------
Hit detailed_search(const Pair &pr, const int *ap, const int *an, int na, double d0s,
                    double sd8, bool use_frag, int iters_override, int max_starts) {
    if (na < 3) return hit_of(kNeg);
    std::vector<double> A(3 * (size_t)na), B(3 * (size_t)na);
    for (int i = 0; i < na; i++) {
        std::memcpy(&A[3 * i], pr.P + 3 * ap[i], 24);
        std::memcpy(&B[3 * i], pr.Q + 3 * an[i], 24);
    }
    std::vector<uint8_t> ones(na, 1), masks;
    seed_masks(na, ones.data(), use_frag, max_starts, masks);
    const int K = (int)(masks.size() / (size_t)na);
    if (K == 0) return hit_of(kNeg);
    static const double mult[6] = {2.0, 1.5, 1.0, 0.75, 0.5, 0.25};
    double cut[6];
    for (int c = 0; c < 6; c++) cut[c] = d0s * mult[c];
    const int iters = iters_override >= 0 ? iters_override : (use_frag ? 20 : 8);
    return seed_bank(A.data(), B.data(), ones.data(), masks.data(), K, cut, 6, pr.d0, sd8,
                     pr.Lnorm, na, iters, true, 1.0);
}

// ---- DP refinement: align.h:54-159 --------------------------------------
// Two chains (gap open -0.6 then 0.0), each iterating: TM-distance score
// matrix over compact rows -> NW -> Kabsch on the aligned rows -> detailed
// search; the running best starts at (identity, 0) and is shared.
Hit dp_refine(const Pair &pr, const Xf &x0, double sd8, int iters, bool lite) {
    Hit best = hit_of(0.0);
    const int pl = pr.plen(), nl = pr.nlen();
    if (pl < 3 || nl < 3) return best;
    const double d0dp = std::max(pr.d0, 1.0);
    const double d0sq_dp = std::max(d0dp * d0dp, 1e-12), cut = sd8 * sd8;
    const double d0sq = std::max(pr.d0 * pr.d0, 1e-12);
    std::vector<double> S((size_t)pl * nl), H, moved(3 * (size_t)pr.N);
    std::vector<int8_t> T;
    std::vector<int> ap(std::min(pl, nl)), an(std::min(pl, nl));
    static const double gaps[2] = {-0.6, 0.0};
    for (double gap : gaps) {
        Xf cur = x0;
        for (int it = 0; it < iters; it++) {
            for (int a = 0; a < pl; a++) {
                const int i = pr.pmap[a];
                for (int b = 0; b < nl; b++) {
                    const double d2 = moved_d2(cur, pr.P, i, pr.Q, pr.nmap[b]);
                    S[(size_t)a * nl + b] = d2 <= cut ? 1.0 / (1.0 + d2 / d0sq_dp) : 0.0;
                }
            }
            nw_fill(S.data(), pl, nl, gap, H, T);
            int na = nw_walk(T, pl, nl, ap.data(), an.data());
            for (int k = 0; k < na; k++) {
                ap[k] = pr.pmap[ap[k]];
                an[k] = pr.nmap[an[k]];
            }
            if (na < 3) continue;
            std::vector<double> A(3 * (size_t)na), B(3 * (size_t)na);
            for (int k = 0; k < na; k++) {
                std::memcpy(&A[3 * k], pr.P + 3 * ap[k], 24);
                std::memcpy(&B[3 * k], pr.Q + 3 * an[k], 24);
            }
            std::vector<uint8_t> ones(na, 1);
            Xf nx = superpose(A.data(), B.data(), ones.data(), na);
            double s = 0.0;
            for (int k = 0; k < na; k++) s += 1.0 / (1.0 + moved_d2(nx, A.data(), k, B.data(), k) / d0sq);
            s /= std::max(pr.Lnorm, 1.0);
            const Hit det = detailed_search(pr, ap.data(), an.data(), na, pr.d0s, sd8, !lite, -1, 50);
            Hit step{nx, s};
            if (det.score > step.score) step = det;
            cur = step.x;
            if (step.score > best.score) best = step;
        }
    }
    return best;
}

// ---- rescue schedules: rescue.h:12-46 -----------------------------------
std::vector<int> rescue_lengths(int plen, int nlen, bool dense) {
    const int ml = std::min(plen, nlen);
    std::vector<int> out;
    if (ml < 4) return out;
    if (dense) {
        const int c[9] = {ml, std::min(100, ml / 2), std::min(32, ml), std::min(24, ml),
                          std::min(20, ml / 3), 16, 12, 8, 4};
        for (int x : c)
            if (x >= 4 && std::find(out.begin(), out.end(), x) == out.end()) out.push_back(x);
    } else {
        const int a = std::min(20, ml / 3), b = std::min(100, ml / 2);
        if (a >= 4) out.push_back(a);
        if (b >= 4 && b != a) out.push_back(b);
    }
    return out;
}

int rescue_stride(int len, bool dense) {
    int j;
    if (dense) {
        j = len > 350 ? 50 : len > 300 ? 30 : len > 200 ? 10 : 5;
        return std::max(1, std::min(j, std::max(1, len / 4)));
    }
    j = len > 250 ? 45 : len > 200 ? 35 : len > 150 ? 25 : 15;
    return std::max(1, std::min(j, std::max(1, len / 3)));
}

// ---- local fragment DP rescue: rescue.h:50-117 --------------------------
Hit local_fragment_dp(const Pair &pr, bool dense) {
    Hit best = hit_of(kNeg);
    const int pl = pr.plen(), nl = pr.nlen();
    const std::vector<int> lens = rescue_lengths(pl, nl, dense);
    if (lens.empty()) return best;
    const int jp = rescue_stride(pl, dense), jn = rescue_stride(nl, dense);
    for (int len : lens) {
        const int pmax = pl - len + 1, nmax = nl - len + 1;
        if (pmax <= 0 || nmax <= 0) continue;
        for (int ps = 0; ps < pmax; ps += jp) {
            for (int ns = 0; ns < nmax; ns += jn) {
                std::vector<double> A(3 * (size_t)len), B(3 * (size_t)len);
                for (int k = 0; k < len; k++) {
                    std::memcpy(&A[3 * k], pr.P + 3 * pr.pmap[ps + k], 24);
                    std::memcpy(&B[3 * k], pr.Q + 3 * pr.nmap[ns + k], 24);
                }
                std::vector<uint8_t> ones(len, 1);
                const Xf x0 = superpose(A.data(), B.data(), ones.data(), len);
                const Hit h = dp_refine(pr, x0, 100.0, 2, true);
                if (h.score > best.score) best = h;
            }
        }
    }
    return best;
}

// ---- threading DP rescue: rescue.h:121-190 ------------------------------
Hit threading_dp(const Pair &pr) {
    Hit best = hit_of(kNeg);
    const int pl = pr.plen(), nl = pr.nlen();
    if (std::min(pl, nl) < 3) return best;
    const int step = std::max(1, std::min(pl, nl) / 50);
    for (int off = -(nl - 3); off < pl - 2; off += step) {
        const int ps = std::max(off, 0), ns = std::max(-off, 0);
        const int ov = std::min(pl - ps, nl - ns);
        if (ov < 3) continue;
        std::vector<int> ap(ov), an(ov);
        for (int k = 0; k < ov; k++) {
            ap[k] = pr.pmap[ps + k];
            an[k] = pr.nmap[ns + k];
        }
        const Hit det = detailed_search(pr, ap.data(), an.data(), ov, pr.d0s, 100.0, true, 8, 5);
        const Hit dp = dp_refine(pr, det.x, 100.0, 1, true);
        Hit cand = dp.score > det.score ? dp : det;
        cand.score = std::max(det.score, dp.score);
        if (cand.score > best.score) best = cand;
    }
    return best;
}

// ---- secondary structure from C1' geometry: ss.h:16-52 ------------------
void assign_ss(const double *X, const uint8_t *v, int N, char *ss) {
    for (int i = 0; i < N; i++) ss[i] = 'L';
    auto dist2 = [&](int a, int b) {
        const double dx = X[3 * a] - X[3 * b], dy = X[3 * a + 1] - X[3 * b + 1],
                     dz = X[3 * a + 2] - X[3 * b + 2];
        return dx * dx + dy * dy + dz * dz;
    };
    const double lo = 81.0, hi = 144.0;
    std::vector<int> mate(N, -1);
    for (int i = 0; i + 4 < N; i++) {
        if (!v[i] || !v[i + 1]) continue;
        for (int j = N - 1; j > i + 3; j--) {
            if (!v[j] || !v[j - 1]) continue;
            const double a = dist2(i, j), b = dist2(i + 1, j - 1);
            if (a >= lo && a <= hi && b >= lo && b <= hi && mate[i] < 0 && mate[j] < 0) {
                mate[i] = j;
                mate[j] = i;
            }
        }
    }
    for (int i = 0; i < N; i++)
        if (mate[i] >= 0) ss[i] = i < mate[i] ? '<' : '>';
}

// NW over a compact score matrix, aligned pairs mapped back to full indices.
int compact_nw(const Pair &pr, const std::vector<double> &S, double gap, int *ap, int *an) {
    std::vector<double> H;
    std::vector<int8_t> T;
    const int pl = pr.plen(), nl = pr.nlen();
    nw_fill(S.data(), pl, nl, gap, H, T);
    const int na = nw_walk(T, pl, nl, ap, an);
    for (int k = 0; k < na; k++) {
        ap[k] = pr.pmap[ap[k]];
        an[k] = pr.nmap[an[k]];
    }
    return na;
}

// ---- SS seeds (IA2 then IA4): ss.h:56-164 -------------------------------
Hit ss_seeds(const Pair &pr, const char *sp, const char *sn, const Xf &x0, int dp_iter) {
    Hit best = hit_of(kNeg);
    const int pl = pr.plen(), nl = pr.nlen();
    std::vector<int> ap(pr.N), an(pr.N);
    if (pl >= 3 && nl >= 3) {
        // IA2: secondary-structure identity, gap -1 (ss.h:56-83, 94-106)
        std::vector<double> S((size_t)pl * nl);
        for (int a = 0; a < pl; a++)
            for (int b = 0; b < nl; b++)
                S[(size_t)a * nl + b] = sp[pr.pmap[a]] == sn[pr.nmap[b]] ? 1.0 : 0.0;
        const int na = compact_nw(pr, S, -1.0, ap.data(), an.data());
        if (na >= 3) {
            const Hit det = detailed_search(pr, ap.data(), an.data(), na, pr.d0s, pr.sd8, true, -1, 50);
            if (det.score > best.score) best = det;
            if (dp_iter > 0) {
                const Hit dp = dp_refine(pr, best.x, pr.sd8, dp_iter, false);
                if (dp.score > best.score) best = dp;
            }
        }
        // IA4: TM-distance at d01 plus 0.5 for equal SS, gap -1 (ss.h:108-153)
        const Xf start = best.score > -1e20 ? best.x : x0;
        const double d01 = std::max(pr.d0 + 1.5, pr.d0s), d01sq = d01 * d01;
        for (int a = 0; a < pl; a++) {
            const int i = pr.pmap[a];
            for (int b = 0; b < nl; b++) {
                const int j = pr.nmap[b];
                double s = 1.0 / (1.0 + moved_d2(start, pr.P, i, pr.Q, j) / d01sq);
                if (sp[i] == sn[j]) s += 0.5;
                S[(size_t)a * nl + b] = s;
            }
        }
        const int nb = compact_nw(pr, S, -1.0, ap.data(), an.data());
        if (nb >= 3) {
            const Hit det = detailed_search(pr, ap.data(), an.data(), nb, pr.d0s, pr.sd8, true, -1, 50);
            if (det.score > best.score) best = det;
            if (dp_iter > 0) {
                const Hit dp = dp_refine(pr, best.x, pr.sd8, dp_iter, false);
                if (dp.score > best.score) best = dp;
            }
        }
    }
    return best;
}

// ---- gapless threading: threading.h:41-80 -------------------------------
Hit gapless(const Pair &pr) {
    Hit best = hit_of(kNeg);
    const int N = pr.N;
    int pl = 0, nl = 0;
    for (int i = 0; i < N; i++) {
        pl += pr.pv[i];
        nl += pr.nv[i];
    }
    const int min_ali = std::max(3, std::min(pl, nl) / 2);
    const int kr = N - min_ali;
    if (kr <= 0) return best;
    const int stride = std::max(1, (2 * kr + 99) / 100);
    std::vector<double> A(3 * (size_t)N), B(3 * (size_t)N);
    std::vector<uint8_t> ones(N, 1), sel(N);
    for (int k = -kr; k <= kr; k += stride) {
        if (k == 0) continue;
        int na = 0;
        for (int i = 0; i < N; i++) {
            const int j = i - k;
            if (j < 0 || j >= N || !pr.pv[i] || !pr.nv[j]) continue;
            std::memcpy(&A[3 * na], pr.P + 3 * i, 24);
            std::memcpy(&B[3 * na], pr.Q + 3 * j, 24);
            na++;
        }
        if (na < min_ali) continue;
        Xf x = superpose(A.data(), B.data(), ones.data(), na);
        cutoff_refine(pr.P, pr.Q, pr.v, pr.d0s, N, 5, x, sel);
        const double s = pr.same_index(x);
        if (s > best.score) best = Hit{x, s};
    }
    return best;
}

// Kabsch on the offset pairing pred[i] <-> native[i-k] (indialign.cpp:236-247).
bool offset_superpose(const Pair &pr, int k, Xf &x) {
    std::vector<double> A, B;
    int na = 0;
    for (int i = 0; i < pr.N; i++) {
        const int j = i - k;
        if (j < 0 || j >= pr.N || !pr.pv[i] || !pr.nv[j]) continue;
        A.insert(A.end(), pr.P + 3 * i, pr.P + 3 * i + 3);
        B.insert(B.end(), pr.Q + 3 * j, pr.Q + 3 * j + 3);
        na++;
    }
    if (na < 3) return false;
    std::vector<uint8_t> ones(na, 1);
    x = superpose(A.data(), B.data(), ones.data(), na);
    return true;
}

void record(double *trace, int stage, const Hit &h) {
    if (trace) pack(h, trace + 13 * stage);
}

// ---- the pipeline: indialign.cpp:29-266 ---------------------------------
Hit search(const Pair &pr, int max_iter, bool use_frag, int dp_iter, double rs1, double rs2,
           double early_term, bool serial_bank, double *trace) {
    if (trace)
        for (int s = 0; s < ORACLE_N_STAGES; s++) {
            pack(hit_of(std::numeric_limits<double>::quiet_NaN()), trace + 13 * s);
        }
    if (pr.Lnorm <= 2.0) return hit_of(0.0);
    const int N = pr.N;
    // d0 search list: d0s x {2,1.5,1,.75,.5,.25} then d0 + {1..5} unless
    // within 0.01 of an earlier entry (indialign.cpp:54-67).
    static const double mult[6] = {2.0, 1.5, 1.0, 0.75, 0.5, 0.25};
    double cut[11];
    int C = 0;
    for (double m : mult) cut[C++] = pr.d0s * m;
    for (int a = 1; a <= 5; a++) {
        const double v = pr.d0 + a;
        bool dup = false;
        for (int e = 0; e < C && !dup; e++) dup = std::fabs(cut[e] - v) < 0.01;
        if (!dup) cut[C++] = v;
    }
    std::vector<uint8_t> masks;
    seed_masks(N, pr.v, use_frag, 50, masks);
    anchor_seeds(pr.P, pr.Q, pr.v, N, pr.d0s, pr.d0, 5, masks);
    const int K = (int)(masks.size() / (size_t)N);
    if (K == 0) return hit_of(0.0);

    Hit best = seed_bank(pr.P, pr.Q, pr.v, masks.data(), K, cut, C, pr.d0, pr.sd8, pr.Lnorm, N,
                         max_iter, false, serial_bank ? early_term : 1e300);
    record(trace, 0, best);
    auto rescored = [&](Hit h) {
        h.score = pr.same_index(h.x);
        return h;
    };

    std::vector<char> sp(N), sn(N);
    assign_ss(pr.P, pr.pv, N, sp.data());
    assign_ss(pr.Q, pr.nv, N, sn.data());
    const Hit ss = rescored(ss_seeds(pr, sp.data(), sn.data(), best.x, dp_iter));
    if (ss.score > best.score) best = ss;
    record(trace, 1, best);

    const Hit gt = gapless(pr);
    if (gt.score > best.score) best = gt;
    record(trace, 2, best);

    bool cross = false;
    for (int i = 0; i < N && !cross; i++) cross = pr.pv[i] != pr.nv[i];
    if (best.score < rs1) {
        const Hit h = rescored(local_fragment_dp(pr, false));
        if (h.score > best.score) best = h;
        record(trace, 3, best);
    }
    if (best.score < rs1) {
        const Hit h = rescored(threading_dp(pr));
        if (h.score > best.score) best = h;
        record(trace, 4, best);
    }
    if (best.score < rs2) {
        const Hit h = rescored(local_fragment_dp(pr, true));
        if (h.score > best.score) best = h;
        record(trace, 5, best);
    }

    {  // top-k distance seed (indialign.cpp:154-177)
        std::vector<double> d2(N);
        for (int i = 0; i < N; i++) d2[i] = pr.v[i] ? moved_d2(best.x, pr.P, i, pr.Q, i) : 1e12;
        std::vector<uint8_t> tk(N);
        topk_mask(d2.data(), pr.v, N, 0.5, tk.data());
        int cnt = 0;
        for (int i = 0; i < N; i++) cnt += tk[i];
        if (cnt >= 3) {
            const Hit tp = seed_bank(pr.P, pr.Q, pr.v, tk.data(), 1, cut, C, pr.d0, pr.sd8,
                                     pr.Lnorm, N, max_iter, false, early_term);
            if (tp.score > best.score) best = tp;
        }
        record(trace, 6, best);
    }

    auto offer = [&](const Xf &x, double s) {
        if (s > best.score) best = Hit{x, s};
    };
    if (dp_iter > 0) {  // indialign.cpp:188-223
        const double seed_score = best.score;
        const Hit dp = dp_refine(pr, best.x, pr.sd8, dp_iter, false);
        const double dp_id = pr.same_index(dp.x);
        offer(dp.x, dp_id);
        offer(best.x, pr.same_index(best.x));
        if (cross) {
            const Hit idp = dp_refine(pr, identity_xf(), pr.sd8, dp_iter, true);
            offer(idp.x, pr.same_index(idp.x));
            const int extra = std::max(0, 5 - dp_iter);
            if (extra > 0 && best.score < 0.36) {
                const Xf from = seed_score > dp_id ? best.x : dp.x;
                const Hit ex = dp_refine(pr, from, pr.sd8, extra, true);
                offer(ex.x, pr.same_index(ex.x));
            }
        }
        record(trace, 7, best);
        // raw cross-offset starts (indialign.cpp:229-256)
        const int stride = std::max(1, N / 10);
        for (int k = -N / 2; k <= N / 2; k += stride) {
            if (k == 0) continue;
            Xf x0;
            if (!offset_superpose(pr, k, x0)) continue;
            const Hit h = dp_refine(pr, x0, pr.sd8, dp_iter, true);
            offer(h.x, pr.same_index(h.x));
        }
        record(trace, 8, best);
    } else {
        offer(best.x, pr.same_index(best.x));
        record(trace, 7, best);
    }
    const Hit w = weighted_refine(pr.P, pr.Q, pr.v, pr.d0, pr.Lnorm, N, 4, best.x);
    offer(w.x, w.score);
    record(trace, 9, best);
    return best;
}

}  // namespace

extern "C" {

void oracle_default_config(OracleConfig *c) {
    c->backend_mode = 0;
    c->use_fragment_search = 1;
    c->max_iter = 20;
    c->dp_iter = 1;
    c->max_mem_gb = 20.0;
    c->request_cuda = 0;
    c->rescue_score_1 = 0.5;
    c->rescue_score_2 = 0.3;
    c->early_term_score = 0.99;
}

int oracle_search_traced(const OraclePair *p, const OracleConfig *cfg, OracleResult *out,
                         int serial_bank, double *stage_trace) {
    if (p->length <= 0) return -1;
    Pair pr(p->length, p->pred_xyz, p->native_xyz, p->valid_mask, p->pred_valid_mask,
            p->native_valid_mask, p->lnorm);
    const Hit h = search(pr, cfg->max_iter, cfg->use_fragment_search != 0, cfg->dp_iter,
                         cfg->rescue_score_1, cfg->rescue_score_2, cfg->early_term_score,
                         serial_bank != 0, stage_trace);
    out->score = h.score;
    std::memcpy(out->rotation, h.x.R, 72);
    std::memcpy(out->translation, h.x.t, 24);
    return 0;
}

int oracle_search_batch(int n, const OraclePair *p, const OracleConfig *cfg, OracleResult *out,
                        int serial_bank) {
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n; i++) oracle_search_traced(&p[i], cfg, &out[i], serial_bank, nullptr);
    return 0;
}

void oracle_d0(double L, double *o) { d0_params(L, o[0], o[1], o[2]); }

// Correctly rounded division by a loop-invariant divisor from its correctly
// rounded reciprocal and one FMA correction (Markstein), the form the GPU
// kernels use for d2 / d0^2.  Counts the cases, among n random ones, where it
// differs from a / b; the draws cover d0^2 of the TM-score formula,
// arbitrary mantissas over 2^-4..2^11 and quotients over 2^-60..2^14.
long long oracle_markstein_mismatches(long long n, unsigned long long seed) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long long bad = 0;
    for (long long k = 0; k < n; k++) {
        double a, b;
        switch (k % 4) {
            case 0: {
                double d0, d0s, sd8;
                d0_params(3.0 + U(rng) * 3000.0, d0, d0s, sd8);
                b = d0 * d0;
                a = U(rng) * 1e4;
                break;
            }
            case 1: {
                uint64_t u = (rng() & 0x000FFFFFFFFFFFFFull) | (0x3FFull << 52), v;
                std::memcpy(&b, &u, 8);
                v = (rng() & 0x000FFFFFFFFFFFFFull) | ((0x3F0ull + rng() % 40) << 52);
                std::memcpy(&a, &v, 8);
                b = std::ldexp(b, (int)(rng() % 16) - 4);
                break;
            }
            case 2:
                b = std::ldexp(1.0 + std::ldexp((double)(rng() % 1024), -52), (int)(rng() % 10) - 3);
                a = U(rng) * 1e3;
                break;
            default:
                b = 0.25 + U(rng) * 400.0;
                a = std::ldexp(U(rng), -(int)(rng() % 60));
        }
        const double inv = 1.0 / b;
        const double q0 = a * inv;
        const double q = std::fma(std::fma(-b, q0, a), inv, q0);
        bad += q != a / b;
    }
    return bad;
}

void oracle_jacobi4(double *A16, double *V16) {
    double a[4][4], v[4][4];
    std::memcpy(a, A16, sizeof a);
    jacobi_sym4(a, v);
    std::memcpy(A16, a, sizeof a);
    std::memcpy(V16, v, sizeof v);
}

void oracle_cov_to_rotation(const double *H9, double *R9) {
    double h[3][3];
    std::memcpy(h, H9, sizeof h);
    horn_rotation(h, R9);
}

void oracle_kabsch(const double *P, const double *Q, const uint8_t *m, int N, double *R,
                   double *t) {
    const Xf x = superpose(P, Q, m, N);
    std::memcpy(R, x.R, 72);
    std::memcpy(t, x.t, 24);
}

double oracle_tm_score(const double *p, const double *q, const uint8_t *v, const double *R,
                       const double *t, double d0, double Lnorm, double sd8, int N,
                       int use_d8) {
    Xf x;
    std::memcpy(x.R, R, 72);
    std::memcpy(x.t, t, 24);
    return tm_eval(p, q, v, x, d0, Lnorm, sd8, N, use_d8 != 0);
}

int oracle_seed_masks(int N, const uint8_t *valid, int use_frag, int max_starts, uint8_t *out,
                      int cap) {
    std::vector<uint8_t> m;
    seed_masks(N, valid, use_frag != 0, max_starts, m);
    const int K = (int)(m.size() / (size_t)N);
    std::memcpy(out, m.data(), (size_t)std::min(K, cap) * N);
    return K;
}

int oracle_all_seeds(const double *pred, const double *native, const uint8_t *valid, int N,
                     double Lnorm, uint8_t *out, int cap, int *n_frag) {
    double d0, d0s, sd8;
    d0_params(Lnorm, d0, d0s, sd8);
    std::vector<uint8_t> m;
    seed_masks(N, valid, true, 50, m);
    *n_frag = (int)(m.size() / (size_t)N);
    anchor_seeds(pred, native, valid, N, d0s, d0, 5, m);
    const int K = (int)(m.size() / (size_t)N);
    std::memcpy(out, m.data(), (size_t)std::min(K, cap) * N);
    return K;
}

void oracle_topk_seed(const double *d2, const uint8_t *valid, int N, double frac,
                      uint8_t *mask) {
    topk_mask(d2, valid, N, frac, mask);
}

void oracle_refine_seed(const double *p, const double *q, const uint8_t *v, const uint8_t *seed,
                        double d0s, int N, int max_iter, double *R, double *t, int *n_passes) {
    std::vector<uint8_t> ws(N), sel(N);
    const Xf x = refine_seed(p, q, v, seed, d0s, N, max_iter, ws, sel, n_passes);
    std::memcpy(R, x.R, 72);
    std::memcpy(t, x.t, 24);
}

void oracle_seed_bank(const double *p, const double *q, const uint8_t *v, const uint8_t *seeds,
                      int K, const double *d0c, int C, double d0, double sd8, double Lnorm,
                      int N, int max_iter, int use_d8, double early_term, double *out13) {
    pack(seed_bank(p, q, v, seeds, K, d0c, C, d0, sd8, Lnorm, N, max_iter, use_d8 != 0,
                   early_term),
         out13);
}

void oracle_nw_dp(const double *smat, int Np, int Nn, double gap, double *H, int8_t *trace) {
    std::vector<double> h;
    std::vector<int8_t> t;
    nw_fill(smat, Np, Nn, gap, h, t);
    std::memcpy(H, h.data(), h.size() * sizeof(double));
    std::memcpy(trace, t.data(), t.size());
}

int oracle_nw_align(const double *smat, int Np, int Nn, double gap, int *ap, int *an) {
    std::vector<double> h;
    std::vector<int8_t> t;
    nw_fill(smat, Np, Nn, gap, h, t);
    return nw_walk(t, Np, Nn, ap, an);
}

static Pair pair_view(int N, const double *p, const double *q, const uint8_t *v,
                      const uint8_t *pv, const uint8_t *nv, double Lnorm) {
    return Pair(N, p, q, v, pv ? pv : v, nv ? nv : v, Lnorm);
}

void oracle_detailed_search(const double *p, const double *q, const int *ap, const int *an,
                            int na, double d0, double d0s, double sd8, double Lnorm,
                            int use_frag, int mi_override, int max_starts, double *out13) {
    // Only coordinates, d0, d0s and Lnorm enter align.h:15-50.
    int N = 0;
    for (int k = 0; k < na; k++) N = std::max(N, std::max(ap[k], an[k]) + 1);
    std::vector<uint8_t> ones(std::max(N, 1), 1);
    Pair pr = pair_view(std::max(N, 1), p, q, ones.data(), nullptr, nullptr, Lnorm);
    pr.d0 = d0;
    pr.d0s = d0s;
    pack(detailed_search(pr, ap, an, na, d0s, sd8, use_frag != 0, mi_override, max_starts),
         out13);
}

void oracle_dp_refine(const double *p, const double *q, const uint8_t *v, const uint8_t *pv,
                      const uint8_t *nv, const double *R0, const double *t0, double d0,
                      double d0s, double sd8, double Lnorm, int N, int max_iter, int lite,
                      double *out13) {
    Pair pr = pair_view(N, p, q, v, pv, nv, Lnorm);
    pr.d0 = d0;
    pr.d0s = d0s;
    Xf x0;
    std::memcpy(x0.R, R0, 72);
    std::memcpy(x0.t, t0, 24);
    pack(dp_refine(pr, x0, sd8, max_iter, lite != 0), out13);
}

void oracle_local_fragment_dp(const double *p, const double *q, const uint8_t *v,
                              const uint8_t *pv, const uint8_t *nv, double d0, double d0s,
                              double Lnorm, int N, int dense, double *out13) {
    Pair pr = pair_view(N, p, q, v, pv, nv, Lnorm);
    pr.d0 = d0;
    pr.d0s = d0s;
    pack(local_fragment_dp(pr, dense != 0), out13);
}

void oracle_threading_dp(const double *p, const double *q, const uint8_t *v, const uint8_t *pv,
                         const uint8_t *nv, double d0, double d0s, double Lnorm, int N,
                         double *out13) {
    Pair pr = pair_view(N, p, q, v, pv, nv, Lnorm);
    pr.d0 = d0;
    pr.d0s = d0s;
    pack(threading_dp(pr), out13);
}

void oracle_assign_ss(const double *xyz, const uint8_t *v, int N, char *ss) {
    assign_ss(xyz, v, N, ss);
}

void oracle_ss_seeds(const double *p, const double *q, const uint8_t *v, const uint8_t *pv,
                     const uint8_t *nv, const double *R0, const double *t0, double d0,
                     double d0s, double sd8, double Lnorm, int N, int dp_iter, double *out13) {
    Pair pr = pair_view(N, p, q, v, pv, nv, Lnorm);
    pr.d0 = d0;
    pr.d0s = d0s;
    pr.sd8 = sd8;
    std::vector<char> sp(N), sn(N);
    assign_ss(p, pr.pv, N, sp.data());
    assign_ss(q, pr.nv, N, sn.data());
    Xf x0;
    std::memcpy(x0.R, R0, 72);
    std::memcpy(x0.t, t0, 24);
    pack(ss_seeds(pr, sp.data(), sn.data(), x0, dp_iter), out13);
}

void oracle_gapless(const double *p, const double *q, const uint8_t *pv, const uint8_t *nv,
                    const uint8_t *v, double d0, double d0s, double Lnorm, int N,
                    double *out13) {
    Pair pr = pair_view(N, p, q, v, pv, nv, Lnorm);
    pr.d0 = d0;
    pr.d0s = d0s;
    pack(gapless(pr), out13);
}

void oracle_weighted_refine(const double *p, const double *q, const uint8_t *v, double d0,
                            double Lnorm, int N, int iters, double *R, double *t, double *sc) {
    Xf x;
    std::memcpy(x.R, R, 72);
    std::memcpy(x.t, t, 24);
    const Hit h = weighted_refine(p, q, v, d0, Lnorm, N, iters, x);
    std::memcpy(R, h.x.R, 72);
    std::memcpy(t, h.x.t, 24);
    *sc = h.score;
}

}  // extern "C"
