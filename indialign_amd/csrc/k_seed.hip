// Seed construction on the device: anchor contact seeds (seeds.h:164-283)
// and C1'-geometry secondary structure (ss.h:16-52).  One workgroup per pair.
#include "ia_kernels.h"
#include "stl_sort.h"

#include <math.h>

namespace ia {

namespace {

constexpr int kMaxSeen = 1024;  // existing seeds + anchors hashed per pair
constexpr int kMaxAnchors = 5;  // indialign.cpp:74
// Pairs of up to this many residues keep the per-residue counts, the sorted
// candidate list (12 KB) and the current neighbourhood mask (1.5 KB) in LDS:
// the candidate build, the sort and the mask hash are one thread's chains of
// dependent accesses, which global memory would serve at L2 latency.
constexpr int kLdsCand = 1536;

struct CountIdx {
    int count, a;
};
struct ByCountDesc {
    IA_HD bool operator()(const CountIdx &x, const CountIdx &y) const { return x.count > y.count; }
};

IA_D uint64_t fnv_step(uint64_t h, uint8_t b) {
    h ^= b;
    h *= 1099511628211ULL;
    return h;
}

// Pair (i, j) counts as a consistent contact at radius r: the reference
// visits j only inside the 27 grid cells of size r around i (cell index =
// floor(x / r) per axis), then requires pd2 <= r2, nd2 <= r2 and
// |pd - nd| <= tol (seeds.h:206-229).  The tests are a pure conjunction, so
// the cheap distance test runs first and only the pairs within r pay for
// the cell indices' divisions.
IA_D bool contact(const Arena &A, int off, int i, int j, double r, double r2, double tol) {
    const double pix = A.px[off + i], piy = A.py[off + i], piz = A.pz[off + i];
    const double pjx = A.px[off + j], pjy = A.py[off + j], pjz = A.pz[off + j];
    const double pdx = pix - pjx, pdy = piy - pjy, pdz = piz - pjz;
    const double pd2 = pdx * pdx + pdy * pdy + pdz * pdz;
    if (pd2 > r2) return false;
    const long long cx = (long long)floor(pix / r) - (long long)floor(pjx / r);
    const long long cy = (long long)floor(piy / r) - (long long)floor(pjy / r);
    const long long cz = (long long)floor(piz / r) - (long long)floor(pjz / r);
    if (cx < -1 || cx > 1 || cy < -1 || cy > 1 || cz < -1 || cz > 1) return false;
    const double ndx = A.qx[off + i] - A.qx[off + j], ndy = A.qy[off + i] - A.qy[off + j],
                 ndz = A.qz[off + i] - A.qz[off + j];
    const double nd2 = ndx * ndx + ndy * ndy + ndz * ndz;
    if (nd2 > r2) return false;
    return fabs(sqrt(pd2) - sqrt(nd2)) <= tol;
}

__global__ __launch_bounds__(256) void anchor_kernel(Arena A, const int *pairs,
                                                     const SeedDesc *frag, const int *frag_off,
                                                     const int *frag_cnt, uint8_t *amask,
                                                     const long long *aoff, int *acnt,
                                                     int *scratch, const long long *soff) {
    const int p = pairs[blockIdx.x];
    const PairInfo &pi = A.info[p];
    const int N = pi.N, off = pi.off, tid = threadIdx.x;
    int *vidx = scratch + soff[blockIdx.x];          // V
    int *counts = vidx + N;                          // V
    CountIdx *cand = (CountIdx *)(counts + N);       // V (2 ints each)
    uint8_t *out = amask + aoff[blockIdx.x];
    __shared__ uint64_t seen[kMaxSeen];
    __shared__ CountIdx lcand[kLdsCand];
    __shared__ uint32_t lmask4[kLdsCand / 4];
    uint8_t *lmask = (uint8_t *)lmask4;
    // dynamic: the valid residues' pred coordinates, 3 doubles each, sized by
    // the launch for min(longest pair, kLdsCand) residues
    extern __shared__ double lpts[];
    __shared__ double radii[6];
    __shared__ int nr, V, nseen, ncand, added, total, idx, dup;
    __shared__ uint64_t hcur;
    if (tid == 0) {
        // radii d0s x {1, 1.5, 2} then d0 + {1, 2, 3} unless within 0.01
        nr = 0;
        radii[nr++] = pi.d0s;
        radii[nr++] = pi.d0s * 1.5;
        radii[nr++] = pi.d0s * 2.0;
        for (int e = 1; e <= 3; e++) {
            const double v = pi.d0 + (double)e;
            bool d = false;
            for (int j = 0; j < nr && !d; j++) d = fabs(radii[j] - v) < 0.01;
            if (!d) radii[nr++] = v;
        }
        int c = 0;
        for (int i = 0; i < N; i++)
            if (A.v[off + i]) vidx[c++] = i;
        V = c;
        nseen = frag_cnt[blockIdx.x];
        total = 0;
    }
    __syncthreads();
    if (V < 3) {
        if (tid == 0) acnt[blockIdx.x] = 0;
        return;
    }
    // hashes of the existing masks: window [a, a+b) carries valid bytes
    for (int k = tid; k < nseen; k += 256) {
        const SeedDesc sd = frag[frag_off[blockIdx.x] + k];
        uint64_t h = 1469598103934665603ULL;
        for (int i = 0; i < N; i++)
            h = fnv_step(h, (i >= sd.a && i < sd.a + sd.b) ? A.v[off + i] : (uint8_t)0);
        seen[k] = h;
    }
    if (N <= kLdsCand) {
        for (int a = tid; a < V; a += 256) {
            const int i = vidx[a];
            lpts[3 * a] = A.px[off + i];
            lpts[3 * a + 1] = A.py[off + i];
            lpts[3 * a + 2] = A.pz[off + i];
        }
    }
    __syncthreads();
    for (int ri = 0; ri < nr; ri++) {
        const double r = radii[ri];
        if (r <= 0) continue;
        const double r2 = r * r, tol = fmax(0.5, r * 0.25);
        const bool in_lds = N <= kLdsCand;  // N >= V
        for (int a = tid; a < V; a += 256) {
            const int i = vidx[a];
            int c = 1;
            if (in_lds) {
                // contact()'s first test from the LDS copy (same operands
                // and operations); the rest only for pairs within r
                const double pix = A.px[off + i], piy = A.py[off + i], piz = A.pz[off + i];
                for (int b = 0; b < V; b++) {
                    const double pdx = pix - lpts[3 * b], pdy = piy - lpts[3 * b + 1],
                                 pdz = piz - lpts[3 * b + 2];
                    const double pd2 = pdx * pdx + pdy * pdy + pdz * pdz;
                    if (pd2 > r2) continue;
                    const int j = vidx[b];
                    if (j != i && contact(A, off, i, j, r, r2, tol)) c++;
                }
            } else {
                for (int b = 0; b < V; b++) {
                    const int j = vidx[b];
                    if (j != i && contact(A, off, i, j, r, r2, tol)) c++;
                }
            }
            if (in_lds) lcand[a] = CountIdx{c, a};
            else counts[a] = c;
        }
        if (!in_lds) __threadfence();  // LDS needs only the barrier
        __syncthreads();
        CountIdx *cl = in_lds ? lcand : cand;
        if (tid == 0) {
            // candidates with count >= 3 in residue order (compacted in
            // place in LDS: entry nc <= a is written after a is read)
            int nc = 0;
            for (int a = 0; a < V; a++) {
                const int c = in_lds ? lcand[a].count : counts[a];
                if (c >= 3) cl[nc++] = CountIdx{c, a};
            }
            ia_sort::sort(cl, cl + nc, ByCountDesc());
            ncand = nc;
            added = 0;
            idx = 0;
        }
        if (!in_lds) __threadfence();  // LDS needs only the barrier
        __syncthreads();
        while (idx < ncand && added < kMaxAnchors) {
            const int i = vidx[cl[idx].a];
            uint8_t *m = out + (long long)total * N;
            for (int j = tid; j < N; j += 256) {
                const uint8_t x = (j == i || (A.v[off + j] && contact(A, off, i, j, r, r2, tol))) ? 1 : 0;
                m[j] = x;
                if (in_lds) lmask[j] = x;
            }
            if (!in_lds) __threadfence();  // LDS needs only the barrier
            __syncthreads();
            if (tid == 0) {
                uint64_t h = 1469598103934665603ULL;
                if (in_lds) {
                    const int n4 = N >> 2;
                    for (int w = 0; w < n4; w++) {
                        const uint32_t x = lmask4[w];
                        h = fnv_step(h, (uint8_t)x);
                        h = fnv_step(h, (uint8_t)(x >> 8));
                        h = fnv_step(h, (uint8_t)(x >> 16));
                        h = fnv_step(h, (uint8_t)(x >> 24));
                    }
                    for (int j = n4 << 2; j < N; j++) h = fnv_step(h, lmask[j]);
                } else {
                    for (int j = 0; j < N; j++) h = fnv_step(h, m[j]);
                }
                hcur = h;
                dup = 0;
            }
            __syncthreads();
            for (int k = tid; k < nseen; k += 256)
                if (seen[k] == hcur) dup = 1;
            __syncthreads();
            if (tid == 0) {
                // the neighbourhood always holds count >= 3 residues
                if (!dup && nseen < kMaxSeen) {
                    seen[nseen++] = hcur;
                    added++;
                    total++;
                }
                idx++;
            }
            if (!in_lds) __threadfence();  // LDS needs only the barrier
            __syncthreads();
        }
    }
    if (tid == 0) acnt[blockIdx.x] = total;
}

// Greedy stem detection (ss.h:16-52): for i ascending, pair i with the
// largest j > i+3 such that (i,j) and (i+1,j-1) are 9-12 A apart and both
// are still unpaired.  Side 0 = pred with pred-valid, 1 = native with
// native-valid.  Writes 'L', '<', '>' by full index.
//
// The geometric test does not depend on the pairing state, so the block
// first finds, in parallel, the kSsCand largest qualifying j of every i
// (cand, count; a count above kSsCand means more exist); one thread then runs
// the greedy pass over those lists in LDS, rescanning only when all stored
// candidates of an i are already paired.
constexpr int kSsCand = 8;
constexpr int kSsLds = 8192;

IA_D bool stem_geom(const double *X, const double *Y, const double *Z, int off, int i, int j) {
    const double dx = X[off + i] - X[off + j], dy = Y[off + i] - Y[off + j],
                 dz = Z[off + i] - Z[off + j];
    const double a = dx * dx + dy * dy + dz * dz;
    const double ex = X[off + i + 1] - X[off + j - 1], ey = Y[off + i + 1] - Y[off + j - 1],
                 ez = Z[off + i + 1] - Z[off + j - 1];
    const double b = ex * ex + ey * ey + ez * ez;
    return a >= 81.0 && a <= 144.0 && b >= 81.0 && b <= 144.0;
}

__global__ __launch_bounds__(256) void ss_kernel(Arena A, char *ssp, char *ssn, int *scratch) {
    const int p = blockIdx.x >> 1, side = blockIdx.x & 1;
    const PairInfo &pi = A.info[p];
    const int N = pi.N, off = pi.off, tid = threadIdx.x;
    const double *X = side ? A.qx : A.px;
    const double *Y = side ? A.qy : A.py;
    const double *Z = side ? A.qz : A.pz;
    const uint8_t *vg = (side ? A.nv : A.pv) + off;
    char *ssg = (side ? ssn : ssp) + off;
    int *cand = scratch + ((long long)2 * off + (long long)side * N) * (kSsCand + 1);
    int *count = cand + (long long)N * kSsCand;
    __shared__ char ss_l[kSsLds];
    __shared__ uint8_t v_l[kSsLds];
    const bool lds = N <= kSsLds;
    char *ss = lds ? ss_l : ssg;
    const uint8_t *v = lds ? v_l : vg;
    for (int i = tid; i < N; i += 256) {
        ss[i] = 'L';
        if (lds) v_l[i] = vg[i];
    }
    // phase 1: candidate lists
    for (int i = tid; i + 4 < N; i += 256) {
        int k = 0;
        if (vg[i] && vg[i + 1]) {
            for (int j = N - 1; j > i + 3 && k <= kSsCand; j--) {
                if (!vg[j] || !vg[j - 1]) continue;
                if (!stem_geom(X, Y, Z, off, i, j)) continue;
                if (k < kSsCand) cand[(long long)i * kSsCand + k] = j;
                k++;
            }
        }
        count[i] = k;
    }
    __threadfence_block();
    __syncthreads();
    // phase 2: the greedy pass, in order
    if (tid == 0) {
        for (int i = 0; i + 4 < N; i++) {
            if (!v[i] || !v[i + 1] || ss[i] != 'L') continue;
            const int k = count[i];
            const int *c = cand + (long long)i * kSsCand;
            int best = -1;
            for (int q = 0; q < k && q < kSsCand; q++)
                if (ss[c[q]] == 'L') {
                    best = c[q];
                    break;
                }
            if (best < 0 && k > kSsCand) {
                for (int j = c[kSsCand - 1] - 1; j > i + 3; j--) {
                    if (!v[j] || !v[j - 1] || ss[j] != 'L') continue;
                    if (stem_geom(X, Y, Z, off, i, j)) {
                        best = j;
                        break;
                    }
                }
            }
            if (best >= 0) {
                ss[i] = '<';
                ss[best] = '>';
            }
        }
    }
    __syncthreads();
    if (lds)
        for (int i = tid; i < N; i += 256) ssg[i] = ss_l[i];
}

}  // namespace

void launch_ss(hipStream_t s, const Arena &A, int n_pairs, char *ssp, char *ssn, int *scratch) {
    if (n_pairs <= 0) return;
    ss_kernel<<<2 * n_pairs, 256, 0, s>>>(A, ssp, ssn, scratch);
}

void launch_anchors(hipStream_t s, const Arena &A, const int *pairs, int n,
                    const SeedDesc *frag_seeds, const int *frag_off, const int *frag_cnt,
                    uint8_t *anchor_masks, const long long *anchor_off, int *anchor_cnt,
                    int *scratch_i, const long long *scratch_off, int max_n) {
    if (n <= 0) return;
    const size_t lds = (size_t)3 * sizeof(double) * (size_t)min(max_n, kLdsCand);
    anchor_kernel<<<n, 256, lds, s>>>(A, pairs, frag_seeds, frag_off, frag_cnt, anchor_masks,
                                    anchor_off, anchor_cnt, scratch_i, scratch_off);
}

}  // namespace ia
