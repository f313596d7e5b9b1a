// Seed-bank refinement (K1) and the small per-pair kernels around it.
//
// K1 maps one candidate (seed k, d0 cutoff c) to one lane.  The 64 lanes of
// a wavefront share the point set, so the per-point loop is uniform (its
// coordinate loads are broadcast) while each lane keeps its own transform,
// inlier count and convergence state.  Every sum runs in index order inside
// the lane, which is the order of the reference loops, so a candidate's
// (R, t, score) equals the FP64 CPU result bit for bit.
#include "ia_kernels.h"
#include "wave.h"

#include <math.h>

namespace ia {

namespace {

// Point i of the bank's point set: one 64-byte load for full-index and list
// points (a pre-gathered list copy), two 32-byte loads for compact windows;
// the address is wave-uniform, so these are scalar loads.  Returns false for
// an invalid residue (PTS_FULL only).
template <int PTS>
IA_D bool get_pt(const Arena &A, int off, const P8 *lq, int a0, int a1, int i, double &x,
                 double &y, double &z, double &qa, double &qb, double &qc) {
    P4 p, q;
    if (PTS == PTS_FULL) {
        const P8 r = A.pq8[off + i];
        p = r.p;
        q = r.q;
    } else if (PTS == PTS_LIST || PTS == PTS_TLIST) {
        const P8 r = lq[i];
        p = r.p;
        q = r.q;
    } else {
        p = A.pc4[off + a0 + i];
        q = A.qc4[off + a1 + i];
    }
    x = p.x; y = p.y; z = p.z;
    qa = q.x; qb = q.y; qc = q.z;
    // full-index points: p.w is the valid flag; PyTorch-engine list records:
    // p.w is the entry's flag (0 for masked entries)
    return PTS == PTS_FULL || PTS == PTS_TLIST ? p.w != 0.0 : true;
}

IA_D int bank_points(const BankDesc &B, const int *list_na) {
    return B.pts == PTS_LIST || B.pts == PTS_TLIST ? list_na[B.a0] : B.n;
}

IA_D int bank_seeds(const BankDesc &B, int n, const int *list_na) {
    if (B.seeds == SEED_EXPLICIT) return B.K;
    if (B.pts == PTS_TLIST) return torch_list_seed_count(B.n, list_na[B.a1], B.max_starts);
    return implicit_seed_count(n, B.max_starts);
}

// Seeds that run: all, or the screened subset.
IA_D int bank_live_seeds(const BankDesc &B, int n, const int *list_na) {
    const int k = bank_seeds(B, n, list_na);
    return B.sel_m > 0 && B.sel_m < k ? B.sel_m : k;
}

// Refinement state of a candidate stopped at the bank's pass cap: the count
// of its last pass and the index of its next pass (its R, t go to its
// candidate slot).
struct ContState {
    int last, it;
};

// Passes it .. it_end-1 of the d0s-cutoff refinement (search.h:38-55): select
// the points within the cutoff, re-superpose on >= 3, stop when the count
// repeats.  Returns true when the candidate's refinement has finished (the
// count repeated or all B.iters passes ran), false when it stopped at it_end.
// Selection bits of the plain refinement's first pass, points 0..kSelCache-1:
// word w of thread t at sel_cache[w * 256 + t].  One 16 KB array per
// 256-thread block, shared by every point-set variant of refine_passes (six
// such blocks, the plain kernel's register-bound occupancy, use 96 KB of a
// CU's 160 KB of LDS; a 1024-point cache costs occupancy on short pairs).
constexpr int kSelCache = 512;
__shared__ unsigned sel_cache[(kSelCache / 32) * 256];

template <int PTS, bool FAST>
IA_D bool refine_passes(const Arena &A, const BankDesc &B, int n, int c, const P8 *lq, double *R,
                        double *t, int &it, int &last, int it_end, int &kops) {
    const int off = A.info[B.pair].off;
    const double cut = B.cut[c];
    const double lim = cut * cut;
    KabschSeq ks;
    for (; it < it_end; it++) {
        ks.reset();
        int cnt = 0;
        if (FAST) {  // one pass: selection + raw moments
            for (int i = 0; i < n; i++) {
                double x, y, z, qa, qb, qc;
                if (!get_pt<PTS>(A, off, lq, B.a0, B.a1, i, x, y, z, qa, qb, qc)) continue;
                if (dist2_moved_fma(R, t, x, y, z, qa, qb, qc) <= lim) {
                    cnt++;
                    ks.add_moment(x, y, z, qa, qb, qc);
                }
            }
            if (cnt >= 3) {
                ks.finish_moments();
                ks.template solve<FAST>(R, t);
            }
        } else {
            // two passes (centroids, then centred covariance); the first
            // records its selection of points 0..m-1 so the second
            // recomputes distances only for points m..n-1
            const int m = n < kSelCache ? n : kSelCache;
            unsigned word = 0;
            for (int i = 0; i < m; i++) {
                double x, y, z, qa, qb, qc;
                bool in = false;
                if (get_pt<PTS>(A, off, lq, B.a0, B.a1, i, x, y, z, qa, qb, qc) &&
                    dist2_moved(R, t, x, y, z, qa, qb, qc) <= lim) {
                    in = true;
                    cnt++;
                    ks.add_centroid(x, y, z, qa, qb, qc);
                }
                word |= (in ? 1u : 0u) << (i & 31);
                if ((i & 31) == 31 || i == m - 1) {
                    sel_cache[(i >> 5) * 256 + threadIdx.x] = word;
                    word = 0;
                }
            }
            for (int i = m; i < n; i++) {
                double x, y, z, qa, qb, qc;
                if (!get_pt<PTS>(A, off, lq, B.a0, B.a1, i, x, y, z, qa, qb, qc)) continue;
                if (dist2_moved(R, t, x, y, z, qa, qb, qc) <= lim) {
                    cnt++;
                    ks.add_centroid(x, y, z, qa, qb, qc);
                }
            }
            if (cnt >= 3) {
                ks.centre();
                for (int i = 0; i < m; i++) {
                    double x, y, z, qa, qb, qc;
                    if ((i & 31) == 0) word = sel_cache[(i >> 5) * 256 + threadIdx.x];
                    get_pt<PTS>(A, off, lq, B.a0, B.a1, i, x, y, z, qa, qb, qc);
                    if ((word >> (i & 31)) & 1u) ks.add_cov(x, y, z, qa, qb, qc);
                }
                for (int i = m; i < n; i++) {
                    double x, y, z, qa, qb, qc;
                    if (!get_pt<PTS>(A, off, lq, B.a0, B.a1, i, x, y, z, qa, qb, qc)) continue;
                    if (dist2_moved(R, t, x, y, z, qa, qb, qc) <= lim) ks.add_cov(x, y, z, qa, qb, qc);
                }
                ks.template solve<FAST>(R, t);
            }
        }
        if (cnt >= 3) kops += 31 * cnt + 3400;  // centroid + covariance + solve
        if (cnt == last && !B.fixed_iters) {
            it++;
            return true;
        }
        last = cnt;
    }
    return it >= B.iters;
}

// TM-score of (R, t) over the bank's point set (core.h:134-166).
template <int PTS, bool FAST>
IA_D double candidate_score(const Arena &A, const BankDesc &B, int n, const P8 *lq,
                            const double *R, const double *t) {
    const PairInfo &pi = A.info[B.pair];
    const int off = pi.off;
    const double d0sq = fmax(pi.d0 * pi.d0, 1e-12), c8 = B.sd8 * B.sd8;
    double acc = 0.0;
    for (int i = 0; i < n; i++) {
        double x, y, z, qa, qb, qc;
        if (!get_pt<PTS>(A, off, lq, B.a0, B.a1, i, x, y, z, qa, qb, qc)) continue;
        const double d2 = FAST ? dist2_moved_fma(R, t, x, y, z, qa, qb, qc)
                               : dist2_moved(R, t, x, y, z, qa, qb, qc);
        if (B.use_d8 && !(d2 <= c8)) continue;
        // fast: one divide, d0^2 / (d0^2 + d2), instead of the reference's two
        acc += FAST ? d0sq / (d0sq + d2) : tm_term(d2, d0sq);
    }
    return acc / fmax(pi.Lnorm, 1.0);
}

// One candidate: iterative_seed_refine (search.h:28-56) or the gapless
// offset start + refine_from_rt (threading.h:15-37, 415-430), then the
// TM-score of search.h:358-362.
// Returns true with out = the finished candidate, or false when its
// refinement stopped at pass_cap (> 0): out holds its (R, t), st the rest.
// kops accumulates the Kabsch operations (initial and per pass); st.it
// receives the passes run.
template <int PTS, bool FAST>
IA_D bool run_candidate(const Arena &A, const BankDesc &B, int n, int j, const SeedDesc *seeds,
                        const uint8_t *masks, const P8 *lq, int pass_cap, Hit &out,
                        ContState &st, int &kops, const int *list_na) {
    const PairInfo &pi = A.info[B.pair];
    const int off = pi.off;
    const int C = B.C;
    const int kk = j / C, c = j - kk * C;
    const int k = B.sel_m > 0 ? A.isel[B.sel_off + kk] : kk;
    SeedDesc sd;
    if (B.seeds == SEED_EXPLICIT) {
        sd = seeds[B.seed_off + k];
    } else {
        int s0, len;
        if (PTS == PTS_TLIST) torch_list_seed(B.n, list_na[B.a1], B.max_starts, k, s0, len);
        else implicit_seed(n, B.max_starts, k, s0, len);
        sd.a = s0;
        sd.b = len;
    }
    double R[9], t[3];
    KabschSeq ks;
    ks.reset();
    if (B.init == INIT_OFFSET) {
        // gapless threading start: pred[i] <-> native[i-k] (threading.h:417-427)
        const int kk = sd.a, N = pi.N;
        for (int i = 0; i < N; i++) {
            const int jj = i - kk;
            if (jj < 0 || jj >= N || !A.pv[off + i] || !A.nv[off + jj]) continue;
            if (FAST)
                ks.add_moment(A.px[off + i], A.py[off + i], A.pz[off + i], A.qx[off + jj],
                              A.qy[off + jj], A.qz[off + jj]);
            else
                ks.add_centroid(A.px[off + i], A.py[off + i], A.pz[off + i], A.qx[off + jj],
                                A.qy[off + jj], A.qz[off + jj]);
        }
        if (ks.cnt < (double)B.min_ali) {
            out.score = -INFINITY;
            set_identity(out.R, out.t);
            st.it = 0;
            return true;
        }
        kops += 31 * (int)ks.cnt + 3400;
        if (FAST) {
            ks.finish_moments();
        } else {
            ks.centre();
            for (int i = 0; i < N; i++) {
                const int jj = i - kk;
                if (jj < 0 || jj >= N || !A.pv[off + i] || !A.nv[off + jj]) continue;
                ks.add_cov(A.px[off + i], A.py[off + i], A.pz[off + i], A.qx[off + jj],
                           A.qy[off + jj], A.qz[off + jj]);
            }
        }
        ks.template solve<FAST>(R, t);
    } else {
        // seed & valid (search.h:34-35)
        const bool win = sd.b > 0;
        const uint8_t *m = masks + (win ? 0 : (long long)sd.a);
        for (int i = 0; i < n; i++) {
            const bool in = win ? (i >= sd.a && i < sd.a + sd.b) : (m[i] != 0);
            if (!in) continue;
            double x, y, z, qa, qb, qc;
            if (!get_pt<PTS>(A, off, lq, B.a0, B.a1, i, x, y, z, qa, qb, qc)) continue;
            if (FAST) ks.add_moment(x, y, z, qa, qb, qc);
            else ks.add_centroid(x, y, z, qa, qb, qc);
        }
        if (ks.cnt >= 3.0) kops += 31 * (int)ks.cnt + 3400;
        if (PTS == PTS_TLIST && ks.cnt < 3.0) {  // the engine drops such seeds (search.py:119-122)
            out.score = kNeg;
            set_identity(out.R, out.t);
            st.it = 0;
            return true;
        }
        if (ks.cnt < 3.0) {
            set_identity(R, t);
        } else if (FAST) {
            ks.finish_moments();
            ks.template solve<FAST>(R, t);
        } else {
            ks.centre();
            for (int i = 0; i < n; i++) {
                const bool in = win ? (i >= sd.a && i < sd.a + sd.b) : (m[i] != 0);
                if (!in) continue;
                double x, y, z, qa, qb, qc;
                if (!get_pt<PTS>(A, off, lq, B.a0, B.a1, i, x, y, z, qa, qb, qc)) continue;
                ks.add_cov(x, y, z, qa, qb, qc);
            }
            ks.template solve<FAST>(R, t);
        }
    }
    // d0s-cutoff refinement; stop when the inlier count repeats.
    int it = 0, last = -1;
    const bool done = refine_passes<PTS, FAST>(A, B, n, c, lq, R, t, it, last,
                                               pass_cap > 0 ? min(pass_cap, B.iters) : B.iters,
                                               kops);
#pragma unroll
    for (int q = 0; q < 9; q++) out.R[q] = R[q];
    out.t[0] = t[0]; out.t[1] = t[1]; out.t[2] = t[2];
    st.last = last;
    st.it = it;
    if (!done) {  // stopped at the pass cap: the continuation kernel resumes it
        out.score = kNeg;
        return false;
    }
    out.score = candidate_score<PTS, FAST>(A, B, n, lq, R, t);
    return true;
}

// With pass_cap > 0, a candidate still refining after pass_cap passes
// leaves its (R, t) in its slot and its state in cont_state, and its index
// joins its bank's list (cont_idx[cand_off + k], k < cont_cnt[bank]) for
// bank_cont_kernel; the lanes of a wavefront then wait for at most pass_cap
// passes of their slowest candidate instead of up to B.iters.
// Occupancy (measured): six wavefronts per SIMD for bank_kernel and four for
// bank_cont_kernel, both methods; the few registers this spills cost less
// than the latency the extra wavefronts hide.
// TL: the instantiation that runs the PyTorch-engine list banks (PTS_TLIST)
// and nothing else; the others never compile that point kind, which keeps
// its per-point flag test out of their loops.
template <bool FAST, bool TL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void bank_kernel(
    Arena A, const BankDesc *banks, const int *blk_bank, const int *blk_first,
    const SeedDesc *seeds, const uint8_t *masks, const P8 *list_pq, const int *list_na,
    int list_cap, Hit *cands, int pass_cap, int *cont_cnt, int *cont_idx,
    ContState *cont_state) {
    const int bank = blk_bank[blockIdx.x];
    const BankDesc &B = banks[bank];
    const int j = blk_first[blockIdx.x] + (int)threadIdx.x;
    const bool live = j < B.cand_cap;
    ContState st{-1, 0};
    bool done = true;
    int kops = 0;
    const int n = bank_points(B, list_na);
    bool ran = false;
    if (live) {
        Hit h;
        const int K = bank_live_seeds(B, n, list_na);
        if (j >= K * B.C || n <= 0) {
            h.score = -INFINITY;
            set_identity(h.R, h.t);
        } else if constexpr (TL) {
            ran = true;
            const long long base = (long long)B.a0 * list_cap;
            done = run_candidate<PTS_TLIST, FAST>(A, B, n, j, seeds, masks, list_pq + base,
                                                  pass_cap, h, st, kops, list_na);
        } else if (B.pts == PTS_FULL) {
            ran = true;
            done = run_candidate<PTS_FULL, FAST>(A, B, n, j, seeds, masks, nullptr, pass_cap, h, st,
                                                 kops, list_na);
        } else if (B.pts == PTS_LIST) {
            ran = true;
            const long long base = (long long)B.a0 * list_cap;
            done = run_candidate<PTS_LIST, FAST>(A, B, n, j, seeds, masks, list_pq + base, pass_cap,
                                                 h, st, kops, list_na);
        } else {
            ran = true;
            done = run_candidate<PTS_CWIN, FAST>(A, B, n, j, seeds, masks, nullptr, pass_cap, h, st,
                                                 kops, list_na);
        }
        cands[B.cand_off + j] = h;
    }
    {   // work counters: one atomic per wavefront (every lane is still here)
        const double ops = ran ? (double)kops + 26.0 * n * st.it + (done ? 29.0 * n : 0.0) : 0.0;
        const double passes = ran ? (double)st.it : 0.0;
        const double so = wave_sum(ops), sp = wave_sum(passes);
        if ((threadIdx.x & 63) == 0 && sp > 0) {
            atomicAdd(&A.bank_work[0], (unsigned long long)so);
            atomicAdd(&A.bank_work[1], (unsigned long long)sp);
        }
    }
    if (!live || pass_cap <= 0) return;
    // append the unfinished candidates of this wavefront (one atomic per wave)
    const unsigned long long m = __ballot(!done);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int lead = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == lead) base = atomicAdd(&cont_cnt[bank], __popcll(m));
    base = __shfl(base, lead, 64);
    if (!done) {
        const int k = base + __popcll(m & ((1ull << lane) - 1ull));
        cont_idx[B.cand_off + k] = j;
        cont_state[B.cand_off + j] = st;
    }
}

// The unfinished candidates of each bank, 256 per block: the rest of their
// refinement (from their saved state) and their TM-score.
template <bool FAST, bool TL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void bank_cont_kernel(Arena A, const BankDesc *banks,
                                                        const int *blk_bank, const int *blk_first,
                                                        const P8 *list_pq, const int *list_na,
                                                        int list_cap, Hit *cands,
                                                        const int *cont_cnt, const int *cont_idx,
                                                        const ContState *cont_state) {
    const int bank = blk_bank[blockIdx.x];
    const BankDesc &B = banks[bank];
    const int k = blk_first[blockIdx.x] + (int)threadIdx.x;
    const bool live = k < cont_cnt[bank];
    const int n = bank_points(B, list_na);
    int kops = 0, it0 = 0, it1 = 0;
    if (live) {
        const int j = cont_idx[B.cand_off + k];
        Hit &h = cands[B.cand_off + j];
        ContState st = cont_state[B.cand_off + j];
        it0 = st.it;
        double R[9], t[3];
#pragma unroll
        for (int q = 0; q < 9; q++) R[q] = h.R[q];
        t[0] = h.t[0]; t[1] = h.t[1]; t[2] = h.t[2];
        const int c = j % B.C;
        double score;
        if constexpr (TL) {
            const P8 *lq = list_pq + (long long)B.a0 * list_cap;
            refine_passes<PTS_TLIST, FAST>(A, B, n, c, lq, R, t, st.it, st.last, B.iters, kops);
            score = candidate_score<PTS_TLIST, FAST>(A, B, n, lq, R, t);
        } else if (B.pts == PTS_FULL) {
            refine_passes<PTS_FULL, FAST>(A, B, n, c, nullptr, R, t, st.it, st.last, B.iters, kops);
            score = candidate_score<PTS_FULL, FAST>(A, B, n, nullptr, R, t);
        } else if (B.pts == PTS_LIST) {
            const P8 *lq = list_pq + (long long)B.a0 * list_cap;
            refine_passes<PTS_LIST, FAST>(A, B, n, c, lq, R, t, st.it, st.last, B.iters, kops);
            score = candidate_score<PTS_LIST, FAST>(A, B, n, lq, R, t);
        } else {
            refine_passes<PTS_CWIN, FAST>(A, B, n, c, nullptr, R, t, st.it, st.last, B.iters, kops);
            score = candidate_score<PTS_CWIN, FAST>(A, B, n, nullptr, R, t);
        }
        it1 = st.it;
        Hit o;
#pragma unroll
        for (int q = 0; q < 9; q++) o.R[q] = R[q];
        o.t[0] = t[0]; o.t[1] = t[1]; o.t[2] = t[2];
        o.score = score;
        h = o;
    }
    const double ops = live ? (double)kops + 26.0 * n * (it1 - it0) + 29.0 * n : 0.0;
    const double passes = live ? (double)(it1 - it0) : 0.0;
    const double so = wave_sum(ops), sp = wave_sum(passes);
    if ((threadIdx.x & 63) == 0 && so > 0) {
        atomicAdd(&A.bank_work[0], (unsigned long long)so);
        atomicAdd(&A.bank_work[1], (unsigned long long)sp);
    }
}

// Copy the aligned pairs of every PTS_LIST bank's traceback list into
// interleaved 64-byte records (list_pq[slot * list_cap + k]).
__global__ __launch_bounds__(256) void list_gather_kernel(Arena A, const BankDesc *banks,
                                                          const int *lists_p, const int *lists_n,
                                                          const int *list_na, int list_cap,
                                                          P8 *list_pq) {
    const BankDesc &B = banks[blockIdx.x];
    if (B.pts != PTS_LIST && B.pts != PTS_TLIST) return;
    const int off = A.info[B.pair].off;
    const long long base = (long long)B.a0 * list_cap;
    const int na = list_na[B.a0];
    for (int k = threadIdx.x; k < na; k += 256) {
        P8 r;
        const int ip = lists_p[base + k], in = lists_n[base + k];
        r.p = A.p4[off + ip];
        r.q = A.q4[off + in];
        // the entry's flag: PyTorch-engine lists keep masked entries
        r.p.w = B.pts != PTS_TLIST || (A.v[off + ip] != 0 && A.v[off + in] != 0) ? 1.0 : 0.0;
        list_pq[base + k] = r;
    }
}

__global__ void bank_seed_count_kernel(const BankDesc *banks, int n_banks, const int *list_na,
                                       int *out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_banks) return;
    const int n = bank_points(banks[b], list_na);
    out[b] = n > 0 ? bank_seeds(banks[b], n, list_na) : 0;
}

// (score, index) first-wins maximum over a block.
IA_D void block_argmax(double &s, long long &idx) {
    __shared__ double ss[256];
    __shared__ long long si[256];
    const int t = threadIdx.x;
    ss[t] = s;
    si[t] = idx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            const double os = ss[t + w];
            const long long oi = si[t + w];
            if (oi >= 0 && (si[t] < 0 || os > ss[t] || (os == ss[t] && oi < si[t]))) {
                ss[t] = os;
                si[t] = oi;
            }
        }
        __syncthreads();
    }
    s = ss[0];
    idx = si[0];
    __syncthreads();
}

__global__ __launch_bounds__(256) void bank_reduce_kernel(const BankDesc *banks,
                                                          const int *list_na, const Hit *cands,
                                                          Hit *out) {
    const BankDesc &B = banks[blockIdx.x];
    const int n = bank_points(B, list_na);
    // never past the bank's own candidate slots
    const int total = n > 0 ? min(bank_live_seeds(B, n, list_na) * B.C, B.cand_cap) : 0;
    const Hit *cb = cands + B.cand_off;
    // serial early termination: stop after the first seed whose candidates
    // lift the running best to >= et (search.h:83)
    __shared__ int first;
    if (threadIdx.x == 0) first = 0x7fffffff;
    __syncthreads();
    for (int j = threadIdx.x; j < total; j += 256)
        if (cb[j].score >= B.et) atomicMin(&first, j);
    __syncthreads();
    const int limit = first == 0x7fffffff ? total : (first / B.C + 1) * B.C;
    double bs = kNeg;
    long long bi = -1;
    for (int j = threadIdx.x; j < limit; j += 256) {
        const double s = cb[j].score;
        if (s > bs) { bs = s; bi = j; }
    }
    block_argmax(bs, bi);
    if (threadIdx.x == 0) {
        Hit h;
        if (bi >= 0 && bs > kNeg) {
            h = cb[bi];
        } else {
            set_identity(h.R, h.t);
            h.score = kNeg;
        }
        out[B.out_idx] = h;
    }
}

__global__ __launch_bounds__(256) void group_reduce_kernel(const long long *offs,
                                                           const int *counts, const Hit *in,
                                                           Hit *out) {
    const Hit *g = in + offs[blockIdx.x];
    const int cnt = counts[blockIdx.x];
    double bs = kNeg;
    long long bi = -1;
    for (int j = threadIdx.x; j < cnt; j += 256) {
        const double s = g[j].score;
        if (s > bs) { bs = s; bi = j; }
    }
    block_argmax(bs, bi);
    if (threadIdx.x == 0) {
        Hit h;
        if (bi >= 0 && bs > kNeg) {
            h = g[bi];
        } else {
            // empty group: the reference's initial (-1e30, identity); a
            // non-empty group of skipped entries stays skipped (-inf)
            set_identity(h.R, h.t);
            h.score = cnt > 0 ? -INFINITY : kNeg;
        }
        out[blockIdx.x] = h;
    }
}

// Same-index TM-score without the d8 cut (core.h:152-166), one lane per item,
// index-order sum.  keep_none: entries without a result (score kNeg) stay
// without one instead of being scored as the identity.
__global__ __launch_bounds__(256) void rescore_kernel(Arena A, const int *pair_of,
                                                      const Hit *in, int n, Hit *out,
                                                      int keep_none) {
    const int it = blockIdx.x * 256 + threadIdx.x;
    if (it >= n) return;
    const PairInfo &pi = A.info[pair_of[it]];
    Hit h = in[it];
    if (h.score == -INFINITY || (keep_none && h.score <= kNeg)) {  // nothing to rescore
        out[it] = h;
        return;
    }
    const double d0sq = fmax(pi.d0 * pi.d0, 1e-12);
    double acc = 0.0;
    for (int i = 0; i < pi.N; i++) {
        const int o = pi.off + i;
        if (!A.v[o]) continue;
        acc += tm_term(dist2_moved(h.R, h.t, A.px[o], A.py[o], A.pz[o], A.qx[o], A.qy[o], A.qz[o]),
                       d0sq);
    }
    h.score = acc / fmax(pi.Lnorm, 1.0);
    out[it] = h;
}

// Top-k distance seed (indialign.cpp:156-161, seeds.h:85-98): the
// k = min(max(3, floor(V/2)), N) valid residues with the smallest (d2, index).
__global__ __launch_bounds__(256) void topk_kernel(Arena A, const int *pairs, const Hit *best,
                                                   const long long *mask_off, uint8_t *masks) {
    const PairInfo &pi = A.info[pairs[blockIdx.x]];
    const Hit h = best[blockIdx.x];
    uint8_t *m = masks + mask_off[blockIdx.x];
    int k = pi.V / 2;  // (int)(V * 0.5)
    if (k < 3) k = 3;
    if (k > pi.N) k = pi.N;
    for (int i = threadIdx.x; i < pi.N; i += 256) {
        const int o = pi.off + i;
        uint8_t sel = 0;
        if (A.v[o]) {
            const double di = dist2_moved(h.R, h.t, A.px[o], A.py[o], A.pz[o], A.qx[o], A.qy[o], A.qz[o]);
            int rank = 0;
            for (int j = 0; j < pi.N && rank < k; j++) {
                const int p = pi.off + j;
                if (!A.v[p] || j == i) continue;
                const double dj = dist2_moved(h.R, h.t, A.px[p], A.py[p], A.pz[p], A.qx[p], A.qy[p], A.qz[p]);
                rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
            }
            sel = rank < k ? 1 : 0;
        }
        m[i] = sel;
    }
}

// Weighted TM refinement, 4 iterations (core.h:183-237); one lane per pair,
// index-order sums.
__global__ __launch_bounds__(64) void weighted_kernel(Arena A, const int *pairs, const Hit *best,
                                                      int n, Hit *out) {
    const int it = blockIdx.x * 64 + threadIdx.x;
    if (it >= n) return;
    const PairInfo &pi = A.info[pairs[it]];
    const int off = pi.off, N = pi.N;
    Hit h = best[it];
    const double d0sq = fmax(pi.d0 * pi.d0, 1e-12);
    for (int iter = 0; iter < 4; iter++) {
        // weights with the current transform, then weighted Kabsch
        double cp[3] = {0, 0, 0}, cq[3] = {0, 0, 0}, ws = 0.0;
        for (int i = 0; i < N; i++) {
            const int o = off + i;
            if (!A.v[o]) continue;
            const double w = tm_term(dist2_moved(h.R, h.t, A.px[o], A.py[o], A.pz[o], A.qx[o], A.qy[o], A.qz[o]), d0sq);
            if (w <= 0) continue;
            cp[0] += w * A.px[o]; cp[1] += w * A.py[o]; cp[2] += w * A.pz[o];
            cq[0] += w * A.qx[o]; cq[1] += w * A.qy[o]; cq[2] += w * A.qz[o];
            ws += w;
        }
        if (ws < 1e-12) {
            set_identity(h.R, h.t);
            continue;
        }
        for (int d = 0; d < 3; d++) { cp[d] /= ws; cq[d] /= ws; }
        double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; i < N; i++) {
            const int o = off + i;
            if (!A.v[o]) continue;
            const double w = tm_term(dist2_moved(h.R, h.t, A.px[o], A.py[o], A.pz[o], A.qx[o], A.qy[o], A.qz[o]), d0sq);
            if (w <= 0) continue;
            const double a0 = A.px[o] - cp[0], a1 = A.py[o] - cp[1], a2 = A.pz[o] - cp[2];
            const double b0 = A.qx[o] - cq[0], b1 = A.qy[o] - cq[1], b2 = A.qz[o] - cq[2];
            H[0] += w * a0 * b0; H[1] += w * a0 * b1; H[2] += w * a0 * b2;
            H[3] += w * a1 * b0; H[4] += w * a1 * b1; H[5] += w * a1 * b2;
            H[6] += w * a2 * b0; H[7] += w * a2 * b1; H[8] += w * a2 * b2;
        }
        double R[9], t[3];
        horn_rotation(H, R);
        finish_t(cp, cq, R, t);
        for (int q = 0; q < 9; q++) h.R[q] = R[q];
        h.t[0] = t[0]; h.t[1] = t[1]; h.t[2] = t[2];
    }
    double acc = 0.0;
    for (int i = 0; i < N; i++) {
        const int o = off + i;
        if (!A.v[o]) continue;
        acc += tm_term(dist2_moved(h.R, h.t, A.px[o], A.py[o], A.pz[o], A.qx[o], A.qy[o], A.qz[o]), d0sq);
    }
    h.score = acc / fmax(pi.Lnorm, 1.0);
    out[it] = h;
}

}  // namespace

void launch_bank(hipStream_t s, const Arena &A, const BankDesc *banks, const int *blk_bank,
                 const int *blk_first, int n_blocks, const SeedDesc *seeds,
                 const uint8_t *masks, const P8 *list_pq, const int *list_na, int list_cap,
                 Hit *cands, bool fast, int pass_cap, int *cont_cnt, int *cont_idx,
                 void *cont_state, bool tlist) {
    if (n_blocks <= 0) return;
    ContState *cs = (ContState *)cont_state;
#define IA_LAUNCH_BANK(F, T)                                                                    \
    bank_kernel<F, T><<<n_blocks, 256, 0, s>>>(A, banks, blk_bank, blk_first, seeds, masks,     \
                                               list_pq, list_na, list_cap, cands, pass_cap,     \
                                               cont_cnt, cont_idx, cs)
    if (tlist) IA_LAUNCH_BANK(false, true);  // PyTorch-engine lists: FP64 only
    else if (fast) IA_LAUNCH_BANK(true, false);
    else IA_LAUNCH_BANK(false, false);
#undef IA_LAUNCH_BANK
}

void launch_bank_cont(hipStream_t s, const Arena &A, const BankDesc *banks, const int *blk_bank,
                      const int *blk_first, int n_blocks, const P8 *list_pq, const int *list_na,
                      int list_cap, Hit *cands, bool fast, const int *cont_cnt,
                      const int *cont_idx, const void *cont_state, bool tlist) {
    if (n_blocks <= 0) return;
    const ContState *cs = (const ContState *)cont_state;
#define IA_LAUNCH_CONT(F, T)                                                                    \
    bank_cont_kernel<F, T><<<n_blocks, 256, 0, s>>>(A, banks, blk_bank, blk_first, list_pq,     \
                                                    list_na, list_cap, cands, cont_cnt,         \
                                                    cont_idx, cs)
    if (tlist) IA_LAUNCH_CONT(false, true);
    else if (fast) IA_LAUNCH_CONT(true, false);
    else IA_LAUNCH_CONT(false, false);
#undef IA_LAUNCH_CONT
}

void launch_list_gather(hipStream_t s, const Arena &A, const BankDesc *banks, int n_banks,
                        const int *lists_p, const int *lists_n, const int *list_na, int list_cap,
                        P8 *list_pq) {
    if (n_banks <= 0) return;
    list_gather_kernel<<<n_banks, 256, 0, s>>>(A, banks, lists_p, lists_n, list_na, list_cap,
                                               list_pq);
}

void launch_bank_seed_count(hipStream_t s, const BankDesc *banks, int n_banks, const int *list_na,
                            int *out) {
    if (n_banks <= 0) return;
    bank_seed_count_kernel<<<(n_banks + 255) / 256, 256, 0, s>>>(banks, n_banks, list_na, out);
}

void launch_bank_reduce(hipStream_t s, const BankDesc *banks, int n_banks, const int *list_na,
                        const Hit *cands, Hit *out) {
    if (n_banks <= 0) return;
    bank_reduce_kernel<<<n_banks, 256, 0, s>>>(banks, list_na, cands, out);
}

void launch_group_reduce(hipStream_t s, const long long *offs, const int *counts, int n,
                         const Hit *in, Hit *out) {
    if (n <= 0) return;
    group_reduce_kernel<<<n, 256, 0, s>>>(offs, counts, in, out);
}

void launch_rescore(hipStream_t s, const Arena &A, const int *pair_of, const Hit *in, int n,
                    Hit *out, bool keep_none) {
    if (n <= 0) return;
    rescore_kernel<<<(n + 255) / 256, 256, 0, s>>>(A, pair_of, in, n, out, keep_none ? 1 : 0);
}

void launch_topk(hipStream_t s, const Arena &A, const int *pairs, const Hit *best, int n,
                 const long long *mask_off, uint8_t *masks) {
    if (n <= 0) return;
    topk_kernel<<<n, 256, 0, s>>>(A, pairs, best, mask_off, masks);
}

void launch_weighted(hipStream_t s, const Arena &A, const int *pairs, const Hit *best, int n,
                     Hit *out) {
    if (n <= 0) return;
    weighted_kernel<<<(n + 63) / 64, 64, 0, s>>>(A, pairs, best, n, out);
}

}  // namespace ia
