// Learned ranking of the local-fragment rescue work items (rescue.h:60-118).
//
// The dense local-fragment pass refines every fragment pair (flen, ps, ns) of
// a pair by two NW chains; almost all of them end below the pair's best.  The
// learned method scores every work item with a small MLP over a 96-value
// descriptor and refines only the top-m items of each pair.  This file holds
//   - the descriptor kernel (one lane per work item),
//   - a BF16-operand / FP32-accumulate GEMM on the matrix cores
//     (v_mfma_f32_32x32x8_bf16) that the MLP's forward and backward passes
//     are made of, plus the small elementwise kernels of training (MSE, head
//     backward, Adam),
//   - the per-pair top-m selection (radix select, ties to the lower index).
#include "ia_kernels.h"
#include "wave.h"

#include <math.h>

namespace ia {

namespace {

// ---- work-item descriptor ---------------------------------------------------
//
// Row layout (kFeat = 96 values, all O(1)):
//   0 flen/ml   1 flen/100   2 ps/plen   3 ns/nlen   4 (ps-ns)/ml
//   5 [ps == ns]   6 Lnorm/300   7 d0/10   8 dense pass   9 prior best score
//  10 fragment RMSD/d0 (<= 10)   11 TM-score of the fragment superposition,
//     same-index pairing   12 same, register pairing pred a <-> native a+ns-ps
//  13..23 fraction of same-index residues within the pair's d0 search cutoffs
//  24 C/11   25 plen/300   26 nlen/300   27 sampled pairs/276   28 log(1+flen)/5
//  29 register-pairing overlap / ml
//  30, 31 TM-score (same-index pairing) after one and two refinements of the
//     superposition: Kabsch over the same-index residues within 2 d0, then
//     within d0, of the previous superposition
//  32..95 histogram of |d_pred(i,j) - d_native(i',j')| over up to 24 evenly
//     spaced fragment positions, 0.25 A bins (last bin open), sum 1.
constexpr int kHistPos = 24;

constexpr int kFeatBlock = kFeatItemsPerBlock;  // lanes per workgroup, all of one pair

// One lane = one work item; the lanes of a workgroup share the pair, so the
// passes over the pair's residues are wave-uniform (scalar loads) and need
// no cross-lane reductions.  Kabsch sums are single-pass FP64 moments;
// TM-score terms are FP32.
__global__ __launch_bounds__(kFeatBlock) void frag_feature_kernel(Arena A, const FeatGroup *groups,
                                                                  const int *blk_group,
                                                                  const int *blk_first, float *X) {
    __shared__ unsigned hist_s[64 * kFeatBlock];
    const FeatGroup &G = groups[blk_group[blockIdx.x]];
    const int item = blk_first[blockIdx.x] + (int)threadIdx.x;
    if (item >= G.items) return;
    const PairInfo &pi = A.info[G.pair];
    const int off = pi.off, N = pi.N;
    int L_ = 0;
    while (L_ + 1 < G.n_lens && G.items_prefix[L_ + 1] <= item) L_++;
    const int r = item - G.items_prefix[L_];
    const int flen = G.lens[L_];
    const int ps = (r / G.n_ns[L_]) * G.jp, ns = (r % G.n_ns[L_]) * G.jn;
    // fragment superposition and its RMSD from the moments
    KabschSeq ks;
    ks.reset();
    double spp = 0.0, sqq = 0.0;
    for (int q = 0; q < flen; q++) {
        const P4 P = A.pc4[off + ps + q], Q = A.qc4[off + ns + q];
        ks.add_moment(P.x, P.y, P.z, Q.x, Q.y, Q.z);
        spp = fma(P.x, P.x, fma(P.y, P.y, fma(P.z, P.z, spp)));
        sqq = fma(Q.x, Q.x, fma(Q.y, Q.y, fma(Q.z, Q.z, sqq)));
    }
    ks.finish_moments();
    double R[9], t[3];
    ks.template solve<true>(R, t);
    double rh = 0.0;
#pragma unroll
    for (int j = 0; j < 9; j++) rh = fma(R[j], ks.h[j], rh);
    const double fl = (double)flen;
    const double cpp = ks.cp[0] * ks.cp[0] + ks.cp[1] * ks.cp[1] + ks.cp[2] * ks.cp[2];
    const double cqq = ks.cq[0] * ks.cq[0] + ks.cq[1] * ks.cq[1] + ks.cq[2] * ks.cq[2];
    const double rmsd = sqrt(fmax(spp - fl * cpp + sqq - fl * cqq - 2.0 * rh, 0.0) / fl);
    // same-index TM-score and cutoff counts
    const double d0sq = pi.d0 * pi.d0;
    const float inv_d0sq = (float)(1.0 / d0sq);
    float cut2[kMaxCut], cnt[kMaxCut];
#pragma unroll
    for (int c = 0; c < kMaxCut; c++) {
        cut2[c] = c < pi.C ? (float)(pi.cut[c] * pi.cut[c]) : -1.f;
        cnt[c] = 0.f;
    }
    float tm = 0.f;
    for (int i = 0; i < N; i++) {
        const P4 P = A.p4[off + i];
        if (P.w == 0.0) continue;
        const P4 Q = A.q4[off + i];
        const float d2 = (float)dist2_moved_fma(R, t, P.x, P.y, P.z, Q.x, Q.y, Q.z);
        tm += __builtin_amdgcn_rcpf(__builtin_fmaf(d2, inv_d0sq, 1.f));
#pragma unroll
        for (int c = 0; c < kMaxCut; c++) cnt[c] += d2 < cut2[c] ? 1.f : 0.f;
    }
    // register pairing a <-> a + (ns - ps) over the compact rows
    const int sh = ns - ps;
    const int a0 = sh < 0 ? -sh : 0, a1 = min(pi.plen, pi.nlen - sh);
    float tmr = 0.f;
    for (int a = a0; a < a1; a++) {
        const P4 P = A.pc4[off + a], Q = A.qc4[off + a + sh];
        const float d2 = (float)dist2_moved_fma(R, t, P.x, P.y, P.z, Q.x, Q.y, Q.z);
        tmr += __builtin_amdgcn_rcpf(__builtin_fmaf(d2, inv_d0sq, 1.f));
    }
    const int ov = a1 > a0 ? a1 - a0 : 0;
    // two cutoff refinements of the same-index superposition (2 d0, d0)
    float tmref[2] = {tm, tm};
    for (int step = 0; step < 2; step++) {
        const double cut = step == 0 ? 4.0 * d0sq : d0sq;
        ks.reset();
        for (int i = 0; i < N; i++) {
            const P4 P = A.p4[off + i];
            if (P.w == 0.0) continue;
            const P4 Q = A.q4[off + i];
            if (dist2_moved_fma(R, t, P.x, P.y, P.z, Q.x, Q.y, Q.z) < cut)
                ks.add_moment(P.x, P.y, P.z, Q.x, Q.y, Q.z);
        }
        if (ks.cnt < 3.0) {
            tmref[step] = step == 0 ? tm : tmref[0];
            continue;
        }
        ks.finish_moments();
        ks.template solve<true>(R, t);
        float sc = 0.f;
        for (int i = 0; i < N; i++) {
            const P4 P = A.p4[off + i];
            if (P.w == 0.0) continue;
            const P4 Q = A.q4[off + i];
            const float d2 = (float)dist2_moved_fma(R, t, P.x, P.y, P.z, Q.x, Q.y, Q.z);
            sc += __builtin_amdgcn_rcpf(__builtin_fmaf(d2, inv_d0sq, 1.f));
        }
        tmref[step] = sc;
    }
    // intra-fragment distance-difference histogram (lane-private LDS bins)
    unsigned *hist = hist_s + threadIdx.x;
#pragma unroll 8
    for (int b = 0; b < 64; b++) hist[b * kFeatBlock] = 0;
    const int m = flen < kHistPos ? flen : kHistPos;
    for (int k1 = 0; k1 + 1 < m; k1++) {
        const int q1 = (k1 * (flen - 1)) / (m - 1);
        const P4 P1 = A.pc4[off + ps + q1], Q1 = A.qc4[off + ns + q1];
        for (int k2 = k1 + 1; k2 < m; k2++) {
            const int q2 = (k2 * (flen - 1)) / (m - 1);
            const P4 P2 = A.pc4[off + ps + q2], Q2 = A.qc4[off + ns + q2];
            const float ax = (float)(P1.x - P2.x), ay = (float)(P1.y - P2.y), az = (float)(P1.z - P2.z);
            const float bx = (float)(Q1.x - Q2.x), by = (float)(Q1.y - Q2.y), bz = (float)(Q1.z - Q2.z);
            const float dd = fabsf(sqrtf(ax * ax + ay * ay + az * az) - sqrtf(bx * bx + by * by + bz * bz));
            const int bin = dd >= 15.75f ? 63 : (int)(dd * 4.f);
            hist[bin * kFeatBlock] += 1u;
        }
    }
    const int np = m * (m - 1) / 2;
    const int ml = pi.plen < pi.nlen ? pi.plen : pi.nlen;
    const double L = pi.Lnorm;
    float f[32];
#pragma unroll
    for (int k = 0; k < 32; k++) f[k] = 0.f;
    f[0] = (float)flen / (float)ml;
    f[1] = (float)flen / 100.f;
    f[2] = (float)ps / (float)pi.plen;
    f[3] = (float)ns / (float)pi.nlen;
    f[4] = (float)(ps - ns) / (float)ml;
    f[5] = ps == ns ? 1.f : 0.f;
    f[6] = (float)(L / 300.0);
    f[7] = (float)(pi.d0 / 10.0);
    f[8] = G.dense;
    f[9] = G.prior;
    f[10] = (float)fmin(rmsd / pi.d0, 10.0);
    f[11] = (float)(tm / L);
    f[12] = (float)(tmr / L);
#pragma unroll
    for (int c = 0; c < kMaxCut; c++) f[13 + c] = (float)(cnt[c] / L);
    f[24] = (float)pi.C / 11.f;
    f[25] = (float)pi.plen / 300.f;
    f[26] = (float)pi.nlen / 300.f;
    f[27] = (float)np / 276.f;
    f[28] = logf(1.f + (float)flen) / 5.f;
    f[29] = (float)ov / (float)ml;
    f[30] = (float)(tmref[0] / L);
    f[31] = (float)(tmref[1] / L);
    float4 *xo = (float4 *)(X + (G.row0 + item) * kFeat);
#pragma unroll
    for (int k = 0; k < 8; k++) xo[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
    const float inv_np = np > 0 ? 1.f / (float)np : 0.f;
#pragma unroll
    for (int k = 0; k < 16; k++)
        xo[8 + k] = make_float4((float)hist[(4 * k) * kFeatBlock] * inv_np,
                                (float)hist[(4 * k + 1) * kFeatBlock] * inv_np,
                                (float)hist[(4 * k + 2) * kFeatBlock] * inv_np,
                                (float)hist[(4 * k + 3) * kFeatBlock] * inv_np);
}

// ---- BF16 GEMM on the matrix cores --------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (finite inputs).
IA_D short bf16_rne(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
// The FP32 value of bf16_rne(f).
IA_D float bf16_val(float f) { return __uint_as_float((unsigned)(unsigned short)bf16_rne(f) << 16); }

// One workgroup = 4 wavefronts = a 64 x 64 tile of C; each wavefront owns a
// 32 x 32 sub-tile and runs v_mfma_f32_32x32x8_bf16 over K in steps of 8.
// Operand layout of the instruction: lane l holds A(row l%32, k 4*(l/32)+j)
// and B(k 4*(l/32)+j, col l%32), j = 0..3; result v of lane l is
// C(8*(v/4) + 4*(l/32) + v%4, l%32).
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long m0 = (long long)blockIdx.y * 64 + (wv >> 1) * 32;
    const int n0 = blockIdx.x * 64 + (wv & 1) * 32;
    const int r = lane & 31, kq = (lane >> 5) * 4;
    const long long am = m0 + r;
    const int bn = n0 + r;
    const bool a_ok = am < g.M, b_ok = bn < g.N;
    const float *Ap = g.A + (a_ok ? am * g.sam : 0);
    const float *Bp = g.B + (b_ok ? (long long)bn * g.sbn : 0);
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; v++) acc[v] = 0.f;
    for (int k0 = 0; k0 < g.K; k0 += 8) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + kq + j;
            const bool kin = k < g.K;
            const float av = (a_ok && kin) ? Ap[(long long)k * g.sak] : 0.f;
            const float bv = (b_ok && kin) ? Bp[(long long)k * g.sbk] : 0.f;
            a[j] = bf16_rne(av);
            b[j] = bf16_rne(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int n = n0 + r;
    if (n >= g.N) return;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const long long m = m0 + 8 * (v >> 2) + (lane >> 5) * 4 + (v & 3);
        if (m >= g.M) continue;
        float x = acc[v];
        const long long ci = m * g.scm + (long long)n * g.scn;
        if (g.epi == EPI_BIAS_RELU) x = fmaxf(x + g.bias[n], 0.f);
        else if (g.epi == EPI_BIAS) x = x + g.bias[n];
        else if (g.epi == EPI_BIAS_SKIP) x = x + g.bias[n] + g.skip[m * g.sskip];
        else if (g.epi == EPI_MASK) x = g.mask[ci] > 0.f ? x : 0.f;
        g.C[ci] = x;
    }
}

// ---- fused ranker inference ------------------------------------------------
// One wavefront = 32 rows, a workgroup = 4 wavefronts.  Layer 1
// (32x96 . 96x256) accumulates in 8 tiles of 32x32; relu(+b1) goes to LDS as
// BF16; layer 2 (32x256 . 256x256) reads its A operand from there; the head
// relu(+b2) . w3 is reduced across the 32 lanes of each column group.  The
// rounding points are those of the GEMM path: every matrix operand
// (X, W1, h1, W2, h2, w3) in BF16, sums in FP32.
constexpr int kInferRows = 128;
constexpr int kH1Stride = 260;  // BF16 elements per LDS row (8-byte aligned, 2-way banks)

__global__ __launch_bounds__(256) void mlp_infer_kernel(const float *X, int n, const short *W1b,
                                                        const short *W2b, const float *P,
                                                        float *y) {
    __shared__ short h1s[4][32 * kH1Stride];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 31, kq = (lane >> 5) * 4;
    const long long m0 = (long long)blockIdx.x * kInferRows + wv * 32;
    const long long row = m0 + r;
    const bool rok = row < n;
    short *H = h1s[wv];
    f32x16 acc[8];
#pragma unroll
    for (int t = 0; t < 8; t++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[t][v] = 0.f;
    // layer 1
    for (int k0 = 0; k0 < kFeat; k0 += 8) {
        bf16x4 a;
        if (rok) {
            const float4 xa = *(const float4 *)(X + row * kFeat + k0 + kq);
            a[0] = bf16_rne(xa.x); a[1] = bf16_rne(xa.y); a[2] = bf16_rne(xa.z); a[3] = bf16_rne(xa.w);
        } else {
            a[0] = a[1] = a[2] = a[3] = 0;
        }
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const bf16x4 b = *(const bf16x4 *)(W1b + (t * 32 + r) * kFeat + k0 + kq);
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
    // relu(+b1) -> LDS (BF16); result v of lane l is row 8(v/4)+4(l/32)+v%4, col l%32
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int col = t * 32 + r;
        const float bias = P[kOffB1 + col];
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int rr = 8 * (v >> 2) + (lane >> 5) * 4 + (v & 3);
            H[rr * kH1Stride + col] = bf16_rne(fmaxf(acc[t][v] + bias, 0.f));
            acc[t][v] = 0.f;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    // layer 2
    for (int k0 = 0; k0 < kHid; k0 += 8) {
        const bf16x4 a = *(const bf16x4 *)(H + r * kH1Stride + k0 + kq);
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const bf16x4 b = *(const bf16x4 *)(W2b + (t * 32 + r) * kHid + k0 + kq);
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
    // head: sum over columns of bf16(relu(h2 + b2)) * bf16(w3), per row
    float part[16];
#pragma unroll
    for (int v = 0; v < 16; v++) part[v] = 0.f;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int col = t * 32 + r;
        const float bias = P[kOffB2 + col];
        const float w = bf16_val(P[kOffW3 + col]);
#pragma unroll
        for (int v = 0; v < 16; v++) part[v] += bf16_val(fmaxf(acc[t][v] + bias, 0.f)) * w;
    }
#pragma unroll
    for (int v = 0; v < 16; v++) {
        float x = part[v];
        for (int o = 16; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        part[v] = x;
    }
    if (r == 0) {
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const long long rr = m0 + 8 * (v >> 2) + (lane >> 5) * 4 + (v & 3);
            if (rr < n) y[rr] = part[v] + P[kOffB3] + X[rr * kFeat + kSkipCol];
        }
    }
}

// W1, W2 of the parameter array -> BF16 copies.
__global__ void weights_bf16_kernel(const float *P, short *W1b, short *W2b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kHid * kFeat) W1b[i] = bf16_rne(P[kOffW1 + i]);
    if (i < kHid * kHid) W2b[i] = bf16_rne(P[kOffW2 + i]);
}

__global__ void head_back_kernel(const float *dy, const float *w3, const float *h2, long long n,
                                 float *dh2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kHid) return;
    const long long m = i / kHid;
    const int j = (int)(i - m * kHid);
    dh2[i] = h2[i] > 0.f ? dy[m] * w3[j] : 0.f;
}

// One workgroup: dy and the mean squared error (FP32, fixed reduction order).
__global__ __launch_bounds__(1024) void mse_kernel(const float *y, const float *t, int n,
                                                   float *dy, float *loss) {
    __shared__ float part[1024];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const float d = y[i] - t[i];
        dy[i] = 2.f * d / (float)n;
        s += d * d;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = part[0] / (float)n;
}

__global__ void adam_kernel(float *p, const float *g, float *m, float *v, long long n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- per-pair top-m selection ---------------------------------------------------
IA_D unsigned order_key(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Block-wide exclusive prefix count of a flag (256 threads, 4 wavefronts).
IA_D int block_excl(bool flag, int *wtot, int &total) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned long long b = __ballot(flag);
    const int below = __popcll(b & ((1ull << lane) - 1ull));
    __syncthreads();
    if (lane == 0) wtot[wv] = __popcll(b);
    __syncthreads();
    int base = 0;
    total = 0;
    for (int w = 0; w < 4; w++) {
        if (w < wv) base += wtot[w];
        total += wtot[w];
    }
    return base + below;
}

// Top m of each group by prediction; a group whose largest prediction is
// below `thr` keeps its top m_low instead (when m_low > m).
__global__ __launch_bounds__(256) void select_top_kernel(const float *pred, const long long *row0,
                                                         const int *items, int m_high, int m_low,
                                                         float thr, int *sel,
                                                         const long long *sel_off, int *count,
                                                         float *pmax) {
    __shared__ unsigned hist[256];
    __shared__ unsigned s_prefix, s_mask, s_need;
    __shared__ int wtot[4];
    __shared__ float wmax[4];
    const int gi = blockIdx.x;
    const int n = items[gi];
    const float *pv = pred + row0[gi];
    int *out = sel + sel_off[gi];
    int m = m_high;
    {   // largest prediction of the group
        float mx = -INFINITY;
        for (int i = threadIdx.x; i < n; i += 256) mx = fmaxf(mx, pv[i]);
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
        __syncthreads();
        const float gmax = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        if (threadIdx.x == 0) pmax[gi] = gmax;
        if (gmax < thr && m_low > m) m = m_low;
    }
    if (n <= m) {
        for (int i = threadIdx.x; i < n; i += 256) out[i] = i;
        if (threadIdx.x == 0) count[gi] = n;
        return;
    }
    if (threadIdx.x == 0) { s_prefix = 0; s_mask = 0; s_need = (unsigned)m; }
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 24 - 8 * pass;
        hist[threadIdx.x] = 0;
        __syncthreads();
        const unsigned prefix = s_prefix, mask = s_mask;
        for (int i = threadIdx.x; i < n; i += 256) {
            const unsigned k = order_key(pv[i]);
            if ((k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned cum = 0, need = s_need;
            for (int d = 255; d >= 0; d--) {
                if (cum + hist[d] >= need) {
                    s_prefix = prefix | ((unsigned)d << shift);
                    s_mask = mask | (255u << shift);
                    s_need = need - cum;
                    break;
                }
                cum += hist[d];
            }
        }
        __syncthreads();
    }
    const unsigned T = s_prefix;
    const int need_eq = (int)s_need;
    int written = 0, eq_seen = 0;
    for (int c0 = 0; c0 < n; c0 += 256) {
        const int i = c0 + threadIdx.x;
        const unsigned k = i < n ? order_key(pv[i]) : 0u;
        const bool gt = i < n && k > T, eq = i < n && k == T;
        int eq_tot = 0;
        const int eq_rank = eq_seen + block_excl(eq, wtot, eq_tot);
        const bool take = gt || (eq && eq_rank < need_eq);
        int take_tot = 0;
        const int pos = written + block_excl(take, wtot, take_tot);
        if (take) out[pos] = i;
        written += take_tot;
        eq_seen += eq_tot;
    }
    if (threadIdx.x == 0) count[gi] = m;
}

__global__ void gather_rows_kernel(const float *src, const float *tsrc, const int *idx, int n,
                                   float *dst, float *tdst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * kFeat) return;
    const int r = (int)(i / kFeat), c = (int)(i - (long long)r * kFeat);
    const long long s = (long long)idx[r];
    dst[i] = src[s * kFeat + c];
    if (c == 0) tdst[r] = tsrc[s];
}

__global__ void offset_screen_kernel(Arena A, const OffsetItem *items, int n, float *score) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const OffsetItem it = items[i];
    const PairInfo &pi = A.info[it.pair];
    const int off = pi.off;
    KabschSeq ks;
    ks.reset();
    for (int q = 0; q < it.ov; q++) {
        const P4 P = A.pc4[off + it.ps + q], Q = A.qc4[off + it.ns + q];
        ks.add_moment(P.x, P.y, P.z, Q.x, Q.y, Q.z);
    }
    ks.finish_moments();
    double R[9], t[3];
    ks.template solve<true>(R, t);
    const double d0sq = pi.d0 * pi.d0;
    double acc = 0.0;
    for (int q = 0; q < it.ov; q++) {
        const P4 P = A.pc4[off + it.ps + q], Q = A.qc4[off + it.ns + q];
        const double d2 = dist2_moved_fma(R, t, P.x, P.y, P.z, Q.x, Q.y, Q.z);
        acc += d0sq / (d0sq + d2);
    }
    score[i] = (float)(acc / fmax(pi.Lnorm, 1.0));
}

__global__ void fill_none_kernel(Hit *out, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Hit h;
    set_identity(h.R, h.t);
    h.score = -INFINITY;
    out[i] = h;
}

__global__ void hit_scores_kernel(const Hit *in, long long n, int stride, float *y) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = (float)in[i * stride].score;
}

__global__ void item_best_kernel(const Hit *out, long long n, float *y) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = out[2 * i].score, b = out[2 * i + 1].score;
    y[i] = (float)(a > b ? a : b);
}

}  // namespace

void launch_offset_screen(hipStream_t s, const Arena &A, const OffsetItem *items, int n,
                          float *score) {
    if (n <= 0) return;
    hipLaunchKernelGGL(offset_screen_kernel, dim3((n + 255) / 256), dim3(256), 0, s, A, items, n,
                       score);
}

void launch_fill_none(hipStream_t s, Hit *out, long long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_none_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, n);
}

void launch_hit_scores(hipStream_t s, const Hit *in, long long n, int stride, float *y) {
    if (n <= 0) return;
    hipLaunchKernelGGL(hit_scores_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in,
                       n, stride, y);
}

void launch_item_best(hipStream_t s, const Hit *out, long long n_items, float *y) {
    if (n_items <= 0) return;
    hipLaunchKernelGGL(item_best_kernel, dim3((unsigned)((n_items + 255) / 256)), dim3(256), 0, s,
                       out, n_items, y);
}

void launch_frag_features(hipStream_t s, const Arena &A, const FeatGroup *groups,
                          const int *blk_group, const int *blk_first, int n_blocks, float *X) {
    if (n_blocks <= 0) return;
    hipLaunchKernelGGL(frag_feature_kernel, dim3(n_blocks), dim3(kFeatBlock), 0, s, A, groups,
                       blk_group, blk_first, X);
}

void launch_gemm(hipStream_t s, const GemmArgs &g) {
    if (g.M <= 0 || g.N <= 0) return;
    const dim3 grid((unsigned)((g.N + 63) / 64), (unsigned)((g.M + 63) / 64));
    hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, s, g);
}

void launch_weights_bf16(hipStream_t s, const float *P, short *W1b, short *W2b) {
    hipLaunchKernelGGL(weights_bf16_kernel, dim3((kHid * kHid + 255) / 256), dim3(256), 0, s, P,
                       W1b, W2b);
}

void launch_mlp_infer(hipStream_t s, const float *X, long long n, const short *W1b,
                      const short *W2b, const float *P, float *y) {
    for (long long r0 = 0; r0 < n; r0 += (1LL << 30)) {
        const int m = (int)((n - r0) < (1LL << 30) ? (n - r0) : (1LL << 30));
        hipLaunchKernelGGL(mlp_infer_kernel, dim3((m + kInferRows - 1) / kInferRows), dim3(256), 0,
                           s, X + r0 * kFeat, m, W1b, W2b, P, y + r0);
    }
}

void launch_head_back(hipStream_t s, const float *dy, const float *w3, const float *h2, int n,
                      float *dh2) {
    const long long tot = (long long)n * kHid;
    if (tot <= 0) return;
    hipLaunchKernelGGL(head_back_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, dy,
                       w3, h2, (long long)n, dh2);
}

void launch_mse(hipStream_t s, const float *y, const float *t, int n, float *dy, float *loss) {
    hipLaunchKernelGGL(mse_kernel, dim3(1), dim3(1024), 0, s, y, t, n, dy, loss);
}

void launch_adam(hipStream_t s, float *p, const float *g, float *m, float *v, long long n,
                 float lr, float b1, float b2, float eps, int step) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, g, m,
                       v, n, lr, b1, b2, eps, c1, c2);
}

void launch_select_top(hipStream_t s, const float *pred, const long long *row0,
                       const int *items, int n_groups, int m, int *sel, const long long *sel_off,
                       int *count, float *pmax, int m_low, float thr) {
    if (n_groups <= 0) return;
    hipLaunchKernelGGL(select_top_kernel, dim3(n_groups), dim3(256), 0, s, pred, row0, items, m,
                       m_low, thr, sel, sel_off, count, pmax);
}

void launch_gather_rows(hipStream_t s, const float *src, const float *tsrc, const int *idx,
                        int n, float *dst, float *tdst) {
    const long long tot = (long long)n * kFeat;
    if (tot <= 0) return;
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       src, tsrc, idx, n, dst, tdst);
}

}  // namespace ia
