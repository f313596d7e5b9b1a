// Device task descriptors and kernel launchers shared by the HIP kernels and
// the host engine.
#pragma once
#include "ia_common.h"

namespace ia {

// ---- seed bank (one lane per candidate) ---------------------------------
enum PointKind { PTS_FULL = 0, PTS_LIST = 1, PTS_CWIN = 2, PTS_TLIST = 3 };
enum SeedKind { SEED_EXPLICIT = 0, SEED_IMPLICIT = 1 };
enum InitKind { INIT_SEED = 0, INIT_OFFSET = 1 };

// A seed: window [a, a+b) when b > 0; explicit byte mask at offset a when
// b == -1; gapless offset pairing pred[i] <-> native[i-a] when b == -2.
struct SeedDesc {
    int a, b;
};

// One bank = one evaluate_seed_bank call (search.h:62-138) over one point set.
//   PTS_FULL: the pair's residues, valid mask applies.
//   PTS_LIST: aligned pairs (lp[k], ln[k]) from an NW traceback list slot
//             (a0 = list slot); the count comes from list_na[a0].
//   PTS_CWIN: compact window pmap[a0 + k] <-> nmap[a1 + k], k < n.
//   PTS_TLIST: a traceback list of the PyTorch engine's DP (SM_TORCH
//             chains), list slot a0: its entries carry a validity flag,
//             window seeds start anywhere in [0, n) (n = the pair's N, the
//             list's padded length) with lengths from the list's valid
//             count list_na[a1], and a seed with fewer than 3 valid entries
//             yields no candidate (search.py:47-122, dp.py:133-183).
//             Only the bank kernels' TLIST instantiations run these banks,
//             and they run nothing else.
struct BankDesc {
    int pair, pts, n, a0, a1;
    int seeds, seed_off, K, max_starts;
    int C, iters, use_d8, init, min_ali;
    int cand_cap;
    int fixed_iters;     // 1: no convergence break, every candidate runs `iters` passes
    long long cand_off;  // first candidate slot
    long long out_idx;   // slot of the reduced result
    double et;           // early termination threshold (search.h:83); huge = never
    double sd8;          // d8 cutoff when use_d8
    double cut[kMaxCut];
    int sel_m;           // > 0: only seeds A.isel[sel_off + i], i < min(seeds, sel_m)
    long long sel_off;   //      (a screened subset in increasing seed order)
};

// ---- NW chains (one wavefront per chain) ---------------------------------
enum ChainInit { CI_XFORM = 0, CI_FRAG = 1, CI_OFFSET = 2 };
enum ChainMode { CM_LITE = 0, CM_LIST = 1 };
// SM_MAT: score read from a supplied matrix (the NW test entry only)
// SM_TORCH: the PyTorch engine's DP score matrix (dp.py:9-28) over all N x N
// residue pairs, d2 = max(|p|^2 + |q|^2 - 2 p.q, 0), 1 / (1 + d2 / d0^2)
// kept where d2 <= score_d8^2 and both residues are valid, NW with no gap
// penalty and ties diag, then up, then left (dp.py:31-56)
enum ScoreMode { SM_TM = 0, SM_SS = 1, SM_IA4 = 2, SM_MAT = 3, SM_TORCH = 4 };

// A group of chains of one pair that share everything but the start.
//   CI_XFORM:  chain of item i, gap g starts from
//              xf[xf_off + i*xf_stride + (xf_per_chain ? g : 0)].
//   CI_FRAG:   local-fragment rescue work items (rescue.h:69-78): item
//              index -> (flen, ps, ns) over compact indices; start = Kabsch
//              of the fragment pair (rescue.h:84-91).  With sel != 0 the
//              chain's item is A.isel[sel_off + i] (a selected subset, in
//              increasing item order) and results stay at index i.
//              CI_XFORM with sel != 0: chain i runs item A.isel[sel_off + i]
//              in place (its start and its result slots).
//   CI_OFFSET: offset k = k0 + (c/2)*kstep; start = Kabsch of the pairing
//              pred[i] <-> native[i-k] (indialign.cpp:234-247); k == 0 or
//              fewer than 3 pairs -> the chain reports "skipped".
// SM_TM chains come in gap-open pairs (g = 0: -0.6, g = 1: 0.0, align.h:80),
// chain c = 2*item + g; SM_SS / SM_IA4 chains are single (gap gap_ss), c =
// item.  The chain result goes to out[out_off + item*out_stride + g]; CM_LIST
// chains write list slot list_slot0 + c instead.
struct ChainGroup {
    int pair, init, mode, smat;
    int n_chains;
    int iters;
    long long chain_off;  // prefix over groups (for task -> group search)
    long long out_off;
    int xf_off, xf_stride, xf_per_chain;
    int out_stride;
    int list_slot0;       // CM_LIST: first list slot (one per chain)
    int k0, kstep;        // CI_OFFSET
    int n_lens, jp, jn;   // CI_FRAG
    int lens[9];
    int items_prefix[10];
    int n_ns[9];
    double sd8;
    double gap_ss;        // gap used by SM_SS / SM_IA4 chains (ss.h:75,144)
    int sel;              // CI_FRAG: 0 = every item, 1 = the items A.isel[sel_off + i]
    long long sel_off;
    int nv_slot0;         // SM_TORCH CM_LIST: list_na[nv_slot0 + c] = the list's valid count
};

// Per-chain extra outputs of CM_LIST chains.
struct ListOut {
    double R[9], t[3];
    double it_score;
    int na;
    int nvalid;  // entries whose residues are both valid (= na except SM_TORCH)
};

// Launchers (defined in the .hip files; all asynchronous on `s`).
// pass_cap > 0: candidates still refining after pass_cap passes are left for
// launch_bank_cont (their index appended to cont_idx[cand_off + k] with
// k < cont_cnt[bank], which must start at zero; cont_state holds 8 bytes per
// candidate slot).
void launch_bank(hipStream_t s, const Arena &A, const BankDesc *banks, const int *blk_bank,
                 const int *blk_first, int n_blocks, const SeedDesc *seeds,
                 const uint8_t *masks, const P8 *list_pq, const int *list_na, int list_cap,
                 Hit *cands, bool fast, int pass_cap, int *cont_cnt, int *cont_idx,
                 void *cont_state, bool tlist);
// The rest of those candidates: block b runs entries blk_first[b] .. +255 of
// bank blk_bank[b]'s list.
void launch_bank_cont(hipStream_t s, const Arena &A, const BankDesc *banks, const int *blk_bank,
                      const int *blk_first, int n_blocks, const P8 *list_pq, const int *list_na,
                      int list_cap, Hit *cands, bool fast, const int *cont_cnt,
                      const int *cont_idx, const void *cont_state, bool tlist);
// Interleaved copies of the traceback lists of the PTS_LIST banks, read by
// launch_bank (list_pq[slot * list_cap + k]).
void launch_list_gather(hipStream_t s, const Arena &A, const BankDesc *banks, int n_banks,
                        const int *lists_p, const int *lists_n, const int *list_na, int list_cap,
                        P8 *list_pq);
// out[b] = number of seeds of bank b (implicit seeds depend on the list
// length, known on the device).
void launch_bank_seed_count(hipStream_t s, const BankDesc *banks, int n_banks, const int *list_na,
                            int *out);
void launch_bank_reduce(hipStream_t s, const BankDesc *banks, int n_banks, const int *list_na,
                        const Hit *cands, Hit *out);
void launch_group_reduce(hipStream_t s, const long long *offs, const int *counts, int n,
                         const Hit *in, Hit *out);
// Same-index TM-score of each entry's (R, t); keep_none leaves entries
// without a result (score kNeg) as they are.
void launch_rescore(hipStream_t s, const Arena &A, const int *pair_of, const Hit *in, int n,
                    Hit *out, bool keep_none = false);
void launch_topk(hipStream_t s, const Arena &A, const int *pairs, const Hit *best, int n,
                 const long long *mask_off, uint8_t *masks);
void launch_weighted(hipStream_t s, const Arena &A, const int *pairs, const Hit *best, int n,
                     Hit *out);
// scratch: 2 * (sum of N) * 9 ints
void launch_ss(hipStream_t s, const Arena &A, int n_pairs, char *ssp, char *ssn, int *scratch);
void launch_anchors(hipStream_t s, const Arena &A, const int *pairs, int n,
                    const SeedDesc *frag_seeds, const int *frag_off, const int *frag_cnt,
                    uint8_t *anchor_masks, const long long *anchor_off, int *anchor_cnt,
                    int *scratch_i, const long long *scratch_off, int max_n);
void launch_chains(hipStream_t s, const Arena &A, const ChainGroup *groups, int n_groups,
                   long long n_chains, const Hit *xf, Hit *out, ListOut *lout, int *lists_p,
                   int *lists_n, int *list_na, int list_cap, int max_plen, int max_nlen,
                   int *queue, unsigned *trace_scratch, long long trace_slot_words,
                   int n_slots, bool cols_in_lds, bool fast, int waves, P8 *gq,
                   long long gq_stride, bool torch);
// Needleman-Wunsch over supplied score matrices (test entry): problem k has
// a plen x nlen row-major matrix at smat + s_off, its H and trace codes go
// to H / T + h_off ((plen + 1) x (nlen + 1), interior only) and its aligned
// pairs (increasing order) to ap / an + a_off; one wavefront per problem.
struct NwProblem {
    long long s_off, h_off, a_off;
    int plen, nlen;
    double gap;
};
// Largest compact native length whose NW chain fits one wavefront's LDS
// (the row hand-over buffer, 16 bytes per column).
int chain_max_nlen();
size_t nw_matrix_lds_bytes(int max_plen, int max_nlen, bool trace_in_lds);
long long nw_matrix_trace_words(int max_plen, int max_nlen);
void launch_nw_matrix(hipStream_t s, const NwProblem *probs, int n, const double *smat, double *H,
                      signed char *T, int *ap, int *an, int *na, int max_plen, int max_nlen,
                      unsigned *trace_g, bool fast);
// ---- learned ranking of local-fragment work items ------------------------
constexpr int kFeat = 96;   // descriptor width
constexpr int kHid = 256;   // hidden units per layer
// MLP parameters, one FP32 array: W1[kHid][kFeat], b1[kHid], W2[kHid][kHid],
// b2[kHid], w3[kHid], b3.
constexpr long long kOffW1 = 0;
constexpr long long kOffB1 = kOffW1 + (long long)kHid * kFeat;
constexpr long long kOffW2 = kOffB1 + kHid;
constexpr long long kOffB2 = kOffW2 + (long long)kHid * kHid;
constexpr long long kOffW3 = kOffB2 + kHid;
constexpr long long kOffB3 = kOffW3 + kHid;
constexpr long long kNParam = kOffB3 + 1;
// The ranker's output is the network's value plus descriptor column
// kSkipCol (the twice-refined same-index TM-score): the network learns a
// correction to that estimate.
constexpr int kSkipCol = 31;

// The local-fragment work items of one pair (the CI_FRAG enumeration of its
// chain group) occupy rows [row0, row0 + items) of the feature matrix.
struct FeatGroup {
    int pair, items, n_lens, jp, jn;
    int lens[9];
    int items_prefix[10];
    int n_ns[9];
    long long row0;
    float prior;   // running best score of the pair before the stage
    float dense;   // 1 for the dense pass, 0 for the first pass
};
// Descriptor rows of all work items; block b covers items
// [blk_first[b], blk_first[b] + 128) of group blk_group[b].
constexpr int kFeatItemsPerBlock = 128;
void launch_frag_features(hipStream_t s, const Arena &A, const FeatGroup *groups,
                          const int *blk_group, const int *blk_first, int n_blocks, float *X);

// C(m,n) = epi(sum_k A(m,k) B(k,n)) with A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn], C(m,n) = C[m*scm + n*scn]; operands rounded to
// BF16, products accumulated in FP32 on the matrix cores.
enum GemmEpi { EPI_STORE = 0, EPI_BIAS_RELU = 1, EPI_BIAS = 2, EPI_MASK = 3, EPI_BIAS_SKIP = 4 };
struct GemmArgs {
    const float *A, *B;
    float *C;
    long long sam, sak, sbk, sbn, scm, scn;
    int M, N, K, epi;
    const float *bias;          // EPI_BIAS*: bias[n]
    const float *mask;          // EPI_MASK: C *= (mask(m,n) > 0), same strides as C
    const float *skip;          // EPI_BIAS_SKIP: C += skip[m * sskip] (FP32, not rounded)
    long long sskip;
};
void launch_gemm(hipStream_t s, const GemmArgs &g);
// Ranker inference fused in one kernel (same rounding points as the GEMM
// path): y[0..n) from X rows, parameters P and their BF16 weight copies.
void launch_weights_bf16(hipStream_t s, const float *P, short *W1b, short *W2b);
void launch_mlp_infer(hipStream_t s, const float *X, long long n, const short *W1b,
                      const short *W2b, const float *P, float *y);
// dh2(m,n) = dy[m] * w3[n] * (h2(m,n) > 0)
void launch_head_back(hipStream_t s, const float *dy, const float *w3, const float *h2, int n,
                      float *dh2);
// dy[m] = 2 (y[m] - t[m]) / n and loss = mean (y - t)^2 (FP32, one value)
void launch_mse(hipStream_t s, const float *y, const float *t, int n, float *dy, float *loss);
void launch_adam(hipStream_t s, float *p, const float *g, float *m, float *v, long long n,
                 float lr, float b1, float b2, float eps, int step);
// Per group: indices (relative to the group) of the top-m predictions, ties
// to the lower index, written in increasing index order to
// sel[sel_off[g] ...]; pmax[g] = largest prediction; a group whose pmax is
// below thr takes m_low instead of m when m_low > m; count[g] = min(that m,
// items).
void launch_select_top(hipStream_t s, const float *pred, const long long *row0,
                       const int *items, int n_groups, int m, int *sel, const long long *sel_off,
                       int *count, float *pmax, int m_low = 0, float thr = 0.f);

// out[i] = "no result" (score -inf, identity) for i < n.
void launch_fill_none(hipStream_t s, Hit *out, long long n);
// y[i] = (float) in[i * stride].score for i < n.
void launch_hit_scores(hipStream_t s, const Hit *in, long long n, int stride, float *y);
// Threading-DP offset screening: per offset window (compact pred rows
// [ps, ps+ov) against native rows [ns, ns+ov)), the TM-score (d0 of the
// pair, normalized by Lnorm) of the window's own superposition; one lane per
// offset.
struct OffsetItem {
    int pair, ps, ns, ov;
};
void launch_offset_screen(hipStream_t s, const Arena &A, const OffsetItem *items, int n,
                          float *score);
// y[i] = best score of the two chains of work item i (out[2i], out[2i+1]).
void launch_item_best(hipStream_t s, const Hit *out, long long n_items, float *y);
// dst[r] = src[idx[r]] (rows of kFeat values), tdst[r] = tsrc[idx[r]]
void launch_gather_rows(hipStream_t s, const float *src, const float *tsrc, const int *idx,
                        int n, float *dst, float *tdst);

size_t chain_lds_bytes(int max_plen, int max_nlen, bool trace_in_lds, bool cols_in_lds);
// 32-bit trace words of the largest chain fill (max_plen x max_nlen)
long long chain_trace_words(int max_plen, int max_nlen);
// resident chain wavefronts per CU for the W = waves variant
int chain_blocks_per_cu(size_t lds, bool fast, int waves, bool torch);

}  // namespace ia
