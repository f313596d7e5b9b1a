// Device task descriptors and kernel launchers shared by the HIP kernels and
// the host engine.
#pragma once
#include "ia_common.h"

namespace ia {

// ---- seed bank (one lane per candidate) ---------------------------------
enum PointKind { PTS_FULL = 0, PTS_LIST = 1, PTS_CWIN = 2 };
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
struct BankDesc {
    int pair, pts, n, a0, a1;
    int seeds, seed_off, K, max_starts;
    int C, iters, use_d8, init, min_ali;
    int cand_cap;
    long long cand_off;  // first candidate slot
    long long out_idx;   // slot of the reduced result
    double et;           // early termination threshold (search.h:83); huge = never
    double sd8;          // d8 cutoff when use_d8
    double cut[kMaxCut];
};

// ---- NW chains (one wavefront per chain) ---------------------------------
enum ChainInit { CI_XFORM = 0, CI_FRAG = 1, CI_OFFSET = 2 };
enum ChainMode { CM_LITE = 0, CM_LIST = 1 };
enum ScoreMode { SM_TM = 0, SM_SS = 1, SM_IA4 = 2 };

// A group of chains of one pair that share everything but the start.
//   CI_XFORM:  chain of item i, gap g starts from
//              xf[xf_off + i*xf_stride + (xf_per_chain ? g : 0)].
//   CI_FRAG:   local-fragment rescue work items (rescue.h:69-78): item
//              index -> (flen, ps, ns) over compact indices; start = Kabsch
//              of the fragment pair (rescue.h:84-91).
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
};

// Per-chain extra outputs of CM_LIST chains.
struct ListOut {
    double R[9], t[3];
    double it_score;
    int na;
    int pad;
};

// Launchers (defined in the .hip files; all asynchronous on `s`).
void launch_bank(hipStream_t s, const Arena &A, const BankDesc *banks, const int *blk_bank,
                 const int *blk_first, int n_blocks, const SeedDesc *seeds,
                 const uint8_t *masks, const int *lists_p, const int *lists_n,
                 const int *list_na, int list_cap, Hit *cands, bool fast);
void launch_bank_reduce(hipStream_t s, const BankDesc *banks, int n_banks, const int *list_na,
                        const Hit *cands, Hit *out);
void launch_group_reduce(hipStream_t s, const long long *offs, const int *counts, int n,
                         const Hit *in, Hit *out);
void launch_rescore(hipStream_t s, const Arena &A, const int *pair_of, const Hit *in, int n,
                    Hit *out);
void launch_topk(hipStream_t s, const Arena &A, const int *pairs, const Hit *best, int n,
                 const long long *mask_off, uint8_t *masks);
void launch_weighted(hipStream_t s, const Arena &A, const int *pairs, const Hit *best, int n,
                     Hit *out);
void launch_ss(hipStream_t s, const Arena &A, int n_pairs, char *ssp, char *ssn);
void launch_anchors(hipStream_t s, const Arena &A, const int *pairs, int n,
                    const SeedDesc *frag_seeds, const int *frag_off, const int *frag_cnt,
                    uint8_t *anchor_masks, const long long *anchor_off, int *anchor_cnt,
                    int *scratch_i, const long long *scratch_off);
void launch_chains(hipStream_t s, const Arena &A, const ChainGroup *groups, int n_groups,
                   long long n_chains, const Hit *xf, Hit *out, ListOut *lout, int *lists_p,
                   int *lists_n, int *list_na, int list_cap, int max_plen, int max_nlen,
                   int *queue, unsigned *trace_scratch, long long trace_slot_words,
                   int n_slots, bool cols_in_lds, bool fast);
size_t chain_lds_bytes(int max_plen, int max_nlen, bool trace_in_lds, bool cols_in_lds);
int chain_blocks_per_cu(size_t lds);

}  // namespace ia
