// Host engine and C ABI of the gfx950 TM-score superposition search.
//
// The engine runs the stages of the reference pipeline (indialign.cpp:29-266)
// in the same order and with the same merge rules, for a whole batch of pairs
// at a time.  Every stage is a batched device launch over all pairs that
// take part in it; the host only enumerates work (seed windows, rescue work
// items, offsets), applies the stage gates and performs the strict-'>' merges
// on per-pair results copied back after each stage.
#include "host_util.h"
#include "ia_kernels.h"
#include "indialign.h"
#include "mlp.h"

#include <dlfcn.h>
#include <sched.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace ia {


namespace {

Hit identity_hit(double s) {
    Hit h;
    set_identity(h.R, h.t);
    h.score = s;
    return h;
}

// d0 normalisation on the host with the C library (core.h:12-25).
void d0_params(double L, double &d0, double &d0s, double &sd8) {
    const double Lc = std::max(L, 1.0);
    if (Lc <= 11.0) d0 = 0.3;
    else if (Lc <= 15.0) d0 = 0.4;
    else if (Lc <= 19.0) d0 = 0.5;
    else if (Lc <= 23.0) d0 = 0.6;
    else d0 = 0.7;
    if (Lc >= 30.0) d0 = 0.6 * std::sqrt(Lc - 0.5) - 2.5;
    if (d0 < 0.3) d0 = 0.3;
    const double v = 1.24 * std::cbrt(std::max(Lc - 15.0, 0.0)) - 1.8;
    d0s = std::clamp(v, 4.5, 8.0);
    sd8 = 1.5 * std::pow(Lc, 0.3) + 3.5;
}

// Rescue schedules (rescue.h:12-46).
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

struct HostPair {
    int src;  // index in the caller's array
    int N, V, plen, nlen;
    bool cross;
    PairInfo info;
    std::vector<int> pre;        // prefix counts of valid residues
    std::vector<SeedDesc> frag;  // global + window seeds (seeds.h:41-82)
};

constexpr double kNever = 1e300;  // early termination disabled

}  // namespace

constexpr int kNStats = 12;
constexpr int kTooLong = -6;      // return code: some pair exceeds chain_max_nlen()
constexpr int kUnsupported = -7;  // return code: method 3 with separate pred/native masks
constexpr int kChainClasses = 4;  // compact native length <= 160, 400, 900, longer

class Engine {
  public:
    Engine() {
        IA_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
        int dev = 0;
        IA_CHECK(hipGetDevice(&dev));
        IA_CHECK(hipDeviceGetAttribute(&n_cu_, hipDeviceAttributeMultiprocessorCount, dev));
        IA_CHECK(hipEventCreate(&e0_));
        IA_CHECK(hipEventCreate(&e1_));
        for (int k = 0; k < kChainClasses; k++) {
            IA_CHECK(hipStreamCreateWithFlags(&cs_[k], hipStreamNonBlocking));
            IA_CHECK(hipEventCreateWithFlags(&cev_[k], hipEventDisableTiming));
        }
    }
    ~Engine() {
        for (DBuf *b : all_bufs()) b->release();
        (void)hipEventDestroy(e0_);
        (void)hipEventDestroy(e1_);
        for (int k = 0; k < kChainClasses; k++) {
            (void)hipEventDestroy(cev_[k]);
            (void)hipStreamDestroy(cs_[k]);
        }
        (void)hipStreamDestroy(s_);
    }

    // Copy a batch into device memory (kept resident until the next load).
    void load(const NativePairInput *inp, int n);
    // Search the resident batch; out has one slot per loaded pair.
    // method 0: plain FP64 (bit-identical to the reference's arithmetic);
    // method 1: fast (FP32 NW fill, closed-form FP64 rotations).
    // method 2: learned (method 1 + only the ranker's top-m local-fragment
    //           work items of each pair are refined).
    // serial_bank: 1 = early termination of the main seed bank applies (the
    // reference's serial path), 0 = it never applies (its OpenMP path);
    // -1 = per pair, the reference's gate (search.h:14-24) with
    // omp_threads host threads: the serial path when omp_threads <= 1 or
    // the bank's K*C candidates are fewer than 4 per thread.
    int search(const NativeConfig *cfg, int serial_bank, NativePairResult *out, int method = 0,
               int omp_threads = 1);
    int loaded() const { return n_loaded_; }
    double stats[kNStats] = {};
    // Running best score of every loaded pair after each of kStages stages
    // (NaN: the stage did not run), row-major [pair][stage].
    static constexpr int kStages = 10;
    std::vector<double> stage_scores;
    // Work-item ranker used by method 2 and its per-pair budget.
    Mlp ranker;
    int top_m = 64;                 // dense local-fragment pass
    int top_m_first = 16;           // first local-fragment pass
    double fallback_margin = -1.0;  // < 0: never fall back to every work item
    int top_offsets = 16;           // threading-DP offsets per pair that start chains
    int top_offset_banks = 16;      // threading-DP offsets per pair that run their bank
    int top_seeds = 0;              // seeds refined per SS-stage and threading bank (0 = all)
    // Fallback budgets: a pair whose largest predicted refined score of a
    // local-fragment pass is below that pass's rescue threshold (the case
    // where the plain search is the fallback) refines this many items.
    int fallback_top_m = 512;        // dense pass (below rescue_score_2)
    int fallback_top_m_first = 64;   // first pass (below rescue_score_1)
    // Method 2: a rescue pass keeps its work item whose transform has the
    // best same-index TM-score (true) or, as the plain search, the item with
    // the best chain score, rescored afterwards (false, the default: it keeps
    // the learned search on the plain search's trajectory).
    bool reduce_rescored = false;
    // Chain-kernel launch shape (indialign_set_launch_tuning): wavefronts per
    // SIMD of the kernel variant (2 or 3), persistent slots per CU (0: as many
    // as are resident), NW trace placement (0: LDS only when that costs no
    // resident wavefront, 1: LDS whenever it fits, 2: always a global slot).
    int chain_waves = 3;
    int chain_slots_per_cu = 0;
    int chain_trace = 0;
    // length classes of the chain launches (1: one launch for every chain,
    // the default: one queue ordered longest first balances the short-pair
    // batch best; the classes only pay off when a few very long chains
    // would set a low-occupancy LDS plan for everything)
    int chain_buckets = 1;
    // Seed banks: refinement passes a candidate runs in the first launch
    // before it is left to the continuation launch (0: one launch).
    int bank_pass_cap = 5;
    // Record sink: when set, every local-fragment work item of the search
    // appends its descriptor row, its refined score (best of its two
    // chains) and its pair index, up to cap rows.
    struct RecordSink {
        float *X, *y;
        int *pair;
        long long cap, n;
    };
    RecordSink *rec = nullptr;
    hipStream_t stream() const { return s_; }

  private:
    hipStream_t s_;
    hipEvent_t e0_, e1_;
    // streams of the concurrent chain launches of the length classes
    hipStream_t cs_[kChainClasses] = {};
    hipEvent_t cev_[kChainClasses] = {};
    int n_cu_ = 256;
    const NativeConfig *cfg_ = nullptr;
    int serial_ = 1;
    int omp_threads_ = 1;
    bool fast_ = false;
    bool learned_ = false;
    // method 3: the seed search of the reference's PyTorch engine
    // (scoring/_tmscore_gpu/search.py:492-761 with dp_iter = 0 and
    // same-index masks): window seeds, no anchor seeds, cutoffs d0_search x
    // {2, 1.5, 1, .75, .5, .25} only, a fixed number of refinement passes
    // (no convergence break, search.py:423-489), candidates ranked by the
    // TM-score with the score_d8 cut (runtime.py:40), then the top-k seed,
    // and the reported score max(TM-score without the cut, ranking score)
    // (search.py:744-757).
    bool torch_sem_ = false;
    void bank_cutoffs(BankDesc &b, const HostPair &h) const {
        static const double mult[6] = {2.0, 1.5, 1.0, 0.75, 0.5, 0.25};
        if (torch_sem_) {
            b.C = 6;
            for (int c = 0; c < 6; c++) b.cut[c] = h.info.d0s * mult[c];
            b.use_d8 = 1;
            b.fixed_iters = 1;
        } else {
            b.C = h.info.C;
            std::memcpy(b.cut, h.info.cut, sizeof(b.cut));
            b.use_d8 = 0;
            b.fixed_iters = 0;
        }
        b.sd8 = h.info.sd8;
    }
    std::vector<HostPair> hp_;
    std::vector<Hit> best_;
    std::vector<int> live_;  // pairs still in the pipeline (index into hp_)
    int n_loaded_ = 0;
    // per loaded pair: 0 search, 1 Lnorm <= 2, -1 N <= 0, 2 more native-valid
    // residues than one chain wavefront's LDS holds (chain_max_nlen())
    std::vector<signed char> status_;
    Arena A_{};
    int list_cap_ = 1;
    int n_res_ = 0;  // residues of all loaded pairs

    void ensure_lists(int slots, int cap) {
        list_cap_ = std::max(cap, 1);
        b_lists_p.get<int>((size_t)slots * list_cap_);
        b_lists_n.get<int>((size_t)slots * list_cap_);
        b_list_na.get<int>(slots);
        b_lout.get<ListOut>(slots);
    }
    // Full detailed search (align.h:15-50, use_frag) over a traceback list.
    BankDesc list_bank(int pair, int slot, int cap, double sd8, long long out_idx) {
        BankDesc b{};
        b.pair = pair;
        b.pts = PTS_LIST;
        b.n = -1;
        b.a0 = slot;
        b.seeds = SEED_IMPLICIT;
        b.max_starts = 50;
        b.C = 6;
        b.iters = 20;
        b.use_d8 = 1;
        b.init = INIT_SEED;
        b.cand_cap = cap;
        b.out_idx = out_idx;
        b.et = 1.0;
        b.sd8 = sd8;
        static const double mult[6] = {2.0, 1.5, 1.0, 0.75, 0.5, 0.25};
        for (int c = 0; c < 6; c++) b.cut[c] = hp_[pair].info.d0s * mult[c];
        return b;
    }
    // The PyTorch engine's detailed search over a DP list (dp.py:133-183):
    // window seeds over the padded list length N with lengths from the
    // list's valid count list_na[nv_idx], six cutoffs d0_search x mult,
    // max_iter = 20 fixed passes, candidates ranked with the score_d8 cut.
    BankDesc torch_list_bank(int pair, int slot, int nv_idx, long long out_idx) {
        BankDesc b = list_bank(pair, slot, 1, hp_[pair].info.sd8, out_idx);
        b.pts = PTS_TLIST;
        b.n = hp_[pair].N;
        b.a1 = nv_idx;
        b.fixed_iters = 1;
        b.et = kNever;
        // the most seeds any valid count of a list of this length yields
        int kmax = 1;
        for (int vl = 0; vl <= b.n; vl++) kmax = std::max(kmax, torch_list_seed_count(b.n, vl, b.max_starts));
        b.cand_cap = kmax * b.C;
        return b;
    }
    // dp_refine of the PyTorch engine (dp.py:186-252) from start, iters
    // iterations, same-index masks; returns the best (R, t, score), score 0
    // and the start transform when no iteration scores above 0.
    std::vector<Hit> torch_dp_refine(const std::vector<int> &pairs, const std::vector<Hit> &start,
                                     int iters);
    void profile(const char *stage, std::chrono::steady_clock::time_point t0);

    // arena
    DBuf b_px, b_py, b_pz, b_qx, b_qy, b_qz, b_v, b_pv, b_nv, b_pcx, b_pcy, b_pcz, b_qcx, b_qcy,
        b_qcz, b_pmap, b_nmap, b_ssp, b_ssn, b_info, b_qfx, b_qfy, b_qfz, b_p4, b_q4, b_pc4,
        b_qc4, b_ss_scr, b_fblk_g, b_fblk_f, b_pq8, b_list_pq, b_sbanks, b_sblk_bank,
        b_sblk_first, b_scands, b_oitems;
    // work buffers
    DBuf b_seeds, b_masks, b_banks, b_blk_bank, b_blk_first, b_cands, b_bank_out, b_lists_p,
        b_lists_n, b_list_na, b_lout, b_groups, b_xf, b_chain_out, b_queue, b_trace, b_red_off,
        b_red_cnt, b_red_out, b_rs_pair, b_rs_in, b_rs_out, b_pairs, b_aoff, b_acnt, b_ascr,
        b_soff, b_fofs, b_fcnt, b_moff, b_feat, b_fgroups, b_pred, b_isel, b_srow, b_sitems,
        b_ssel, b_scnt, b_itemy, b_pmax, b_cont_cnt, b_cont_idx, b_cont_state, b_cblk_bank,
        b_cblk_first, b_rr_pair, b_rr_out, b_bank_work, b_gq;

    std::vector<DBuf *> all_bufs() {
        return {&b_px, &b_py, &b_pz, &b_qx, &b_qy, &b_qz, &b_v, &b_pv, &b_nv, &b_pcx, &b_pcy,
                &b_pcz, &b_qcx, &b_qcy, &b_qcz, &b_pmap, &b_nmap, &b_ssp, &b_ssn, &b_info,
                &b_qfx, &b_qfy, &b_qfz, &b_p4, &b_q4, &b_pc4, &b_qc4, &b_ss_scr, &b_fblk_g, &b_fblk_f, &b_pq8, &b_list_pq, &b_sbanks, &b_sblk_bank, &b_sblk_first,
                &b_scands, &b_oitems,
                &b_seeds, &b_masks, &b_banks, &b_blk_bank, &b_blk_first, &b_cands,
                &b_bank_out, &b_lists_p, &b_lists_n, &b_list_na, &b_lout, &b_groups, &b_xf,
                &b_chain_out, &b_queue, &b_trace, &b_red_off, &b_red_cnt, &b_red_out,
                &b_rs_pair, &b_rs_in, &b_rs_out, &b_pairs, &b_aoff, &b_acnt, &b_ascr, &b_soff,
                &b_fofs, &b_fcnt, &b_moff, &b_feat, &b_fgroups, &b_pred, &b_isel, &b_srow,
                &b_sitems, &b_ssel, &b_scnt, &b_itemy, &b_pmax, &b_cont_cnt, &b_cont_idx,
                &b_cont_state, &b_cblk_bank, &b_cblk_first, &b_rr_pair, &b_rr_out, &b_bank_work,
                &b_gq};
    }

    template <class T>
    T *upload(DBuf &b, const std::vector<T> &v) {
        T *p = b.get<T>(v.size());
        if (!v.empty())
            IA_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s_));
        return p;
    }
    template <class T>
    std::vector<T> download(const T *p, size_t n) {
        std::vector<T> v(n);
        if (n) IA_CHECK(hipMemcpyAsync(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost, s_));
        IA_CHECK(hipStreamSynchronize(s_));
        return v;
    }
    void launched() { stats[4] += 1; }

    void prepare(const NativePairInput *inp, int n);
    void enumerate_seeds();
    void stage_seed_bank();
    void stage_ss();
    void stage_gapless();
    void stage_rescue();
    void stage_topk();
    void stage_dp();
    void stage_weighted();
    void torch_final();

    // Evaluate banks; returns device results indexed by out_idx (n_out).
    // screen > 0: first score every seed's initial superposition (one
    // cutoff, no refinement), then refine only each bank's top `screen`
    // seeds at every cutoff.
    Hit *run_banks(std::vector<BankDesc> &banks, const std::vector<SeedDesc> &seeds,
                   size_t n_out, Hit *out_dev = nullptr, int screen = 0);
    void screen_seeds(std::vector<BankDesc> &banks, const SeedDesc *seeds_dev, int screen);
    // Launch chain groups (chain_off filled here); returns number of chains.
    long long run_chains(std::vector<ChainGroup> &groups, const Hit *xf_dev, Hit *out_dev);
    std::vector<Hit> rescore(const std::vector<int> &pairs, const std::vector<Hit> &in);
    std::vector<Hit> group_reduce(const std::vector<long long> &offs,
                                 const std::vector<int> &cnts, const Hit *in_dev);
    // Per group, the entry with the best same-index TM-score (groups with no
    // entry: the rescored (kNeg, identity) of the plain path); owner[k] is
    // the pair of group k.
    std::vector<Hit> reduce_rescored_groups(const std::vector<long long> &offs,
                                            const std::vector<int> &cnts, const Hit *in_dev,
                                            const std::vector<int> &owner);
    // dp_refine (align.h:149-159) for a list of (pair, start); lite selects the
    // fused one-launch path, non-lite the per-iteration list + bank path.
    std::vector<Hit> dp_refine(const std::vector<int> &pairs, const std::vector<Hit> &start,
                               int iters, bool lite, int screen = 0);
    void offer(int a, const Hit &h) {
        if (h.score > best_[a].score) best_[a] = h;
    }
    void cells_nw(int a, long long chains_times_iters) {
        stats[2] += (double)hp_[a].plen * hp_[a].nlen * (double)chains_times_iters;
    }
    std::vector<int> max_seed_table(int nmax, int ms);
};

void Engine::prepare(const NativePairInput *inp, int n) {
    hp_.clear();
    std::vector<double> px, py, pz, qx, qy, qz, pcx, pcy, pcz, qcx, qcy, qcz;
    std::vector<uint8_t> v, pv, nv;
    std::vector<int> pmap, nmap;
    std::vector<PairInfo> info;
    int off = 0;
    for (int k = 0; k < n; k++) {
        const NativePairInput &I = inp[k];
        HostPair h{};
        h.src = k;
        h.N = I.length;
        if (status_[k] != 0) continue;  // handled by the caller
        const int N = h.N;
        h.info.N = N;
        h.info.off = off;
        h.info.Lnorm = I.lnorm;
        d0_params(I.lnorm, h.info.d0, h.info.d0s, h.info.sd8);
        // d0 search list (indialign.cpp:54-67)
        static const double mult[6] = {2.0, 1.5, 1.0, 0.75, 0.5, 0.25};
        int C = 0;
        for (double m : mult) h.info.cut[C++] = h.info.d0s * m;
        for (int a = 1; a <= 5; a++) {
            const double val = h.info.d0 + a;
            bool dup = false;
            for (int e = 0; e < C && !dup; e++) dup = std::fabs(h.info.cut[e] - val) < 0.01;
            if (!dup) h.info.cut[C++] = val;
        }
        h.info.C = C;
        int V = 0, pl = 0, nl = 0;
        bool cross = false;
        std::vector<int> pre(N + 1, 0);
        for (int i = 0; i < N; i++) {
            const uint8_t vi = I.valid_mask[i] ? 1 : 0;
            const uint8_t pi = I.pred_valid_mask[i] ? 1 : 0;
            const uint8_t ni = I.native_valid_mask[i] ? 1 : 0;
            V += vi;
            pre[i + 1] = pre[i] + vi;
            cross |= pi != ni;
            px.push_back(I.pred_xyz[3 * i]);
            py.push_back(I.pred_xyz[3 * i + 1]);
            pz.push_back(I.pred_xyz[3 * i + 2]);
            qx.push_back(I.native_xyz[3 * i]);
            qy.push_back(I.native_xyz[3 * i + 1]);
            qz.push_back(I.native_xyz[3 * i + 2]);
            v.push_back(vi);
            pv.push_back(pi);
            nv.push_back(ni);
        }
        // compact pred-valid / native-valid rows (align.h:66-71)
        for (int i = 0; i < N; i++)
            if (pv[off + i]) {
                pcx.push_back(px[off + i]); pcy.push_back(py[off + i]); pcz.push_back(pz[off + i]);
                pmap.push_back(i);
                pl++;
            }
        for (int i = pl; i < N; i++) {
            pcx.push_back(0); pcy.push_back(0); pcz.push_back(0); pmap.push_back(0);
        }
        for (int i = 0; i < N; i++)
            if (nv[off + i]) {
                qcx.push_back(qx[off + i]); qcy.push_back(qy[off + i]); qcz.push_back(qz[off + i]);
                nmap.push_back(i);
                nl++;
            }
        for (int i = nl; i < N; i++) {
            qcx.push_back(0); qcy.push_back(0); qcz.push_back(0); nmap.push_back(0);
        }
        h.V = V;
        h.plen = pl;
        h.nlen = nl;
        h.cross = cross;
        h.info.plen = pl;
        h.info.nlen = nl;
        h.info.V = V;
        h.pre = std::move(pre);
        info.push_back(h.info);
        hp_.push_back(std::move(h));
        off += N;
    }
    A_.px = upload(b_px, px); A_.py = upload(b_py, py); A_.pz = upload(b_pz, pz);
    A_.qx = upload(b_qx, qx); A_.qy = upload(b_qy, qy); A_.qz = upload(b_qz, qz);
    A_.v = upload(b_v, v); A_.pv = upload(b_pv, pv); A_.nv = upload(b_nv, nv);
    A_.pcx = upload(b_pcx, pcx); A_.pcy = upload(b_pcy, pcy); A_.pcz = upload(b_pcz, pcz);
    A_.qcx = upload(b_qcx, qcx); A_.qcy = upload(b_qcy, qcy); A_.qcz = upload(b_qcz, qcz);
    A_.pmap = upload(b_pmap, pmap); A_.nmap = upload(b_nmap, nmap);
    {   // 32-byte AoS copies (one wide load per point in the bank kernel)
        std::vector<P4> p4(px.size()), q4(px.size()), pc4(px.size()), qc4(px.size());
        for (size_t i = 0; i < px.size(); i++) {
            p4[i] = P4{px[i], py[i], pz[i], v[i] ? 1.0 : 0.0};
            q4[i] = P4{qx[i], qy[i], qz[i], 0.0};
            pc4[i] = P4{pcx[i], pcy[i], pcz[i], 0.0};
            qc4[i] = P4{qcx[i], qcy[i], qcz[i], 0.0};
        }
        std::vector<P8> pq8(px.size());
        for (size_t i = 0; i < px.size(); i++) pq8[i] = P8{p4[i], q4[i]};
        A_.pq8 = upload(b_pq8, pq8);
        A_.p4 = upload(b_p4, p4);
        A_.q4 = upload(b_q4, q4);
        A_.pc4 = upload(b_pc4, pc4);
        A_.qc4 = upload(b_qc4, qc4);
    }
    {   // FP32 copies of the compact native coordinates (fast method)
        std::vector<float> fx(qcx.begin(), qcx.end()), fy(qcy.begin(), qcy.end()),
            fz(qcz.begin(), qcz.end());
        A_.qfx = upload(b_qfx, fx);
        A_.qfy = upload(b_qfy, fy);
        A_.qfz = upload(b_qfz, fz);
        IA_CHECK(hipStreamSynchronize(s_));
    }
    n_res_ = off;
    A_.ssp = b_ssp.get<char>(std::max(off, 1));
    A_.ssn = b_ssn.get<char>(std::max(off, 1));
    A_.info = upload(b_info, info);
    A_.bank_work = b_bank_work.get<unsigned long long>(3);
    A_.chain_work = A_.bank_work + 2;
    IA_CHECK(hipStreamSynchronize(s_));
}

// Global + window seeds over index positions with >= 3 valid residues
// (seeds.h:41-82); depends on use_fragment_search, so done per search.
void Engine::enumerate_seeds() {
    for (HostPair &h : hp_) {
        h.frag.clear();
        const int N = h.N, V = h.V;
        if (V >= 3) h.frag.push_back(SeedDesc{0, N});
        if (!cfg_->use_fragment_search || V < 4) continue;
        int lens[9];
        const int nlens = window_lengths(V, lens);
        for (int L = 0; L < nlens; L++) {
            const int cnt = window_start_count(N, lens[L], 50);
            for (int q = 0; q < cnt; q++) {
                const int s0 = window_start(N, lens[L], 50, q);
                const int e = std::min(N, s0 + lens[L]);
                if (h.pre[e] - h.pre[s0] >= 3) h.frag.push_back(SeedDesc{s0, lens[L]});
            }
        }
    }
}

std::vector<int> Engine::max_seed_table(int nmax, int ms) {
    std::vector<int> t(nmax + 1, 0);
    int m = 0;
    for (int n = 0; n <= nmax; n++) {
        m = std::max(m, implicit_seed_count(n, ms));
        t[n] = m;
    }
    return t;
}

void Engine::screen_seeds(std::vector<BankDesc> &banks, const SeedDesc *seeds_dev, int screen) {
    std::vector<BankDesc> sb(banks);
    std::vector<int> blk_bank, blk_first;
    std::vector<long long> row0(banks.size());
    long long cand = 0;
    for (size_t b = 0; b < sb.size(); b++) {
        BankDesc &x = sb[b];
        x.cand_cap = banks[b].cand_cap / std::max(banks[b].C, 1);  // seeds (upper bound)
        x.C = 1;
        x.iters = 0;
        x.et = 1e300;
        x.sel_m = 0;
        x.cand_off = cand;
        row0[b] = cand;
        cand += x.cand_cap;
        for (int j = 0; j < x.cand_cap; j += 256) {
            blk_bank.push_back((int)b);
            blk_first.push_back(j);
        }
    }
    if (cand == 0) return;
    int kmax = 1;  // a budget above every bank's seed count keeps them all
    for (const BankDesc &x : sb) kmax = std::max(kmax, x.cand_cap);
    screen = std::min(screen, kmax);
    const BankDesc *db = upload(b_sbanks, sb);
    const int *dbb = upload(b_sblk_bank, blk_bank);
    const int *dbf = upload(b_sblk_first, blk_first);
    Hit *dc = b_scands.get<Hit>(cand);
    launch_bank(s_, A_, db, dbb, dbf, (int)blk_bank.size(), seeds_dev, b_masks.get<uint8_t>(1),
                b_list_pq.get<P8>(1), b_list_na.get<int>(1), list_cap_, dc, fast_, 0, nullptr,
                nullptr, nullptr, false);  // screened banks are never PyTorch-engine lists
    launched();
    float *sc = b_pred.get<float>(cand);
    launch_hit_scores(s_, dc, cand, 1, sc);
    int *items = b_sitems.get<int>(sb.size());
    launch_bank_seed_count(s_, db, (int)sb.size(), b_list_na.get<int>(1), items);
    std::vector<long long> sel_off(sb.size());
    for (size_t b = 0; b < sb.size(); b++) sel_off[b] = (long long)b * screen;
    int *isel = b_isel.get<int>(sb.size() * (size_t)screen);
    launch_select_top(s_, sc, upload(b_srow, row0), items, (int)sb.size(), screen, isel,
                      upload(b_ssel, sel_off), b_scnt.get<int>(sb.size()),
                      b_pmax.get<float>(sb.size()));
    launched();
    IA_CHECK(hipGetLastError());
    A_.isel = isel;
    for (size_t b = 0; b < banks.size(); b++) {
        banks[b].sel_m = screen;
        banks[b].sel_off = sel_off[b];
        banks[b].cand_cap = std::min(banks[b].cand_cap, screen * banks[b].C);
    }
}

Hit *Engine::run_banks(std::vector<BankDesc> &banks, const std::vector<SeedDesc> &seeds,
                       size_t n_out, Hit *out_dev, int screen) {
    if (screen > 0 && !banks.empty()) {
        // the list copies screen_seeds reads are gathered below as well
        int max_slot = -1;
        for (const BankDesc &b : banks)
            if (b.pts == PTS_LIST || b.pts == PTS_TLIST) max_slot = std::max(max_slot, b.a0);
        if (max_slot >= 0) {
            P8 *lpq = b_list_pq.get<P8>((size_t)(max_slot + 1) * list_cap_);
            launch_list_gather(s_, A_, upload(b_banks, banks), (int)banks.size(),
                               b_lists_p.get<int>(1), b_lists_n.get<int>(1),
                               b_list_na.get<int>(1), list_cap_, lpq);
            launched();
        }
        screen_seeds(banks, upload(b_seeds, seeds), screen);
    }
    // Blocks go out in decreasing cost of their bank (points x passes; a
    // list's length is bounded by the shorter compact side), stable among
    // equals, so a launch ends on cheap blocks.  Every block writes only its
    // own candidates' slots, so the order changes no result.
    std::vector<int> bank_order(banks.size());
    std::vector<long long> bank_cost(banks.size());
    const bool tlist = !banks.empty() && banks[0].pts == PTS_TLIST;
    long long cand = 0, max_cost = 0;
    for (size_t b = 0; b < banks.size(); b++) {
        banks[b].cand_off = cand;
        cand += banks[b].cand_cap;
        stats[1] += banks[b].cand_cap;
        const HostPair &h = hp_[banks[b].pair];
        const int n = banks[b].pts == PTS_LIST ? std::min(h.plen, h.nlen) : banks[b].n;
        if ((banks[b].pts == PTS_TLIST) != tlist)
            throw DevError("PyTorch-engine list banks run in a launch of their own");
        bank_cost[b] = std::max(0LL, (long long)n * banks[b].iters);
        max_cost = std::max(max_cost, bank_cost[b]);
        bank_order[b] = (int)b;
    }
    if (max_cost < (1 << 20)) {
        // stable counting sort, decreasing cost: the order stable_sort gives
        // (the threading stage's banks number in the hundreds of thousands)
        std::vector<int> at((size_t)max_cost + 1, 0);
        for (long long c : bank_cost) at[(size_t)c]++;
        int p = 0;
        for (long long c = max_cost; c >= 0; c--) {
            const int k = at[(size_t)c];
            at[(size_t)c] = p;
            p += k;
        }
        for (size_t b = 0; b < banks.size(); b++) bank_order[at[(size_t)bank_cost[b]]++] = (int)b;
    } else {
        std::stable_sort(bank_order.begin(), bank_order.end(),
                         [&](int a, int b) { return bank_cost[a] > bank_cost[b]; });
    }
    std::vector<int> blk_bank, blk_first;
    blk_bank.reserve(banks.size());
    blk_first.reserve(banks.size());
    for (int b : bank_order)
        for (int j = 0; j < banks[b].cand_cap; j += 256) {
            blk_bank.push_back(b);
            blk_first.push_back(j);
        }
    if (!out_dev) out_dev = b_bank_out.get<Hit>(n_out);
    if (banks.empty()) return out_dev;
    const BankDesc *db = upload(b_banks, banks);
    const int *dbb = upload(b_blk_bank, blk_bank);
    const int *dbf = upload(b_blk_first, blk_first);
    const SeedDesc *ds = upload(b_seeds, seeds);
    Hit *dc = b_cands.get<Hit>(cand);
    int max_slot = -1;
    for (const BankDesc &b : banks)
        if (b.pts == PTS_LIST || b.pts == PTS_TLIST) max_slot = std::max(max_slot, b.a0);
    P8 *lpq = b_list_pq.get<P8>((size_t)(max_slot + 1) * list_cap_);
    if (max_slot >= 0 && screen <= 0) {
        launch_list_gather(s_, A_, db, (int)banks.size(), b_lists_p.get<int>(1),
                           b_lists_n.get<int>(1), b_list_na.get<int>(1), list_cap_, lpq);
        launched();
    }
    // Candidates still refining after bank_pass_cap passes run on in a second,
    // compacted launch (their bank's lanes would otherwise idle up to B.iters).
    bool cont = false;
    if (bank_pass_cap > 0)
        for (const BankDesc &b : banks) cont = cont || b.iters > bank_pass_cap;
    int *ccnt = nullptr, *cidx = nullptr;
    long long *cst = nullptr;
    if (cont) {
        ccnt = b_cont_cnt.get<int>(banks.size());
        cidx = b_cont_idx.get<int>(cand);
        cst = b_cont_state.get<long long>(cand);  // one (last count, next pass) pair per slot
        IA_CHECK(hipMemsetAsync(ccnt, 0, banks.size() * sizeof(int), s_));
    }
    IA_CHECK(hipEventRecord(e0_, s_));
    launch_bank(s_, A_, db, dbb, dbf, (int)blk_bank.size(), ds, b_masks.get<uint8_t>(1), lpq,
                b_list_na.get<int>(1), list_cap_, dc, fast_, cont ? bank_pass_cap : 0, ccnt, cidx,
                cst, tlist);
    launched();
    IA_CHECK(hipGetLastError());
    if (cont) {
        const std::vector<int> cnt = download(ccnt, banks.size());
        std::vector<int> cb, cf;
        for (int b : bank_order)
            for (int k = 0; k < cnt[b]; k += 256) {
                cb.push_back(b);
                cf.push_back(k);
            }
        if (!cb.empty()) {
            launch_bank_cont(s_, A_, db, upload(b_cblk_bank, cb), upload(b_cblk_first, cf),
                             (int)cb.size(), lpq, b_list_na.get<int>(1), list_cap_, dc, fast_,
                             ccnt, cidx, cst, tlist);
            launched();
            IA_CHECK(hipGetLastError());
        }
    }
    IA_CHECK(hipEventRecord(e1_, s_));
    launch_bank_reduce(s_, db, (int)banks.size(), b_list_na.get<int>(1), dc, out_dev);
    launched();
    IA_CHECK(hipGetLastError());
    IA_CHECK(hipEventSynchronize(e1_));
    float ms = 0;
    IA_CHECK(hipEventElapsedTime(&ms, e0_, e1_));
    stats[6] += ms;
    return out_dev;
}

long long Engine::run_chains(std::vector<ChainGroup> &groups, const Hit *xf_dev, Hit *out_dev) {
    // Length classes: chains are split by their pair's compact native length
    // into up to chain_buckets launches, each with the LDS plan (trace in
    // LDS or a global slot, staged columns) and occupancy of its own longest
    // chain, so one long pair no longer sets the plan of every short chain.
    // The launches run concurrently on their own streams (default priority);
    // each persistent grid pulls from its own queue.  Within a class the
    // queue hands out chains longest first (NW cells x iterations per
    // chain, stable among equals), so a launch ends on its cheapest chains.
    // Each chain writes only its own outputs, so neither the classes nor the
    // order change any result; chain_off is the prefix in queue order.
    static const int kBound[kChainClasses] = {160, 400, 900, 1 << 30};
    const int ncls = std::max(1, std::min(chain_buckets, kChainClasses));
    auto cls_of = [&](const ChainGroup &g) {
        if (ncls == 1) return 0;
        const int n = g.smat == SM_TORCH ? hp_[g.pair].N : hp_[g.pair].nlen;
        int c = 0;
        while (c < kChainClasses - 1 && n > kBound[c]) c++;
        return std::min(c, ncls - 1);
    };
    // a chain's matrix: compact plen x nlen, or N x N for SM_TORCH chains
    auto rows_of = [&](const ChainGroup &g) { return g.smat == SM_TORCH ? hp_[g.pair].N : hp_[g.pair].plen; };
    auto cols_of = [&](const ChainGroup &g) { return g.smat == SM_TORCH ? hp_[g.pair].N : hp_[g.pair].nlen; };
    auto chain_cost = [&](const ChainGroup &g) {
        return (double)rows_of(g) * cols_of(g) * g.iters;
    };
    // SM_TORCH chains run in a kernel instantiation of their own
    const bool torch = !groups.empty() && groups[0].smat == SM_TORCH;
    for (const ChainGroup &g : groups)
        if ((g.smat == SM_TORCH) != torch)
            throw DevError("PyTorch-engine DP chains run in a launch of their own");
    std::vector<int> order(groups.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const int ca = cls_of(groups[a]), cb = cls_of(groups[b]);
        if (ca != cb) return ca < cb;
        return chain_cost(groups[a]) > chain_cost(groups[b]);
    });
    struct Cls {
        long long g0 = 0, ng = 0, total = 0;
        int mp = 1, mn = 1;
        bool in_lds = false, cols_lds = false;
        long long slots = 0, slot_words = 0, trace_off = 0, gq_off = 0;
        size_t lds = 0;
    };
    std::vector<Cls> cl(kChainClasses);
    std::vector<ChainGroup> queued;
    queued.reserve(groups.size());
    long long total_all = 0;
    for (size_t q = 0; q < order.size(); q++) {
        ChainGroup &g = groups[order[q]];
        Cls &c = cl[cls_of(g)];
        if (c.ng == 0) c.g0 = (long long)q;
        c.ng++;
        g.chain_off = c.total;
        queued.push_back(g);
        c.total += g.n_chains;
        total_all += g.n_chains;
        c.mp = std::max(c.mp, rows_of(g));
        c.mn = std::max(c.mn, cols_of(g));
        long long filled = g.n_chains;  // chains that run the NW fill
        if (g.init == CI_OFFSET && g.k0 <= 0 && (-g.k0) % g.kstep == 0 &&
            (-g.k0) / g.kstep < g.n_chains / 2)
            filled -= 2;  // offset k = 0 is skipped (indialign.cpp:236)
        stats[3] += g.n_chains;
        stats[2] += (double)rows_of(g) * cols_of(g) * (double)filled * g.iters;
    }
    if (total_all == 0) return 0;
    // LDS layout per class: the trace and the staged columns each go to LDS
    // only when that costs no resident wavefront (the fill is latency-bound,
    // so occupancy beats the global slots' extra traffic); otherwise the
    // trace lives in a per-wavefront global slot and the columns are read
    // from the arena.
    constexpr size_t kLdsMax = 160 * 1024;
    int active = 0;
    for (const Cls &c : cl) active += c.total > 0 ? 1 : 0;
    const long long budget = (8LL << 30) / std::max(1, active);  // bytes of global trace slots
    long long trace_words_all = 0, gq_all = 0;
    for (Cls &c : cl) {
        if (c.total == 0) continue;
        auto occupancy = [&](bool tr, bool co) {
            const size_t b = chain_lds_bytes(c.mp, c.mn, tr, co);
            return b > kLdsMax ? 0 : chain_blocks_per_cu(b, fast_, chain_waves, torch);
        };
        const int occ_free = occupancy(false, false);
        if (occ_free <= 0) throw DevError("chain too long for one workgroup's LDS");
        const long long trace_w = chain_trace_words(c.mp, c.mn);
        const bool lds_fits = trace_w * 4 <= 96 * 1024 && occupancy(true, false) > 0;
        c.in_lds = chain_trace == 2 ? false
                   : chain_trace == 1 ? lds_fits
                                      : lds_fits && occupancy(true, false) >= occ_free;
        c.cols_lds = occupancy(c.in_lds, true) >= occupancy(c.in_lds, false);
        c.lds = chain_lds_bytes(c.mp, c.mn, c.in_lds, c.cols_lds);
        int per_cu = std::max(1, chain_blocks_per_cu(c.lds, fast_, chain_waves, torch));
        per_cu = std::min(per_cu, 16);
        if (chain_slots_per_cu > 0) per_cu = std::min(per_cu, chain_slots_per_cu);
        c.slots = (long long)n_cu_ * per_cu;
        if (!c.in_lds) {
            c.slot_words = trace_w;
            c.slots = std::max(1LL, std::min(c.slots, budget / (c.slot_words * 4)));
        }
        c.slots = std::min<long long>(c.slots, c.total);
        c.trace_off = trace_words_all;
        if (!c.in_lds) trace_words_all += c.slots * c.slot_words;
        c.gq_off = gq_all;
        gq_all += c.slots * (c.mn + 32);  // aligned-pair records (na <= nlen) + 32 zeros
        static const bool prof = std::getenv("INDIALIGN_PROFILE") != nullptr;
        if (prof)
            std::fprintf(stderr,
                         "[indialign] chain launch: groups=%lld chains=%lld max %dx%d lds=%zu "
                         "trace=%s slots=%lld\n",
                         c.ng, c.total, c.mp, c.mn, c.lds, c.in_lds ? "lds" : "global", c.slots);
    }
    // every buffer is sized before the first launch: the launches of the
    // classes are in flight together
    const ChainGroup *dg = upload(b_groups, queued);
    int *q = b_queue.get<int>(kChainClasses);
    unsigned *trace = trace_words_all ? b_trace.get<unsigned>((size_t)trace_words_all) : nullptr;
    P8 *gq = b_gq.get<P8>((size_t)std::max(gq_all, 1LL));
    ListOut *lout = b_lout.get<ListOut>(1);
    int *lp = b_lists_p.get<int>(1), *ln = b_lists_n.get<int>(1), *lna = b_list_na.get<int>(1);
    IA_CHECK(hipMemsetAsync(q, 0, sizeof(int) * kChainClasses, s_));
    IA_CHECK(hipEventRecord(e0_, s_));
    int k_stream = 0;
    for (int ci = 0; ci < kChainClasses; ci++) {
        const Cls &c = cl[ci];
        if (c.total == 0) continue;
        hipStream_t st = s_;
        if (active > 1) {
            st = cs_[k_stream];
            IA_CHECK(hipStreamWaitEvent(st, e0_, 0));
        }
        launch_chains(st, A_, dg + c.g0, (int)c.ng, c.total, xf_dev, out_dev, lout, lp, ln, lna,
                      list_cap_, c.mp, c.mn, q + ci, c.in_lds ? nullptr : trace + c.trace_off,
                      c.slot_words, (int)c.slots, c.cols_lds, fast_, chain_waves, gq + c.gq_off,
                      (long long)(c.mn + 32), torch);
        launched();
        IA_CHECK(hipGetLastError());
        if (active > 1) {
            IA_CHECK(hipEventRecord(cev_[k_stream], st));
            IA_CHECK(hipStreamWaitEvent(s_, cev_[k_stream], 0));
            k_stream++;
        }
    }
    IA_CHECK(hipEventRecord(e1_, s_));
    IA_CHECK(hipEventSynchronize(e1_));
    float ms = 0;
    IA_CHECK(hipEventElapsedTime(&ms, e0_, e1_));
    stats[5] += ms;
    return total_all;
}

std::vector<Hit> Engine::rescore(const std::vector<int> &pairs, const std::vector<Hit> &in) {
    if (pairs.empty()) return {};
    const int *dp = upload(b_rs_pair, pairs);
    const Hit *di = upload(b_rs_in, in);
    Hit *dout = b_rs_out.get<Hit>(in.size());
    launch_rescore(s_, A_, dp, di, (int)in.size(), dout);
    launched();
    IA_CHECK(hipGetLastError());
    return download(dout, in.size());
}

std::vector<Hit> Engine::group_reduce(const std::vector<long long> &offs,
                                      const std::vector<int> &cnts, const Hit *in_dev) {
    if (offs.empty()) return {};
    const long long *d_off = upload(b_red_off, offs);
    const int *d_cnt = upload(b_red_cnt, cnts);
    Hit *dout = b_red_out.get<Hit>(offs.size());
    launch_group_reduce(s_, d_off, d_cnt, (int)offs.size(), in_dev, dout);
    launched();
    IA_CHECK(hipGetLastError());
    return download(dout, offs.size());
}

std::vector<Hit> Engine::reduce_rescored_groups(const std::vector<long long> &offs,
                                                const std::vector<int> &cnts, const Hit *in_dev,
                                                const std::vector<int> &owner) {
    if (offs.empty()) return {};
    long long n = 0;
    for (size_t k = 0; k < offs.size(); k++) n = std::max(n, offs[k] + cnts[k]);
    std::vector<int> pair_of((size_t)n, owner[0]);
    for (size_t k = 0; k < offs.size(); k++)
        for (int j = 0; j < cnts[k]; j++) pair_of[(size_t)(offs[k] + j)] = owner[k];
    Hit *rs = b_rr_out.get<Hit>((size_t)std::max(n, 1LL));
    if (n > 0) {
        launch_rescore(s_, A_, upload(b_rr_pair, pair_of), in_dev, (int)n, rs, true);
        launched();
        IA_CHECK(hipGetLastError());
    }
    std::vector<Hit> red = group_reduce(offs, cnts, rs);
    std::vector<int> ep;
    std::vector<Hit> eh;
    std::vector<size_t> ek;
    for (size_t k = 0; k < red.size(); k++)
        if (cnts[k] == 0) {
            ep.push_back(owner[k]);
            eh.push_back(red[k]);
            ek.push_back(k);
        }
    if (!ep.empty()) {
        const std::vector<Hit> r = rescore(ep, eh);
        for (size_t i = 0; i < ek.size(); i++) red[ek[i]] = r[i];
    }
    return red;
}

std::vector<Hit> Engine::dp_refine(const std::vector<int> &pairs, const std::vector<Hit> &start,
                                   int iters, bool lite, int screen) {
    const size_t n = pairs.size();
    std::vector<Hit> res(n, identity_hit(0.0));  // align.h:61-64
    std::vector<int> run;                        // entries with plen, nlen >= 3
    for (size_t e = 0; e < n; e++)
        if (hp_[pairs[e]].plen >= 3 && hp_[pairs[e]].nlen >= 3 && iters > 0) run.push_back((int)e);
    if (run.empty()) return res;
    if (lite) {
        std::vector<ChainGroup> groups;
        std::vector<Hit> xf;
        for (int e : run) {
            ChainGroup g{};
            g.pair = pairs[e];
            g.init = CI_XFORM;
            g.mode = CM_LITE;
            g.smat = SM_TM;
            g.n_chains = 2;
            g.iters = iters;
            g.xf_off = (int)xf.size();
            g.xf_stride = 1;
            g.out_off = 2LL * (long long)xf.size();
            g.out_stride = 2;
            g.sd8 = hp_[pairs[e]].info.sd8;
            groups.push_back(g);
            xf.push_back(start[e]);
        }
        const Hit *dxf = upload(b_xf, xf);
        Hit *dout = b_chain_out.get<Hit>(2 * xf.size());
        run_chains(groups, dxf, dout);
        std::vector<long long> offs;
        std::vector<int> cnts;
        for (size_t k = 0; k < xf.size(); k++) {
            offs.push_back(2LL * k);
            cnts.push_back(2);
        }
        const std::vector<Hit> m = group_reduce(offs, cnts, dout);
        for (size_t k = 0; k < run.size(); k++) res[run[k]] = m[k];
        return res;
    }
    // non-lite: per iteration, NW lists for both chains, full detailed search
    // banks over the lists, then the merge of align.h:114-141 on the host.
    const size_t nc = 2 * run.size();
    std::vector<Hit> cur(nc), cb(nc, identity_hit(0.0));
    for (size_t k = 0; k < run.size(); k++) cur[2 * k] = cur[2 * k + 1] = start[run[k]];
    int lmax = 1, nmax = 1;
    for (int e : run) {
        lmax = std::max(lmax, std::min(hp_[pairs[e]].plen, hp_[pairs[e]].nlen));
        nmax = std::max(nmax, hp_[pairs[e]].N);
    }
    const std::vector<int> kmax = max_seed_table(lmax, 50);
    for (int it = 0; it < iters; it++) {
        ensure_lists((int)nc, lmax);
        std::vector<ChainGroup> groups;
        for (size_t k = 0; k < run.size(); k++) {
            ChainGroup g{};
            g.pair = pairs[run[k]];
            g.init = CI_XFORM;
            g.mode = CM_LIST;
            g.smat = SM_TM;
            g.n_chains = 2;
            g.iters = 1;
            g.xf_off = (int)(2 * k);
            g.xf_stride = 2;
            g.xf_per_chain = 1;
            g.out_off = 2LL * (long long)k;
            g.out_stride = 2;
            g.list_slot0 = (int)(2 * k);
            g.sd8 = hp_[g.pair].info.sd8;
            groups.push_back(g);
        }
        const Hit *dxf = upload(b_xf, cur);
        Hit *dout = b_chain_out.get<Hit>(nc);
        run_chains(groups, dxf, dout);
        std::vector<BankDesc> banks;
        for (size_t ch = 0; ch < nc; ch++) {
            const HostPair &h = hp_[pairs[run[ch / 2]]];
            banks.push_back(list_bank(pairs[run[ch / 2]], (int)ch,
                                      kmax[std::min(h.plen, h.nlen)] * 6, h.info.sd8, (long long)ch));
        }
        Hit *ddet = run_banks(banks, {}, nc, nullptr, screen);
        const std::vector<Hit> det = download(ddet, nc);
        const std::vector<ListOut> lo = download(b_lout.get<ListOut>(nc), nc);
        for (size_t ch = 0; ch < nc; ch++) {
            if (lo[ch].na < 3) continue;  // align.h:106
            Hit step;
            std::memcpy(step.R, lo[ch].R, 72);
            std::memcpy(step.t, lo[ch].t, 24);
            step.score = lo[ch].it_score;
            if (det[ch].score > step.score) step = det[ch];
            cur[ch] = step;
            if (step.score > cb[ch].score) cb[ch] = step;
        }
    }
    for (size_t k = 0; k < run.size(); k++)
        res[run[k]] = cb[2 * k + 1].score > cb[2 * k].score ? cb[2 * k + 1] : cb[2 * k];
    return res;
}


std::vector<Hit> Engine::torch_dp_refine(const std::vector<int> &pairs,
                                         const std::vector<Hit> &start, int iters) {
    const size_t n = pairs.size();
    std::vector<Hit> best(start), cur(start);
    for (Hit &h : best) h.score = 0.0;  // best_score starts at 0 (dp.py:212)
    if (n == 0 || iters <= 0) return best;
    int nmax = 1;
    for (int a : pairs) nmax = std::max(nmax, hp_[a].N);
    for (int it = 0; it < iters; it++) {
        // slots 0..n-1: the lists; n..2n-1: their valid counts
        ensure_lists((int)(2 * n), nmax);
        std::vector<ChainGroup> groups;
        for (size_t k = 0; k < n; k++) {
            ChainGroup g{};
            g.pair = pairs[k];
            g.init = CI_XFORM;
            g.mode = CM_LIST;
            g.smat = SM_TORCH;
            g.n_chains = 1;
            g.iters = 1;
            g.xf_off = (int)k;
            g.xf_stride = 1;
            g.out_off = (long long)k;
            g.out_stride = 1;
            g.list_slot0 = (int)k;
            g.nv_slot0 = (int)(n + k);
            g.gap_ss = 0.0;  // no gap penalty (dp.py:31-33)
            g.sd8 = hp_[g.pair].info.sd8;
            groups.push_back(g);
        }
        const Hit *dxf = upload(b_xf, cur);
        Hit *dout = b_chain_out.get<Hit>(n);
        run_chains(groups, dxf, dout);
        std::vector<BankDesc> banks;
        for (size_t k = 0; k < n; k++)
            banks.push_back(torch_list_bank(pairs[k], (int)k, (int)(n + k), (long long)k));
        Hit *ddet = run_banks(banks, {}, n);
        const std::vector<Hit> det = download(ddet, n);
        const std::vector<ListOut> lo = download(b_lout.get<ListOut>(n), n);
        for (size_t k = 0; k < n; k++) {
            // dp.py:221-247: the detailed search replaces the NW superposition
            // when it scores higher; a list with fewer than 3 valid entries
            // changes nothing
            if (lo[k].nvalid < 3) continue;
            Hit step;
            std::memcpy(step.R, lo[k].R, 72);
            std::memcpy(step.t, lo[k].t, 24);
            step.score = lo[k].it_score;
            if (det[k].score > step.score) step = det[k];
            cur[k] = step;
            if (step.score > best[k].score) best[k] = step;
        }
    }
    return best;
}

void Engine::profile(const char *stage, std::chrono::steady_clock::time_point t0) {
    static const char *names[kStages] = {"seed_bank",   "ss",           "gapless_threading",
                                         "local_fragment", "threading_dp", "dense_local_fragment",
                                         "topk_seed",   "dp_refine",    "offset_dp",
                                         "weighted_refine"};
    for (int k = 0; k < kStages; k++)
        if (std::strcmp(stage, names[k]) == 0)
            for (const HostPair &h : hp_)
                stage_scores[(size_t)h.src * kStages + k] = best_[&h - hp_.data()].score;
    static const bool on = std::getenv("INDIALIGN_PROFILE") != nullptr;
    if (!on) return;
    const char *tag = std::getenv("INDIALIGN_PROFILE_TAG");
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "[indialign]%s pairs=%zu stage=%s ms=%.3f\n", tag ? tag : "", hp_.size(),
                 stage, ms);
}

// 1. seeds (seeds.h:41-82 on the host, anchors seeds.h:164-283 on the device)
//    and the main seed bank (indialign.cpp:69-104).
void Engine::stage_seed_bank() {
    auto t0 = std::chrono::steady_clock::now();
    const int n = (int)hp_.size();
    std::vector<int> pairs(n);
    std::vector<SeedDesc> frag;
    std::vector<int> foff(n), fcnt(n);
    std::vector<long long> aoff(n), soff(n);
    long long abytes = 0, sints = 0;
    int max_n = 0;
    for (int a = 0; a < n; a++) {
        pairs[a] = a;
        max_n = std::max(max_n, hp_[a].N);
        foff[a] = (int)frag.size();
        fcnt[a] = (int)hp_[a].frag.size();
        if (fcnt[a] > 990) throw DevError("too many fragment seeds for one pair");
        frag.insert(frag.end(), hp_[a].frag.begin(), hp_[a].frag.end());
        aoff[a] = abytes;
        abytes += 30LL * hp_[a].N;
        soff[a] = sints;
        sints += 4LL * hp_[a].N;
    }
    const int *dpairs = upload(b_pairs, pairs);
    const SeedDesc *dfrag = upload(b_seeds, frag);
    const int *dfo = upload(b_fofs, foff);
    const int *dfc = upload(b_fcnt, fcnt);
    const long long *dao = upload(b_aoff, aoff);
    const long long *dso = upload(b_soff, soff);
    uint8_t *masks = b_masks.get<uint8_t>(abytes);
    int *dac = b_acnt.get<int>(n);
    std::vector<int> acnt(n, 0);
    if (!torch_sem_) {  // the PyTorch engine builds no anchor seeds by default
        launch_anchors(s_, A_, dpairs, n, dfrag, dfo, dfc, masks, dao, dac, b_ascr.get<int>(sints),
                       dso, max_n);
        launched();
        IA_CHECK(hipGetLastError());
        acnt = download(dac, n);
    }
    std::vector<SeedDesc> seeds;
    std::vector<BankDesc> banks;
    for (int a = 0; a < n; a++) {
        const HostPair &h = hp_[a];
        const int K = fcnt[a] + acnt[a];
        if (K == 0) continue;  // indialign.cpp:77-81: score 0, identity
        BankDesc b{};
        b.pair = a;
        b.pts = PTS_FULL;
        b.n = h.N;
        b.seeds = SEED_EXPLICIT;
        b.seed_off = (int)seeds.size();
        b.K = K;
        bank_cutoffs(b, h);
        b.iters = cfg_->max_iter;
        b.init = INIT_SEED;
        b.cand_cap = K * b.C;
        b.out_idx = a;
        const bool serial = serial_ >= 0 ? serial_ != 0
                                          : !(omp_threads_ > 1 && K * h.info.C >= 4 * omp_threads_);
        b.et = serial && !torch_sem_ ? cfg_->early_term_score : kNever;
        banks.push_back(b);
        seeds.insert(seeds.end(), h.frag.begin(), h.frag.end());
        for (int r = 0; r < acnt[a]; r++) seeds.push_back(SeedDesc{(int)(aoff[a] + (long long)r * h.N), -1});
        live_.push_back(a);
    }
    Hit *d = run_banks(banks, seeds, n);
    const std::vector<Hit> res = download(d, n);
    for (int a : live_) best_[a] = res[a];
    for (int a = 0; a < n; a++)
        if (fcnt[a] + acnt[a] == 0) best_[a] = identity_hit(0.0);
    profile("seed_bank", t0);
}

// 2b. secondary-structure seeds IA2 + IA4 (ss.h:87-164), rescored
//     (indialign.cpp:106-116).
void Engine::stage_ss() {
    auto t0 = std::chrono::steady_clock::now();
    const int dp_iter = cfg_->dp_iter;
    std::vector<Hit> ssb(hp_.size(), identity_hit(kNeg));
    std::vector<int> E;
    int lmax = 1;
    for (int a : live_)
        if (hp_[a].plen >= 3 && hp_[a].nlen >= 3) {
            E.push_back(a);
            lmax = std::max(lmax, std::min(hp_[a].plen, hp_[a].nlen));
        }
    if (!E.empty()) {
        launch_ss(s_, A_, (int)hp_.size(), (char *)A_.ssp, (char *)A_.ssn,
                  b_ss_scr.get<int>((size_t)2 * 9 * std::max(n_res_, 1)));
        launched();
        IA_CHECK(hipGetLastError());
        const std::vector<int> kmax = max_seed_table(lmax, 50);
        for (int pass = 0; pass < 2; pass++) {  // 0 = IA2, 1 = IA4
            ensure_lists((int)E.size(), lmax);
            std::vector<ChainGroup> groups;
            std::vector<Hit> xf;
            for (size_t e = 0; e < E.size(); e++) {
                const int a = E[e];
                ChainGroup g{};
                g.pair = a;
                g.init = CI_XFORM;
                g.mode = CM_LIST;
                g.smat = pass == 0 ? SM_SS : SM_IA4;
                g.n_chains = 1;
                g.iters = 1;
                g.xf_off = (int)e;
                g.xf_stride = 1;
                g.out_off = (long long)e;
                g.out_stride = 1;
                g.list_slot0 = (int)e;
                g.sd8 = hp_[a].info.sd8;
                g.gap_ss = -1.0;  // ss.h:75,144
                groups.push_back(g);
                // IA4 starts from the IA2 result when there is one (ss.h:114-115)
                xf.push_back(pass == 0 ? identity_hit(0.0) : (ssb[a].score > -1e20 ? ssb[a] : best_[a]));
            }
            const Hit *dxf = upload(b_xf, xf);
            run_chains(groups, dxf, b_chain_out.get<Hit>(E.size()));
            std::vector<BankDesc> banks;
            for (size_t e = 0; e < E.size(); e++) {
                const HostPair &h = hp_[E[e]];
                banks.push_back(list_bank(E[e], (int)e, kmax[std::min(h.plen, h.nlen)] * 6, h.info.sd8, (long long)e));
            }
            Hit *dd = run_banks(banks, {}, E.size(), nullptr, learned_ ? top_seeds : 0);
            const std::vector<Hit> det = download(dd, E.size());
            const std::vector<ListOut> lo = download(b_lout.get<ListOut>(E.size()), E.size());
            std::vector<int> dpp;
            std::vector<Hit> dps;
            for (size_t e = 0; e < E.size(); e++) {
                const int a = E[e];
                if (lo[e].na < 3) continue;
                if (det[e].score > ssb[a].score) ssb[a] = det[e];
                if (dp_iter > 0) {
                    dpp.push_back(a);
                    dps.push_back(ssb[a]);
                }
            }
            if (!dpp.empty()) {
                const std::vector<Hit> dp =
                    dp_refine(dpp, dps, dp_iter, false, learned_ ? top_seeds : 0);
                for (size_t k = 0; k < dpp.size(); k++)
                    if (dp[k].score > ssb[dpp[k]].score) ssb[dpp[k]] = dp[k];
            }
        }
    }
    std::vector<Hit> in;
    for (int a : live_) in.push_back(ssb[a]);
    const std::vector<Hit> rs = rescore(live_, in);
    for (size_t k = 0; k < live_.size(); k++) offer(live_[k], rs[k]);
    profile("ss", t0);
}

// 2c. gapless threading (threading.h:41-80), same-index scores.
void Engine::stage_gapless() {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<BankDesc> banks;
    std::vector<SeedDesc> seeds;
    std::vector<int> who;
    for (int a : live_) {
        const HostPair &h = hp_[a];
        const int min_ali = std::max(3, std::min(h.plen, h.nlen) / 2);
        const int kr = h.N - min_ali;
        if (kr <= 0) continue;
        const int stride = std::max(1, (2 * kr + 99) / 100);
        BankDesc b{};
        b.pair = a;
        b.pts = PTS_FULL;
        b.n = h.N;
        b.seeds = SEED_EXPLICIT;
        b.seed_off = (int)seeds.size();
        for (int k = -kr; k <= kr; k += stride)
            if (k != 0) seeds.push_back(SeedDesc{k, -2});
        b.K = (int)seeds.size() - b.seed_off;
        b.C = 1;
        b.cut[0] = h.info.d0s;
        b.iters = 5;
        b.use_d8 = 0;
        b.init = INIT_OFFSET;
        b.min_ali = min_ali;
        b.cand_cap = b.K;
        b.out_idx = (long long)who.size();
        b.et = kNever;
        b.sd8 = h.info.sd8;
        if (b.K == 0) continue;
        banks.push_back(b);
        who.push_back(a);
    }
    if (!banks.empty()) {
        Hit *d = run_banks(banks, seeds, who.size());
        const std::vector<Hit> res = download(d, who.size());
        for (size_t k = 0; k < who.size(); k++) offer(who[k], res[k]);
    }
    profile("gapless_threading", t0);
}

// 3. rescue passes (indialign.cpp:125-152; rescue.h).
void Engine::stage_rescue() {
    const double rs1 = cfg_->rescue_score_1, rs2 = cfg_->rescue_score_2;
    auto local_frag = [&](bool dense, const char *name) {
        auto t0 = std::chrono::steady_clock::now();
        std::vector<int> P;
        for (int a : live_)
            if (best_[a].score < (dense ? rs2 : rs1)) P.push_back(a);
        if (P.empty()) return;
        std::vector<ChainGroup> groups;
        std::vector<FeatGroup> fgs;
        std::vector<long long> offs;
        std::vector<int> cnts;
        long long base = 0;
        for (int a : P) {
            const HostPair &h = hp_[a];
            ChainGroup g{};
            g.pair = a;
            g.init = CI_FRAG;
            g.mode = CM_LITE;
            g.smat = SM_TM;
            g.iters = 2;        // rescue.h:92-94
            g.sd8 = 100.0;
            g.jp = rescue_stride(h.plen, dense);
            g.jn = rescue_stride(h.nlen, dense);
            int items = 0;
            for (int len : rescue_lengths(h.plen, h.nlen, dense)) {
                const int pmax = h.plen - len + 1, nmax = h.nlen - len + 1;
                if (pmax <= 0 || nmax <= 0) continue;
                const int nps = (pmax + g.jp - 1) / g.jp, nns = (nmax + g.jn - 1) / g.jn;
                g.lens[g.n_lens] = len;
                g.items_prefix[g.n_lens] = items;
                g.n_ns[g.n_lens] = nns;
                g.n_lens++;
                items += nps * nns;
            }
            g.items_prefix[g.n_lens] = items;
            g.n_chains = 2 * items;
            g.out_off = base;
            g.out_stride = 2;
            offs.push_back(base);
            cnts.push_back(2 * items);
            if (items > 0) {
                groups.push_back(g);
                FeatGroup f{};
                f.pair = a;
                f.items = items;
                f.n_lens = g.n_lens;
                f.jp = g.jp;
                f.jn = g.jn;
                std::memcpy(f.lens, g.lens, sizeof(f.lens));
                std::memcpy(f.items_prefix, g.items_prefix, sizeof(f.items_prefix));
                std::memcpy(f.n_ns, g.n_ns, sizeof(f.n_ns));
                f.row0 = base / 2;
                f.prior = (float)best_[a].score;
                f.dense = dense ? 1.f : 0.f;
                fgs.push_back(f);
            }
            base += 2LL * items;
        }
        const long long rows = base / 2;
        const bool rank = learned_ && !groups.empty();
        float *X = nullptr;
        if ((rank || rec) && rows > 0) {
            X = b_feat.get<float>((size_t)rows * kFeat);
            const FeatGroup *fdev = upload(b_fgroups, fgs);
            std::vector<int> bg, bf;
            for (size_t k = 0; k < fgs.size(); k++)
                for (int f = 0; f < fgs[k].items; f += kFeatItemsPerBlock) {
                    bg.push_back((int)k);
                    bf.push_back(f);
                }
            launch_frag_features(s_, A_, fdev, upload(b_fblk_g, bg), upload(b_fblk_f, bf),
                                 (int)bg.size(), X);
            launched();
        }
        if (rank) {
            // refine only the ranker's top-m work items of every pair; a pair
            // whose predicted best stays below this pass's rescue threshold
            // (where the plain search would be the fallback) gets the wider
            // fallback budget
            const int top_m = dense ? this->top_m : top_m_first;
            const int fb_m = dense ? fallback_top_m : fallback_top_m_first;
            const int m_cap = std::max(top_m, fb_m);
            float *pred = b_pred.get<float>(rows);
            ranker.forward(s_, X, rows, pred);
            std::vector<long long> row0(fgs.size()), sel_off(fgs.size());
            std::vector<int> items(fgs.size());
            long long ns = 0;
            for (size_t k = 0; k < fgs.size(); k++) {
                row0[k] = fgs[k].row0;
                items[k] = fgs[k].items;
                sel_off[k] = ns;
                ns += std::min(fgs[k].items, m_cap);
            }
            int *isel = b_isel.get<int>(ns);
            float *pmax_d = b_pmax.get<float>(fgs.size());
            int *cnt_d = b_scnt.get<int>(fgs.size());
            launch_select_top(s_, pred, upload(b_srow, row0), upload(b_sitems, items),
                              (int)fgs.size(), top_m, isel, upload(b_ssel, sel_off), cnt_d, pmax_d,
                              fb_m, (float)(dense ? rs2 : rs1));
            launched();
            const std::vector<int> sel_cnt = download(cnt_d, fgs.size());
            // margin fallback: a pair whose best predicted refined score
            // clears its running best by the margin refines every work item
            std::vector<float> pmax;
            if (fallback_margin >= 0) pmax = download(pmax_d, fgs.size());
            A_.isel = isel;
            offs.clear();
            cnts.clear();
            size_t k = 0;
            long long ob = 0;
            for (int a : P) {
                if (k < groups.size() && groups[k].pair == a) {
                    ChainGroup &g = groups[k];
                    const bool full = fallback_margin >= 0 &&
                                      (double)pmax[k] >= best_[a].score + fallback_margin;
                    const int c = full ? fgs[k].items : sel_cnt[k];
                    g.sel = full ? 0 : 1;
                    g.sel_off = sel_off[k];
                    g.n_chains = 2 * c;
                    g.out_off = ob;
                    offs.push_back(ob);
                    cnts.push_back(2 * c);
                    ob += 2LL * c;
                    stats[7] += full ? 1 : 0;
                    stats[8] += (!full && c > std::min(fgs[k].items, top_m)) ? 1 : 0;
                    k++;
                } else {
                    offs.push_back(ob);
                    cnts.push_back(0);
                }
            }
            base = ob;
        }
        Hit *dout = b_chain_out.get<Hit>(base);
        run_chains(groups, nullptr, dout);
        if (rec && !rank && rows > 0) {
            float *y = b_itemy.get<float>(rows);
            launch_item_best(s_, dout, rows, y);
            const long long take = std::min(rows, rec->cap - rec->n);
            if (take > 0) {
                IA_CHECK(hipMemcpyAsync(rec->X + rec->n * kFeat, X, sizeof(float) * take * kFeat,
                                        hipMemcpyDeviceToHost, s_));
                IA_CHECK(hipMemcpyAsync(rec->y + rec->n, y, sizeof(float) * take,
                                        hipMemcpyDeviceToHost, s_));
                IA_CHECK(hipStreamSynchronize(s_));
                long long r = rec->n;
                for (const FeatGroup &f : fgs)
                    for (int i = 0; i < f.items && r < rec->n + take; i++) rec->pair[r++] = hp_[f.pair].src;
                rec->n += take;
            }
        }
        const std::vector<Hit> rs = learned_ && reduce_rescored
                                        ? reduce_rescored_groups(offs, cnts, dout, P)
                                        : rescore(P, group_reduce(offs, cnts, dout));
        for (size_t k = 0; k < P.size(); k++) offer(P[k], rs[k]);
        profile(name, t0);
    };
    local_frag(false, "local_fragment");

    {   // threading DP (rescue.h:121-190)
        auto t0 = std::chrono::steady_clock::now();
        std::vector<int> P;
        for (int a : live_)
            if (best_[a].score < rs1) P.push_back(a);
        if (!P.empty()) {
            std::vector<ChainGroup> groups;
            std::vector<long long> offs;
            std::vector<int> cnts;
            std::vector<OffsetItem> items;  // every offset window, pair by pair
            std::vector<long long> slot;    // its result slot (3 per offset)
            items.reserve(P.size() * 160);  // about 2 (plen + nlen) / (min(plen, nlen) / 50) each
            slot.reserve(P.size() * 160);
            long long base = 0;
            static const double mult[6] = {2.0, 1.5, 1.0, 0.75, 0.5, 0.25};
            for (int a : P) {
                const HostPair &h = hp_[a];
                const int pl = h.plen, nl = h.nlen;
                int no = 0;
                if (std::min(pl, nl) >= 3) {
                    const int step = std::max(1, std::min(pl, nl) / 50);
                    for (int off = -(nl - 3); off < pl - 2; off += step) {
                        const int ps = std::max(off, 0), ns = std::max(-off, 0);
                        const int ov = std::min(pl - ps, nl - ns);
                        if (ov < 3) continue;
                        items.push_back(OffsetItem{a, ps, ns, ov});
                        slot.push_back(base + 3LL * no);
                        no++;
                    }
                }
                if (no > 0) {
                    ChainGroup g{};
                    g.pair = a;
                    g.init = CI_XFORM;
                    g.mode = CM_LITE;
                    g.smat = SM_TM;
                    g.iters = 1;
                    g.sd8 = 100.0;
                    g.n_chains = 2 * no;
                    g.xf_off = (int)base;
                    g.xf_stride = 3;
                    g.out_off = base + 1;
                    g.out_stride = 3;
                    groups.push_back(g);
                }
                offs.push_back(base);
                cnts.push_back(3 * no);
                base += 3LL * no;
            }
            Hit *dthr = b_chain_out.get<Hit>(base);
            // learned method: only the top_offsets offsets of each pair by
            // their bank score start NW chains; the other chain slots keep
            // "no result"
            const bool prune = learned_ && top_offsets > 0 && !groups.empty();
            std::vector<char> keep(items.size(), 1);
            if (prune && top_offset_banks > 0) {
                // and only the top_offset_banks offsets of each pair by the
                // TM-score of their window's own superposition run the bank
                float *sc = b_pred.get<float>(items.size());
                launch_offset_screen(s_, A_, upload(b_oitems, items), (int)items.size(), sc);
                launched();
                const std::vector<float> score = download(sc, items.size());
                std::vector<int> ord;
                size_t b0 = 0;
                while (b0 < items.size()) {
                    size_t b1 = b0;
                    while (b1 < items.size() && items[b1].pair == items[b0].pair) b1++;
                    if ((int)(b1 - b0) > top_offset_banks) {
                        ord.resize(b1 - b0);
                        for (size_t k = 0; k < ord.size(); k++) ord[k] = (int)k;
                        std::nth_element(ord.begin(), ord.begin() + top_offset_banks, ord.end(),
                                         [&](int x, int y) {
                                             const float sx = score[b0 + x], sy = score[b0 + y];
                                             return sx > sy || (sx == sy && x < y);
                                         });
                        for (size_t k = b0; k < b1; k++) keep[k] = 0;
                        for (int k = 0; k < top_offset_banks; k++) keep[b0 + ord[k]] = 1;
                    }
                    b0 = b1;
                }
            }
            std::vector<BankDesc> banks;
            banks.reserve(items.size());
            std::vector<int> seeds_of;  // implicit_seed_count(ov, 5) by overlap length
            for (size_t k = 0; k < items.size(); k++) {
                if (!keep[k]) continue;
                const OffsetItem &it = items[k];
                BankDesc b{};
                b.pair = it.pair;
                b.pts = PTS_CWIN;
                b.n = it.ov;
                b.a0 = it.ps;
                b.a1 = it.ns;
                b.seeds = SEED_IMPLICIT;
                b.max_starts = 5;  // rescue.h:159-161
                b.C = 6;
                for (int c = 0; c < 6; c++) b.cut[c] = hp_[it.pair].info.d0s * mult[c];
                b.iters = 8;
                b.use_d8 = 1;
                b.sd8 = 100.0;
                b.init = INIT_SEED;
                b.et = 1.0;
                if ((size_t)it.ov >= seeds_of.size()) seeds_of.resize((size_t)it.ov + 1, -1);
                if (seeds_of[it.ov] < 0) seeds_of[it.ov] = implicit_seed_count(it.ov, 5);
                b.cand_cap = seeds_of[it.ov] * 6;
                b.out_idx = slot[k];
                banks.push_back(b);
            }
            if (prune) {
                launch_fill_none(s_, dthr, base);
                launched();
            }
            run_banks(banks, {}, base, dthr, learned_ ? top_seeds : 0);
            if (prune) {
                const long long rows = base / 3;
                float *sc = b_pred.get<float>(rows);
                launch_hit_scores(s_, dthr, rows, 3, sc);
                std::vector<long long> row0(groups.size()), sel_off(groups.size());
                std::vector<int> items(groups.size());
                long long ns = 0;
                for (size_t k = 0; k < groups.size(); k++) {
                    row0[k] = groups[k].xf_off / 3;
                    items[k] = groups[k].n_chains / 2;
                    sel_off[k] = ns;
                    ns += std::min(items[k], top_offsets);
                }
                int *isel = b_isel.get<int>(ns);
                launch_select_top(s_, sc, upload(b_srow, row0), upload(b_sitems, items),
                                  (int)groups.size(), top_offsets, isel, upload(b_ssel, sel_off),
                                  b_scnt.get<int>(groups.size()), b_pmax.get<float>(groups.size()));
                launched();
                A_.isel = isel;
                for (size_t k = 0; k < groups.size(); k++) {
                    groups[k].sel = 1;
                    groups[k].sel_off = sel_off[k];
                    groups[k].n_chains = 2 * std::min(items[k], top_offsets);
                }
            }
            run_chains(groups, dthr, dthr);
            const std::vector<Hit> rs = learned_ && reduce_rescored
                                            ? reduce_rescored_groups(offs, cnts, dthr, P)
                                            : rescore(P, group_reduce(offs, cnts, dthr));
            for (size_t k = 0; k < P.size(); k++) offer(P[k], rs[k]);
            profile("threading_dp", t0);
        }
    }
    local_frag(true, "dense_local_fragment");
}

// 4. top-k distance seed (indialign.cpp:154-177).
void Engine::stage_topk() {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int> P;
    std::vector<Hit> bst;
    std::vector<long long> moff;
    long long m = 0;
    for (int a : live_)
        if (hp_[a].V >= 3) {
            P.push_back(a);
            bst.push_back(best_[a]);
            moff.push_back(m);
            m += hp_[a].N;
        }
    if (!P.empty()) {
        const int *dp = upload(b_pairs, P);
        const Hit *db = upload(b_xf, bst);
        const long long *dm = upload(b_moff, moff);
        uint8_t *masks = b_masks.get<uint8_t>(m);
        launch_topk(s_, A_, dp, db, (int)P.size(), dm, masks);
        launched();
        IA_CHECK(hipGetLastError());
        std::vector<BankDesc> banks;
        std::vector<SeedDesc> seeds;
        for (size_t k = 0; k < P.size(); k++) {
            const HostPair &h = hp_[P[k]];
            BankDesc b{};
            b.pair = P[k];
            b.pts = PTS_FULL;
            b.n = h.N;
            b.seeds = SEED_EXPLICIT;
            b.seed_off = (int)seeds.size();
            seeds.push_back(SeedDesc{(int)moff[k], -1});
            b.K = 1;
            bank_cutoffs(b, h);
            b.iters = cfg_->max_iter;
            b.init = INIT_SEED;
            b.cand_cap = b.C;
            b.out_idx = (long long)k;
            b.et = torch_sem_ ? kNever : cfg_->early_term_score;
            banks.push_back(b);
        }
        Hit *d = run_banks(banks, seeds, P.size());
        const std::vector<Hit> res = download(d, P.size());
        for (size_t k = 0; k < P.size(); k++) offer(P[k], res[k]);
    }
    profile("topk_seed", t0);
}

// 5. DP refinement and 5b. raw offset starts (indialign.cpp:179-256).
void Engine::stage_dp() {
    auto t0 = std::chrono::steady_clock::now();
    const int dp_iter = cfg_->dp_iter;
    if (dp_iter <= 0) {
        std::vector<Hit> in;
        for (int a : live_) in.push_back(best_[a]);
        const std::vector<Hit> rs = rescore(live_, in);
        for (size_t k = 0; k < live_.size(); k++) offer(live_[k], rs[k]);
        profile("dp_refine", t0);
        return;
    }
    std::vector<double> seed_score(hp_.size());
    std::vector<Hit> start;
    for (int a : live_) {
        seed_score[a] = best_[a].score;
        start.push_back(best_[a]);
    }
    const std::vector<Hit> dp = dp_refine(live_, start, dp_iter, false);
    const std::vector<Hit> dp_id = rescore(live_, dp);
    for (size_t k = 0; k < live_.size(); k++) offer(live_[k], dp_id[k]);
    {
        std::vector<Hit> cur;
        for (int a : live_) cur.push_back(best_[a]);
        const std::vector<Hit> orig = rescore(live_, cur);
        for (size_t k = 0; k < live_.size(); k++) offer(live_[k], orig[k]);
    }
    std::vector<int> X;
    std::vector<size_t> xk;
    for (size_t k = 0; k < live_.size(); k++)
        if (hp_[live_[k]].cross) {
            X.push_back(live_[k]);
            xk.push_back(k);
        }
    if (!X.empty()) {
        const std::vector<Hit> idp = dp_refine(X, std::vector<Hit>(X.size(), identity_hit(0.0)), dp_iter, true);
        const std::vector<Hit> rs = rescore(X, idp);
        for (size_t k = 0; k < X.size(); k++) offer(X[k], rs[k]);
        const int extra = std::max(0, 5 - dp_iter);
        if (extra > 0) {
            std::vector<int> Y;
            std::vector<Hit> from;
            for (size_t k = 0; k < X.size(); k++) {
                const int a = X[k];
                if (!(best_[a].score < 0.36)) continue;
                Y.push_back(a);
                const size_t lk = xk[k];
                from.push_back(seed_score[a] > dp_id[lk].score ? best_[a] : dp[lk]);
            }
            if (!Y.empty()) {
                const std::vector<Hit> ex = dp_refine(Y, from, extra, true);
                const std::vector<Hit> rs2 = rescore(Y, ex);
                for (size_t k = 0; k < Y.size(); k++) offer(Y[k], rs2[k]);
            }
        }
    }
    profile("dp_refine", t0);
    t0 = std::chrono::steady_clock::now();
    // offsets k = -N/2 .. N/2 step max(1, N/10), k != 0, >= 3 offset pairs
    std::vector<ChainGroup> groups;
    std::vector<long long> offs;
    std::vector<int> cnts, owner;
    long long base = 0;
    for (int a : live_) {
        const HostPair &h = hp_[a];
        if (h.plen < 3 || h.nlen < 3) continue;
        const int kstep = std::max(1, h.N / 10), k0 = -h.N / 2;
        int count = 0;
        for (int k = k0; k <= h.N / 2; k += kstep) count++;
        ChainGroup g{};
        g.pair = a;
        g.init = CI_OFFSET;
        g.mode = CM_LITE;
        g.smat = SM_TM;
        g.iters = dp_iter;
        g.sd8 = h.info.sd8;
        g.k0 = k0;
        g.kstep = kstep;
        g.n_chains = 2 * count;
        g.out_off = base;
        g.out_stride = 2;
        groups.push_back(g);
        for (int o = 0; o < count; o++) {
            offs.push_back(base + 2LL * o);
            cnts.push_back(2);
            owner.push_back(a);
        }
        base += 2LL * count;
    }
    if (!groups.empty()) {
        Hit *dout = b_chain_out.get<Hit>(base);
        run_chains(groups, nullptr, dout);
        const std::vector<Hit> red = group_reduce(offs, cnts, dout);
        const std::vector<Hit> rs = rescore(owner, red);
        for (size_t k = 0; k < owner.size(); k++)
            if (rs[k].score != -INFINITY) offer(owner[k], rs[k]);  // skipped offsets
    }
    profile("offset_dp", t0);
}

// 6. weighted TM refinement (indialign.cpp:258-264).
void Engine::stage_weighted() {
    auto t0 = std::chrono::steady_clock::now();
    if (live_.empty()) return;
    std::vector<Hit> in;
    for (int a : live_) in.push_back(best_[a]);
    const int *dp = upload(b_pairs, live_);
    const Hit *di = upload(b_xf, in);
    Hit *dout = b_rs_out.get<Hit>(in.size());
    launch_weighted(s_, A_, dp, di, (int)in.size(), dout);
    launched();
    IA_CHECK(hipGetLastError());
    const std::vector<Hit> w = download(dout, in.size());
    for (size_t k = 0; k < live_.size(); k++) offer(live_[k], w[k]);
    profile("weighted_refine", t0);
}

// The PyTorch engine's final score (search.py:690-757): without DP, the best
// of the TM-score without the d8 cut and the ranking score; with dp_iter >
// 0, DP refinement from the best seed, then the best of its score, the
// TM-scores of its and of the seed's transforms, and the ranking score.
// The reported transform is the first one that reaches the reported score.
void Engine::torch_final() {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<Hit> seed;
    for (int a : live_) seed.push_back(best_[a]);
    const std::vector<Hit> orig = rescore(live_, seed);
    const int dp_iter = cfg_->dp_iter;
    if (dp_iter <= 0) {
        for (size_t k = 0; k < live_.size(); k++)
            if (orig[k].score > best_[live_[k]].score) best_[live_[k]].score = orig[k].score;
        profile("dp_refine", t0);
        return;
    }
    const std::vector<Hit> dp = torch_dp_refine(live_, seed, dp_iter);
    const std::vector<Hit> dp_id = rescore(live_, dp);
    for (size_t k = 0; k < live_.size(); k++) {
        const Hit *cand[4] = {&dp[k], &dp_id[k], &orig[k], &seed[k]};
        const Hit *pick = cand[0];
        for (int q = 1; q < 4; q++)
            if (cand[q]->score > pick->score) pick = cand[q];
        best_[live_[k]] = *pick;
    }
    profile("dp_refine", t0);
}

void Engine::load(const NativePairInput *inp, int n) {
    n_loaded_ = n;
    status_.assign(n, 0);
    for (int k = 0; k < n; k++) {
        if (inp[k].length <= 0) {
            status_[k] = -1;
        } else if (inp[k].lnorm <= 2.0) {
            status_[k] = 1;
        } else {
            int nl = 0;
            for (int i = 0; i < inp[k].length; i++) nl += inp[k].native_valid_mask[i] ? 1 : 0;
            if (nl > chain_max_nlen()) status_[k] = 2;
        }
    }
    prepare(inp, n);
}

int Engine::search(const NativeConfig *cfg, int serial_bank, NativePairResult *out,
                   int method, int omp_threads) {
    const auto t0 = std::chrono::steady_clock::now();
    for (double &x : stats) x = 0;
    cfg_ = cfg;
    serial_ = serial_bank;
    omp_threads_ = omp_threads;
    fast_ = method == 1 || method == 2;
    learned_ = method == 2;
    torch_sem_ = method == 3;
    if (torch_sem_) {  // the engine's cross-index rescue (separate masks) is not rebuilt
        for (const HostPair &h : hp_)
            if (h.cross) return kUnsupported;
    }
    if (learned_ && !ranker.loaded())
        throw DevError("learned method requested but no ranker is loaded (indialign_set_ranker)");
    stage_scores.assign((size_t)n_loaded_ * kStages, std::nan(""));
    live_.clear();
    best_.assign(hp_.size(), identity_hit(kNeg));
    // Lnorm <= 2 -> score 0, identity (indialign.cpp:47-51); a pair too
    // long for the chain kernel -> score NaN, identity, and the call returns
    // kTooLong after searching the others
    int rc = 0;
    for (int k = 0; k < n_loaded_; k++) {
        if (status_[k] != 1 && status_[k] != 2) continue;
        out[k].score = status_[k] == 1 ? 0.0 : std::nan("");
        set_identity(out[k].rotation, out[k].translation);
        if (status_[k] == 2) rc = kTooLong;
    }
    if (!hp_.empty()) {
        IA_CHECK(hipMemsetAsync(A_.bank_work, 0, 3 * sizeof(unsigned long long), s_));
        enumerate_seeds();
        stage_seed_bank();
        if (torch_sem_) {
            stage_topk();
            torch_final();
            goto done;
        }
        stage_ss();
        stage_gapless();
        stage_rescue();
        stage_topk();
        stage_dp();
        stage_weighted();
    done:
        const std::vector<unsigned long long> bw = download(A_.bank_work, 3);
        stats[9] = (double)bw[1];
        stats[10] = (double)bw[0];
        stats[11] = (double)bw[2];
        for (size_t a = 0; a < hp_.size(); a++) {
            NativePairResult &o = out[hp_[a].src];
            o.score = best_[a].score;
            std::memcpy(o.rotation, best_[a].R, 72);
            std::memcpy(o.translation, best_[a].t, 24);
        }
    }
    stats[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

}  // namespace ia

// ---- C ABI ----------------------------------------------------------------
namespace {
std::mutex g_mu;
std::unique_ptr<ia::Engine> g_engine;
double g_stats[ia::kNStats];
// Ranker parameters installed by indialign_set_ranker (host copy, so that an
// engine re-created after a device error gets them again).
std::vector<float> g_ranker;
int g_top_m = 64;
int g_top_m_first = 16;
double g_margin = -1.0;
int g_top_offsets = 16;
int g_top_offset_banks = 16;
int g_top_seeds = 0;
unsigned g_ranker_ver = 0;
unsigned g_engine_ranker_ver = 0;

// Launch shape: {chain waves per SIMD, chain slots per CU, chain trace
// placement, seed-bank pass cap} (Engine::chain_*, Engine::bank_pass_cap).  Read once from a tuning file written by the autotuner,
// then settable with indialign_set_launch_tuning.
int g_tune[4] = {3, 0, 0, 5};
int g_chain_buckets = 1;
bool g_tune_loaded = false;

bool tune_valid(int waves, int slots, int trace, int cap) {
    return (waves == 2 || waves == 3) && slots >= 0 && slots <= 64 && trace >= 0 && trace <= 2 &&
           cap >= 0 && cap <= 64;
}

// The tuning file is $INDIALIGN_TUNING, or indialign_tuning.txt beside this
// library; lines "chain_waves 3", "chain_slots_per_cu 0", "chain_trace 0",
// "bank_pass_cap 0" ('#' starts a comment).  A missing file keeps the defaults; a malformed
// one is reported and ignored.
void load_tuning_file() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    std::string path;
    const char *env = std::getenv("INDIALIGN_TUNING");
    if (env && *env) {
        path = env;
    } else {
        Dl_info info;
        if (!dladdr((void *)&load_tuning_file, &info) || !info.dli_fname) return;
        path = info.dli_fname;
        const size_t cut = path.find_last_of('/');
        path = (cut == std::string::npos ? std::string(".") : path.substr(0, cut)) +
               "/indialign_tuning.txt";
    }
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return;
    int t[4] = {g_tune[0], g_tune[1], g_tune[2], g_tune[3]};
    char line[256];
    bool ok = true;
    while (std::fgets(line, sizeof line, f)) {
        char key[64];
        int v = 0;
        if (line[0] == '#' || std::sscanf(line, "%63s", key) != 1) continue;
        if (std::sscanf(line, "%63s %d", key, &v) != 2) { ok = false; break; }
        if (!std::strcmp(key, "chain_waves")) t[0] = v;
        else if (!std::strcmp(key, "chain_slots_per_cu")) t[1] = v;
        else if (!std::strcmp(key, "chain_trace")) t[2] = v;
        else if (!std::strcmp(key, "bank_pass_cap")) t[3] = v;
        else if (!std::strcmp(key, "chain_buckets") && v >= 1 && v <= ia::kChainClasses)
            g_chain_buckets = v;
    }
    std::fclose(f);
    if (ok && tune_valid(t[0], t[1], t[2], t[3])) {
        std::memcpy(g_tune, t, sizeof t);
    } else {
        std::fprintf(stderr, "indialign_amd: ignoring malformed tuning file %s\n", path.c_str());
    }
}

// With g_mu held: the engine, created on first use.  A new engine (also
// after a device error dropped the old one) has no ranker, budgets or launch
// shape yet; sync_engine installs them.
bool g_engine_synced = false;
// The loaded batch is the caller's resident batch (indialign_batch_load) and
// no other call has replaced it since.
bool g_resident = false;
int g_resident_n = 0;
bool g_learned_rescored = false;
int g_fallback_m = 512;
int g_fallback_m_first = 64;

ia::Engine &engine() {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        throw ia::DevError("no HIP device available");
    if (!g_engine) {
        g_engine.reset(new ia::Engine());
        g_engine_synced = false;
    }
    return *g_engine;
}

void drop_engine() {
    g_engine.reset();
    g_engine_synced = false;
    g_resident = false;
}

// With g_mu held and g_engine alive: install the current ranker, budgets and
// launch shape.
void sync_engine() {
    load_tuning_file();
    g_engine->chain_waves = g_tune[0];
    g_engine->chain_slots_per_cu = g_tune[1];
    g_engine->chain_trace = g_tune[2];
    g_engine->bank_pass_cap = g_tune[3];
    g_engine->chain_buckets = g_chain_buckets;
    if (g_engine_synced && g_engine_ranker_ver == g_ranker_ver) return;
    if (g_ranker.empty()) g_engine->ranker.unload();
    else g_engine->ranker.set_params(g_engine->stream(), g_ranker.data());
    g_engine->top_m = g_top_m;
    g_engine->top_m_first = g_top_m_first;
    g_engine->fallback_margin = g_margin;
    g_engine->top_offsets = g_top_offsets;
    g_engine->top_offset_banks = g_top_offset_banks;
    g_engine->top_seeds = g_top_seeds;
    g_engine->reduce_rescored = g_learned_rescored;
    g_engine->fallback_top_m = g_fallback_m;
    g_engine->fallback_top_m_first = g_fallback_m_first;
    g_engine_ranker_ver = g_ranker_ver;
    g_engine_synced = true;
}

// Host threads the reference's OpenMP runtime would use for a call from the
// caller's thread (omp_get_max_threads with default settings): the first
// entry of OMP_NUM_THREADS, else the CPUs of this process's affinity mask.
int host_omp_threads() {
    const char *env = std::getenv("OMP_NUM_THREADS");
    if (env && *env) {
        const int v = std::atoi(env);
        if (v > 0) return v;
    }
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof set, &set) == 0) {
        const int c = CPU_COUNT(&set);
        if (c > 0) return c;
    }
    return 1;
}

// Standalone ranker for training and direct evaluation (own stream).
struct Trainer {
    hipStream_t s = nullptr;
    ia::Mlp mlp;
    ia::DBuf X, t, perm, Xb, tb, y, losses;
    Trainer() {
        int dev = 0;
        IA_CHECK(hipGetDevice(&dev));
        IA_CHECK(hipStreamCreate(&s));
    }
};
std::mutex g_tmu;
std::unique_ptr<Trainer> g_trainer;

Trainer &trainer() {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        throw ia::DevError("no HIP device available");
    if (!g_trainer) g_trainer.reset(new Trainer());
    return *g_trainer;
}

// After a device error: score NaN and the identity for every pair the call
// would have written (length > 0), so a caller that ignores the return code
// (the reference's batch callers do) cannot read stale or uninitialised
// results as scores.
void fail_outputs(const NativePairInput *inp, int n, NativePairResult *out) {
    for (int k = 0; k < n; k++) {
        if (inp && inp[k].length <= 0) continue;
        out[k].score = std::nan("");
        ia::set_identity(out[k].rotation, out[k].translation);
    }
}

int run_search(int n, const NativePairInput *inp, const NativeConfig *cfg, NativePairResult *out,
               int serial, int method = 0, int omp_threads = 1) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (method == 2 && g_ranker.empty()) return -4;
    try {
        ia::Engine &E = engine();
        sync_engine();
        g_resident = false;
        E.load(inp, n);
        const int rc = E.search(cfg, serial, out, method, omp_threads);
        std::memcpy(g_stats, E.stats, sizeof(g_stats));
        return rc;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "indialign_amd: %s\n", e.what());
        drop_engine();
        fail_outputs(inp, n, out);
        return -2;
    }
}
}  // namespace

extern "C" {

void indialign_default_config(NativeConfig *cfg) {
    cfg->backend_mode = 0;
    cfg->use_fragment_search = 1;
    cfg->max_iter = 20;
    cfg->dp_iter = 1;
    cfg->max_mem_gb = 20.0;
    cfg->request_cuda = 0;
    cfg->rescue_score_1 = 0.5;
    cfg->rescue_score_2 = 0.3;
    cfg->early_term_score = 0.99;
}

int indialign_tmscore_search(const NativePairInput *inp, const NativeConfig *cfg,
                             NativePairResult *out) {
    if (inp->length <= 0) return -1;
    return run_search(1, inp, cfg, out, -1, 0, host_omp_threads());
}

int indialign_tmscore_search_batch(int n_pairs, const NativePairInput *inputs,
                                   const NativeConfig *cfg, NativePairResult *outputs) {
    if (n_pairs <= 0) return 0;
    return run_search(n_pairs, inputs, cfg, outputs, 1);
}

int indialign_search_ex(int n_pairs, const NativePairInput *inputs, const NativeConfig *cfg,
                        NativePairResult *outputs, int serial_bank, int method) {
    if (n_pairs <= 0) return 0;
    if (method < 0 || method > 3) return -3;
    return run_search(n_pairs, inputs, cfg, outputs, serial_bank != 0 ? 1 : 0, method);
}

int indialign_batch_load(int n_pairs, const NativePairInput *inputs) {
    if (n_pairs < 0) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    try {
        ia::Engine &E = engine();
        g_resident = false;
        E.load(inputs, n_pairs);
        g_resident = true;
        g_resident_n = n_pairs;
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "indialign_amd: %s\n", e.what());
        drop_engine();
        return -2;
    }
}

int indialign_batch_run(int n_pairs, const NativeConfig *cfg, NativePairResult *outputs,
                        int serial_bank, int method) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (method < 0 || method > 3) return -3;
    if (method == 2 && g_ranker.empty()) return -4;
    if (!g_engine || !g_resident || n_pairs != g_resident_n) return -5;
    try {
        sync_engine();
        const int rc = g_engine->search(cfg, serial_bank != 0 ? 1 : 0, outputs, method);
        std::memcpy(g_stats, g_engine->stats, sizeof(g_stats));
        return rc;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "indialign_amd: %s\n", e.what());
        drop_engine();
        fail_outputs(nullptr, n_pairs, outputs);
        return -2;
    }
}

int indialign_release(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    drop_engine();
    return 0;
}

int indialign_nw_matrix(int n, const double *smat, const int *plen, const int *nlen,
                        const double *gap, int fast, int trace_global, double *H, int8_t *T,
                        int *ap, int *an, int *na) {
    if (n < 0 || fast < 0 || fast > 1 || trace_global < 0 || trace_global > 1) return -1;
    if (n == 0) return 0;
    std::vector<ia::NwProblem> pr(n);
    long long so = 0, ho = 0, ao = 0;
    int mp = 1, mn = 1;
    for (int k = 0; k < n; k++) {
        if (plen[k] < 1 || nlen[k] < 1 || plen[k] > 32767 || nlen[k] > 32767) return -1;
        pr[k] = ia::NwProblem{so, ho, ao, plen[k], nlen[k], gap[k]};
        so += (long long)plen[k] * nlen[k];
        ho += (long long)(plen[k] + 1) * (nlen[k] + 1);
        ao += std::min(plen[k], nlen[k]);
        mp = std::max(mp, plen[k]);
        mn = std::max(mn, nlen[k]);
    }
    if (ia::nw_matrix_lds_bytes(mp, mn, trace_global == 0) > 160 * 1024) return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return -2;
    std::lock_guard<std::mutex> lk(g_tmu);
    void *buf[8] = {};
    hipStream_t st = nullptr;
    int rc = 0;
    try {
        const long long tw = trace_global ? ia::nw_matrix_trace_words(mp, mn) * n : 0;
        const size_t sz[8] = {sizeof(ia::NwProblem) * n, sizeof(double) * so, sizeof(double) * ho,
                              (size_t)ho, sizeof(int) * ao, sizeof(int) * ao, sizeof(int) * n,
                              sizeof(unsigned) * (size_t)tw};
        for (int b = 0; b < 8; b++)
            if (sz[b]) IA_CHECK(hipMalloc(&buf[b], sz[b]));
        IA_CHECK(hipStreamCreate(&st));
        IA_CHECK(hipMemcpyAsync(buf[0], pr.data(), sz[0], hipMemcpyHostToDevice, st));
        IA_CHECK(hipMemcpyAsync(buf[1], smat, sz[1], hipMemcpyHostToDevice, st));
        IA_CHECK(hipMemsetAsync(buf[2], 0, sz[2], st));     // H boundary 0 (nw.h:14-17)
        IA_CHECK(hipMemsetAsync(buf[3], 0xff, sz[3], st));  // trace boundary -1
        ia::launch_nw_matrix(st, (const ia::NwProblem *)buf[0], n, (const double *)buf[1],
                             (double *)buf[2], (signed char *)buf[3], (int *)buf[4], (int *)buf[5],
                             (int *)buf[6], mp, mn, trace_global ? (unsigned *)buf[7] : nullptr,
                             fast != 0);
        IA_CHECK(hipGetLastError());
        IA_CHECK(hipMemcpyAsync(H, buf[2], sz[2], hipMemcpyDeviceToHost, st));
        IA_CHECK(hipMemcpyAsync(T, buf[3], sz[3], hipMemcpyDeviceToHost, st));
        if (ao) {
            IA_CHECK(hipMemcpyAsync(ap, buf[4], sz[4], hipMemcpyDeviceToHost, st));
            IA_CHECK(hipMemcpyAsync(an, buf[5], sz[5], hipMemcpyDeviceToHost, st));
        }
        IA_CHECK(hipMemcpyAsync(na, buf[6], sz[6], hipMemcpyDeviceToHost, st));
        IA_CHECK(hipStreamSynchronize(st));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "indialign_amd: %s\n", e.what());
        rc = -2;
    }
    if (st) (void)hipStreamDestroy(st);
    for (void *b : buf)
        if (b) (void)hipFree(b);
    return rc;
}

int indialign_cuda_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0 ? 1 : 0;
}

int indialign_last_stats(double *out, int n) {
    std::lock_guard<std::mutex> lk(g_mu);
    const int m = n < ia::kNStats ? n : ia::kNStats;
    for (int k = 0; k < m; k++) out[k] = g_stats[k];
    return m;
}

int indialign_rotation_from_covariance(const double *h, int method, double *R) {
    if (method == 0) {
        ia::horn_rotation(h, R);
        return 0;
    }
    if (method == 1) return ia::qcp_rotation(h, R) ? 0 : 1;
    return -1;
}

int indialign_last_stage_scores(double *out, int n_pairs) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_engine || n_pairs <= 0) return 0;
    const int K = ia::Engine::kStages;
    const int m = std::min<long long>(n_pairs, (long long)g_engine->stage_scores.size() / K);
    std::memcpy(out, g_engine->stage_scores.data(), sizeof(double) * (size_t)m * K);
    return m;
}

int indialign_set_ranker(const float *params) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (params) g_ranker.assign(params, params + ia::kNParam);
    else g_ranker.clear();
    g_ranker_ver++;
    return 0;
}

int indialign_set_learned_budgets(int top_m, int top_m_first, int top_offsets,
                                  int top_offset_banks, int top_seeds, double fallback_margin) {
    if (top_m < 1 || top_m_first < 1 || top_offsets < 1 || top_offset_banks < 0 ||
        (top_offset_banks > 0 && top_offset_banks < top_offsets) || top_seeds < 0)
        return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    g_top_offset_banks = top_offset_banks;
    g_top_m = top_m;
    g_top_m_first = top_m_first;
    g_top_offsets = top_offsets;
    g_top_seeds = top_seeds;
    g_margin = fallback_margin;
    g_ranker_ver++;
    return 0;
}

int indialign_set_learned_options(int reduce_rescored, int fallback_top_m,
                                  int fallback_top_m_first) {
    if (reduce_rescored < 0 || reduce_rescored > 1 || fallback_top_m < 0 ||
        fallback_top_m_first < 0)
        return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    g_learned_rescored = reduce_rescored != 0;
    g_fallback_m = fallback_top_m;
    g_fallback_m_first = fallback_top_m_first;
    g_ranker_ver++;
    return 0;
}

int indialign_set_launch_tuning(int chain_waves, int chain_slots_per_cu, int chain_trace,
                                int bank_pass_cap) {
    if (!tune_valid(chain_waves, chain_slots_per_cu, chain_trace, bank_pass_cap)) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    g_tune_loaded = true;  // an explicit setting wins over the file
    g_tune[0] = chain_waves;
    g_tune[1] = chain_slots_per_cu;
    g_tune[2] = chain_trace;
    g_tune[3] = bank_pass_cap;
    return 0;
}

int indialign_set_chain_buckets(int n_classes) {
    if (n_classes < 1 || n_classes > ia::kChainClasses) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    load_tuning_file();
    g_chain_buckets = n_classes;
    return 0;
}

int indialign_get_launch_tuning(int *out4) {
    std::lock_guard<std::mutex> lk(g_mu);
    load_tuning_file();
    for (int k = 0; k < 4; k++) out4[k] = g_tune[k];
    return 0;
}

int indialign_collect_records(int n_pairs, const NativePairInput *inputs, const NativeConfig *cfg,
                              NativePairResult *outputs, float *X, float *y, int *pair_idx,
                              long long cap, long long *n_records) {
    if (n_pairs <= 0 || cap < 0) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    try {
        engine();
        sync_engine();
        g_resident = false;
        ia::Engine::RecordSink sink{X, y, pair_idx, cap, 0};
        g_engine->rec = &sink;
        g_engine->load(inputs, n_pairs);
        int rc = 0;
        try {
            rc = g_engine->search(cfg, true, outputs, 0);
        } catch (...) {
            g_engine->rec = nullptr;
            throw;
        }
        g_engine->rec = nullptr;
        std::memcpy(g_stats, g_engine->stats, sizeof(g_stats));
        *n_records = sink.n;
        return rc;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "indialign_amd: %s\n", e.what());
        drop_engine();
        return -2;
    }
}

int indialign_mlp_forward(const float *params, const float *X, long long n, float *y) {
    if (n < 0) return -1;
    if (n == 0) return 0;
    std::lock_guard<std::mutex> lk(g_tmu);
    try {
        Trainer &T = trainer();
        T.mlp.set_params(T.s, params);
        float *dX = T.X.get<float>((size_t)n * ia::kFeat);
        float *dy = T.y.get<float>((size_t)n);
        IA_CHECK(hipMemcpyAsync(dX, X, sizeof(float) * n * ia::kFeat, hipMemcpyHostToDevice, T.s));
        T.mlp.forward(T.s, dX, n, dy);
        IA_CHECK(hipMemcpyAsync(y, dy, sizeof(float) * n, hipMemcpyDeviceToHost, T.s));
        IA_CHECK(hipStreamSynchronize(T.s));
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "indialign_amd: %s\n", e.what());
        g_trainer.reset();
        return -2;
    }
}

int indialign_mlp_grad(const float *params, const float *X, const float *t, int n, float *grad,
                       float *loss) {
    if (n <= 0 || n > ia::Mlp::kChunk) return -1;
    std::lock_guard<std::mutex> lk(g_tmu);
    try {
        Trainer &T = trainer();
        T.mlp.set_params(T.s, params);
        float *dX = T.X.get<float>((size_t)n * ia::kFeat);
        float *dt = T.t.get<float>((size_t)n);
        IA_CHECK(hipMemcpyAsync(dX, X, sizeof(float) * (size_t)n * ia::kFeat,
                                hipMemcpyHostToDevice, T.s));
        IA_CHECK(hipMemcpyAsync(dt, t, sizeof(float) * n, hipMemcpyHostToDevice, T.s));
        T.mlp.backward(T.s, dX, dt, n);
        IA_CHECK(hipMemcpyAsync(grad, T.mlp.grad(), sizeof(float) * ia::kNParam,
                                hipMemcpyDeviceToHost, T.s));
        IA_CHECK(hipMemcpyAsync(loss, T.mlp.loss_dev(), sizeof(float), hipMemcpyDeviceToHost, T.s));
        IA_CHECK(hipStreamSynchronize(T.s));
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "indialign_amd: %s\n", e.what());
        g_trainer.reset();
        return -2;
    }
}

int indialign_mlp_train(float *params, const float *X, const float *t, long long n, int epochs,
                        int batch, float lr, unsigned seed, float *epoch_loss) {
    if (n <= 0 || epochs < 0 || batch <= 0 || batch > ia::Mlp::kChunk) return -1;
    std::lock_guard<std::mutex> lk(g_tmu);
    try {
        Trainer &T = trainer();
        T.mlp.set_params(T.s, params);
        T.mlp.reset_adam(T.s);
        float *dX = T.X.get<float>((size_t)n * ia::kFeat);
        float *dt = T.t.get<float>((size_t)n);
        IA_CHECK(hipMemcpyAsync(dX, X, sizeof(float) * n * ia::kFeat, hipMemcpyHostToDevice, T.s));
        IA_CHECK(hipMemcpyAsync(dt, t, sizeof(float) * n, hipMemcpyHostToDevice, T.s));
        float *Xb = T.Xb.get<float>((size_t)batch * ia::kFeat);
        float *tb = T.tb.get<float>((size_t)batch);
        const long long nb = (n + batch - 1) / batch;
        float *losses = T.losses.get<float>((size_t)nb);
        std::vector<int> perm(n);
        std::vector<float> lh(nb);
        for (int e = 0; e < epochs; e++) {
            for (long long i = 0; i < n; i++) perm[i] = (int)i;
            std::mt19937 rng(seed + 7919u * (unsigned)e);
            for (long long i = n - 1; i > 0; i--) {  // Fisher-Yates
                const long long j = (long long)(rng() % (unsigned long long)(i + 1));
                std::swap(perm[i], perm[j]);
            }
            int *dperm = T.perm.get<int>((size_t)n);
            IA_CHECK(hipMemcpyAsync(dperm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                                    T.s));
            for (long long b = 0; b < nb; b++) {
                const int m = (int)std::min<long long>(batch, n - b * batch);
                ia::launch_gather_rows(T.s, dX, dt, dperm + b * batch, m, Xb, tb);
                T.mlp.backward(T.s, Xb, tb, m);
                IA_CHECK(hipMemcpyAsync(losses + b, T.mlp.loss_dev(), sizeof(float),
                                        hipMemcpyDeviceToDevice, T.s));
                T.mlp.adam_step(T.s, lr);
            }
            IA_CHECK(hipMemcpyAsync(lh.data(), losses, sizeof(float) * nb, hipMemcpyDeviceToHost,
                                    T.s));
            IA_CHECK(hipStreamSynchronize(T.s));
            double acc = 0;
            for (float v : lh) acc += v;
            if (epoch_loss) epoch_loss[e] = (float)(acc / (double)nb);
        }
        T.mlp.get_params(T.s, params);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "indialign_amd: %s\n", e.what());
        g_trainer.reset();
        return -2;
    }
}

}  // extern "C"
