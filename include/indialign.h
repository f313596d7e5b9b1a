/* indialign_amd — MI355X (gfx950) TM-score superposition search.
 *
 * Drop-in replacement for the reference library's C ABI.  The struct layouts
 * and the four reference entry points are identical to
 * indialign_c/api.h:10-46, so existing ctypes callers (tests/test_cpp_smoke.py,
 * benchmark/fair_benchmark.py:21-46, benchmark/pdb_*.py) load
 * libindialign.so from this repository unchanged.
 *
 * All pointers are borrowed, read-only, caller-owned and must stay valid for
 * the duration of the call.  Coordinates are interleaved AoS double[3N];
 * masks are uint8 with values 0 or 1.  The rotation is applied as
 * out_d = sum_k P_k * rotation[3k+d] + translation[d].
 */
#ifndef INDIALIGN_AMD_H
#define INDIALIGN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* api.h:10-20.  backend_mode and max_mem_gb are accepted and ignored (the
 * reference ignores them too, indialign.cpp:290-297); request_cuda is
 * ignored: the search always runs on the GPU. */
typedef struct {
    int backend_mode;
    int use_fragment_search;
    int max_iter;
    int dp_iter;
    double max_mem_gb;
    int request_cuda;
    double rescue_score_1;
    double rescue_score_2;
    double early_term_score;
} NativeConfig;

/* api.h:22-30 */
typedef struct {
    int length;
    const double *pred_xyz;
    const double *native_xyz;
    const uint8_t *valid_mask;
    const uint8_t *pred_valid_mask;
    const uint8_t *native_valid_mask;
    double lnorm;
} NativePairInput;

/* api.h:32-36 */
typedef struct {
    double score;
    double rotation[9];
    double translation[3];
} NativePairResult;

/* Replaces indialign_default_config (api.h:38, indialign.cpp:272-282):
 * same defaults. */
void indialign_default_config(NativeConfig *cfg);

/* Replaces indialign_tmscore_search (api.h:39-41, indialign.cpp:284-303).
 * Returns -1 when length <= 0 (as the reference), -2 when no usable GPU is
 * present or a device error occurred (message on stderr; out then holds
 * score NaN and the identity; the reference never had this case because it
 * fell back to its CPU path), else 0.
 * The main seed bank follows the reference's per-call switch
 * (search.h:14-24, 83-136): early termination at early_term_score (its
 * serial path) when the reference's OpenMP runtime would run one thread or
 * the bank has fewer than 4 candidates (K seeds x C cutoffs) per thread,
 * otherwise none (its multithreaded path).  The thread count is the one
 * libgomp's omp_get_max_threads() gives a caller with default settings: the
 * first entry of OMP_NUM_THREADS, else the CPUs of the process's affinity
 * mask (a caller that changed it with omp_set_num_threads is not seen).
 * A pair with more than 10239 native-valid residues (the NW chain kernel
 * keeps one FP64 (H, G) pair per compact native residue in a wavefront's
 * 160 KB of LDS) gets score NaN and the identity transform, and the call
 * returns -6 (the reference has no such limit). */
int indialign_tmscore_search(const NativePairInput *inp, const NativeConfig *cfg,
                             NativePairResult *out);

/* Replaces indialign_tmscore_search_batch (api.h:42-45,
 * indialign.cpp:305-313).  All pairs run together in batched launches; the
 * main seed bank follows the reference's batch (serial) semantics, i.e.
 * early_term_score applies (search.h:83).  Returns 0, or -2 on a device
 * error (every pair with length > 0 then holds score NaN and the identity,
 * visible to callers that ignore the return code, as the reference's batch
 * callers do).  Pairs with length <= 0 are left untouched (the reference
 * ignored per-pair return codes).  Pairs over the 10239 native-valid residue limit
 * (see indialign_tmscore_search) get score NaN and the identity while the
 * others are searched; the call then returns -6. */
int indialign_tmscore_search_batch(int n_pairs, const NativePairInput *inputs,
                                   const NativeConfig *cfg, NativePairResult *outputs);

/* Replaces indialign_cuda_available (api.h:46, indialign.cpp:315-321):
 * 1 when a HIP device is present, else 0 (the name is kept for the ABI). */
int indialign_cuda_available(void);

/* ---- extensions (no counterpart in the reference) ---------------------- */

/* Batch search with an explicit seed-bank convention and method.
 * serial_bank = 1 applies early_term_score to the main seed bank (the
 * reference's batch / one-thread semantics), 0 never stops early (its
 * multithreaded single-pair semantics).
 * method = 0: plain FP64, the reference's arithmetic (results identical to
 * the reference CPU path up to its own rounding); method = 1: fast (FP32
 * Needleman-Wunsch fill, closed-form FP64 rotations; every reported score is
 * still an FP64 TM-score of the returned transform); method = 2: learned
 * (method 1, and of the local-fragment rescue work items (rescue.h:60-118)
 * of each pair only the top-m by the ranker installed with
 * indialign_set_ranker are refined; the threading-DP offsets and the seeds
 * of the nested detailed searches are screened the same way, see
 * indialign_set_learned_budgets).  method = 3: the search of the reference's
 * separate PyTorch engine with same-index masks
 * (scoring/_tmscore_gpu/search.py:492-761), each pair as a batch of one:
 * window seeds without anchor seeds, cutoffs d0_search x {2, 1.5, 1, .75,
 * .5, .25}, exactly max_iter refinement passes per candidate (no
 * convergence break), candidates ranked by the TM-score with the score_d8
 * cut, then the top-k distance seed; with dp_iter > 0 the engine's DP
 * refinement (dp.py:186-252) from the best seed; the reported score is the
 * maximum of the DP score, the TM-scores without the cut of the DP and seed
 * transforms and the ranking score, with the first transform that reaches
 * it; FP64.  Its cross-index rescue is not rebuilt: method 3 returns -7
 * when a pair has pred_valid != native_valid.  Same return codes as the
 * batch entry, -3 for an unknown method, -4 for method 2 without a
 * ranker. */
int indialign_search_ex(int n_pairs, const NativePairInput *inputs, const NativeConfig *cfg,
                        NativePairResult *outputs, int serial_bank, int method);

/* Resident batch: copy n_pairs inputs into device memory once
 * (indialign_batch_load), then search them any number of times without
 * host-to-device traffic (indialign_batch_run; n_pairs must be the loaded
 * count and outputs has n_pairs slots; same conventions as
 * indialign_search_ex).  Any other search call (or indialign_release)
 * replaces the resident batch; indialign_batch_run then returns -5 until the
 * next load, as it does for an n_pairs other than the loaded count.
 * Return 0, -1 for n_pairs < 0 (load), -2 on a device error, -3 / -4 as
 * indialign_search_ex. */
int indialign_batch_load(int n_pairs, const NativePairInput *inputs);
int indialign_batch_run(int n_pairs, const NativeConfig *cfg, NativePairResult *outputs,
                        int serial_bank, int method);

/* Free all device memory the library holds (the engine is created again by
 * the next call; the ranker, its budgets and the launch shape are kept).
 * Returns 0. */
int indialign_release(void);

/* Work and time of the last search.  Fills up to n (at most 11)
 * doubles:
 *   [0] wall ms  [1] candidates refined  [2] NW cells filled
 *   [3] chains run  [4] kernel launches
 *   [5] chain-kernel ms  [6] bank-kernel ms  (device time, HIP events)
 *   [7] local-fragment passes that fell back to every work item (method 2,
 *       margin rule)
 *   [8] local-fragment passes that took the fallback budget (method 2)
 *   [9] seed-bank refinement passes (iterations of the refine loop,
 *       search.h:40-55, counted on the device)
 *  [10] seed-bank algorithmic FP64 operations, counted on the device: per
 *       pass 26 n (transform + distance) + 31 n_sel + 3400 (centroids,
 *       covariance, key matrix, Jacobi solve) when n_sel >= 3, the initial
 *       Kabsch 31 n_seed + 3400, and 29 n per TM-score (n the bank's points)
 *  [11] chain-kernel algorithmic FP64 operations outside the NW cells,
 *       counted on the device in the same units: per NW iteration the
 *       Kabsch on the aligned set (31 n + 3400) and its TM-score (29 n);
 *       per lite detailed search and d0 cutoff, 26 n per selection pass,
 *       31 n_sel + 3400 per solved Kabsch and 29 n for the TM-score; the
 *       Kabsch of fragment and offset chain starts (n the set's points);
 *       FP64 methods (0 and 3) only, 0 for methods 1 and 2
 * Returns the number of values written. */
int indialign_last_stats(double *out, int n);

/* Running best score of each pair of the last search after each of the ten
 * pipeline stages (indialign.cpp:69-264): 0 main seed bank, 1 SS seeds,
 * 2 gapless threading, 3 local-fragment rescue, 4 threading DP, 5 dense
 * local-fragment rescue, 6 top-k seed, 7 DP refinement, 8 offset DP,
 * 9 weighted refinement; NaN where the stage did not run.  out receives
 * n_pairs x 10 doubles, row-major by pair.  Returns the number of pairs
 * written. */
int indialign_last_stage_scores(double *out, int n_pairs);

/* The Needleman-Wunsch fill and traceback of the search's NW chain kernel
 * (nw.h:10-68 semantics) on n supplied score matrices, for tests: matrix k
 * is plen[k] x nlen[k] (1..32767), row-major, concatenated in smat; gap[k]
 * is its gap open.  Outputs, concatenated per matrix: H and T
 * ((plen + 1) x (nlen + 1) each, the reference's nw_dp layout: boundary H 0
 * and trace -1, interior trace 0 diag / 1 up / 2 left) and the aligned
 * pairs in increasing order (min(plen, nlen) slots per matrix in ap / an,
 * count in na[k]).  fast = 1 runs the FP32 fill of the fast method;
 * trace_global = 1 keeps the trace in a global slot (as chains too long for
 * LDS do), 0 in LDS.  Returns 0, -1 for bad sizes (or LDS too small for the
 * requested placement), -2 without a GPU or on a device error. */
int indialign_nw_matrix(int n, const double *smat, const int *plen, const int *nlen,
                        const double *gap, int fast, int trace_global, double *H, int8_t *T,
                        int *ap, int *an, int *na);

/* The rotation solvers of the search, run on the host (the device code is
 * the same source): from the 3x3 covariance h[3k+d] = sum (p_k - cp_k)
 * (q_d - cq_d), the rotation R maximizing sum (R p') . q' (out_d =
 * sum_k p_k R[3k+d]).  method 0: Horn key matrix + 8-sweep Jacobi, the
 * reference's solver (core.h:64-123, the plain method); method 1: the
 * closed-form quaternion solve of the fast method (largest key-matrix
 * eigenvalue by Newton on its characteristic quartic, eigenvector from the
 * adjugate).  Returns 0; for method 1, 1 when the closed form declines (near
 * repeated top eigenvalue; the fast method then uses method 0); -1 for an
 * unknown method. */
int indialign_rotation_from_covariance(const double *h, int method, double *R);

/* ---- learned ranking of the local-fragment work items ---------------------
 * The ranker is an MLP 96 -> 256 -> 256 -> 1 (ReLU) over a work-item
 * descriptor; its parameters are one float array of INDIALIGN_MLP_NPARAM
 * values: W1[256][96], b1[256], W2[256][256], b2[256], w3[256], b3.  Matrix
 * products use BF16 operands with FP32 accumulation. */
#define INDIALIGN_MLP_NFEAT 96
#define INDIALIGN_MLP_NPARAM 90881

/* Install (params != NULL, copied) or remove (NULL) the ranker used by
 * method 2.  Returns 0. */
int indialign_set_ranker(const float *params);

/* Work budgets of method 2 (defaults 64, 16, 16, 16, 0, -1):
 *   top_m >= 1: dense local-fragment work items refined per pair (rescue.h
 *     dense pass), the ranker's top ones;
 *   top_m_first >= 1: the same for the first local-fragment pass;
 *   top_offsets >= 1: threading-DP offsets (rescue.h:121-190) per pair that
 *     start NW chains, chosen by their seed-bank score;
 *   top_offset_banks: threading-DP offsets per pair that run their seed
 *     bank, chosen by the TM-score of their window's own superposition
 *     (0: every offset; else at least top_offsets);
 *   top_seeds >= 0: seeds refined per bank of the SS-seed stage (its
 *     detailed searches and DP refinement, ss.h:56-164) and of threading DP,
 *     chosen by the TM-score of each seed's initial superposition; 0 refines
 *     every seed (the main DP refinement, indialign.cpp:179-223, always
 *     refines every seed);
 *   fallback_margin: a pair whose largest predicted refined score reaches
 *     its running best + fallback_margin refines every local-fragment work
 *     item of that pass, as the plain search does (< 0: never).
 * Returns 0, or -1 for a budget out of range. */
int indialign_set_learned_budgets(int top_m, int top_m_first, int top_offsets,
                                  int top_offset_banks, int top_seeds, double fallback_margin);

/* Further options of method 2 (defaults 0, 512, 64):
 *   reduce_rescored = 1 makes every local-fragment and threading-DP pass
 *     keep the work item whose transform has the best same-index TM-score
 *     (the score the pipeline keeps), 0 the plain search's rule: the item
 *     with the best chain score, rescored afterwards (rescue.h:98-116,
 *     indialign.cpp:129-152);
 *   fallback_top_m, fallback_top_m_first: the fallback budgets.  A pair
 *     whose largest predicted refined score of the dense (first)
 *     local-fragment pass is below rescue_score_2 (rescue_score_1), i.e.
 *     the ranker predicts the pass will not lift it over the threshold at
 *     which the plain search keeps searching, refines its top
 *     max(top_m, fallback_top_m) (max(top_m_first, fallback_top_m_first))
 *     items instead of top_m (top_m_first): a bounded form of falling back
 *     to the plain search's full pass.  0 disables it.
 * Returns 0, or -1 for a value out of range. */
int indialign_set_learned_options(int reduce_rescored, int fallback_top_m,
                                  int fallback_top_m_first);

/* Launch shape of the NW chain kernel and the seed-bank kernel (no
 * counterpart in the reference: its CPU loops have no launch shape).  For
 * callers and the project's autotuner (indialign_amd/autotune.py), which
 * picks the fastest shape on a sample of the workload and writes it to
 * indialign_tuning.txt beside the library (or to the file named by
 * INDIALIGN_TUNING), read at the first search.
 *   chain_waves 2 or 3: wavefronts per SIMD the chain kernel variant is
 *     compiled for (register budget 256 or 168 VGPRs);
 *   chain_slots_per_cu 0..64: persistent chain wavefronts per CU (0: as many
 *     as are resident);
 *   chain_trace 0, 1 or 2: NW trace in LDS only when that costs no resident
 *     wavefront / in LDS whenever it fits / always in a global slot;
 *   bank_pass_cap 0..64: refinement passes a seed-bank candidate runs in the
 *     bank launch; candidates still refining continue in a second launch
 *     that packs them densely (0: a single launch).
 * Results do not depend on the shape.  Returns 0, or -1 for a value out of
 * range (the setting is then unchanged). */
int indialign_set_launch_tuning(int chain_waves, int chain_slots_per_cu, int chain_trace,
                                int bank_pass_cap);
/* Length classes of the NW chain launches (default 1, also read from the
 * tuning file as "chain_buckets N"): chains are split by their pair's
 * compact native length (<= 160, <= 400, <= 900, longer; n_classes merges
 * the longest classes) into concurrent launches, each with the LDS plan and
 * occupancy of its own longest chain; 1 = one launch for every chain.
 * Results do not depend on it.  Returns 0, or -1 for n_classes outside 1..4. */
int indialign_set_chain_buckets(int n_classes);

/* Current launch shape into out4[0..3] (file or defaults 3, 0, 0, 5). */
int indialign_get_launch_tuning(int *out4);

/* Plain search (method 0, serial seed bank) that also records every
 * local-fragment work item: its descriptor row (X, INDIALIGN_MLP_NFEAT floats
 * per row), its refined score (y, best of its two chains) and the index of
 * its pair (pair_idx), up to cap rows; *n_records receives the count.
 * Returns 0, -1 on bad arguments, -2 on a device error. */
int indialign_collect_records(int n_pairs, const NativePairInput *inputs, const NativeConfig *cfg,
                              NativePairResult *outputs, float *X, float *y, int *pair_idx,
                              long long cap, long long *n_records);

/* Ranker forward pass on the GPU over n host rows: y[i] = f(X[i]). */
int indialign_mlp_forward(const float *params, const float *X, long long n, float *y);

/* Gradient of the mean squared error mean_i (f(X[i]) - t[i])^2 over n rows
 * (1 <= n <= 262144) with respect to params (grad, INDIALIGN_MLP_NPARAM
 * floats) and the loss, from the GPU forward and backward kernels. */
int indialign_mlp_grad(const float *params, const float *X, const float *t, int n, float *grad,
                       float *loss);

/* No counterpart in the reference: indialign_mlp_grad for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * training loops, harnesses that forbid hipMalloc).  params
 * (INDIALIGN_MLP_NPARAM floats), X (n x INDIALIGN_MLP_NFEAT), t (n), grad
 * (INDIALIGN_MLP_NPARAM) and loss (1 float) are device pointers; the work
 * is enqueued on `stream` (a hipStream_t; NULL = the null stream) and the
 * call returns without waiting.  The work buffers are a preallocated
 * per-GPU array of about 16.8 MB (held until the process ends), so
 * 1 <= n <= 4096; other n return -1 with a message.  Nothing in them needs
 * clearing (every output is stored, not accumulated).  All calls on one GPU
 * share that array: calls that overlap on a GPU (other threads, other
 * streams) race.  The first call on a device also fills a constant slice of
 * the array and waits for it.  Returns 0, -1 for bad arguments, -2 on a
 * device error. */
int indialign_mlp_grad_noalloc(const float *params, const float *X, const float *t, int n,
                               float *grad, float *loss, void *stream);

/* Adam training on the GPU (beta 0.9 / 0.999, eps 1e-8, constant lr):
 * `epochs` passes over the n rows in minibatches of `batch` rows, shuffled by
 * a seeded permutation; params is updated in place, epoch_loss (may be NULL)
 * receives the mean minibatch loss of each epoch. */
int indialign_mlp_train(float *params, const float *X, const float *t, long long n, int epochs,
                        int batch, float lr, unsigned seed, float *epoch_loss);

#ifdef __cplusplus
}
#endif

#endif
