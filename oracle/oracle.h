/* CPU FP64 oracle for the TM-score superposition search.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch restatement of the
 * reference algorithm (indialign_c/core.h, seeds.h, search.h, nw.h, align.h,
 * rescue.h, ss.h, threading.h, indialign.cpp:29-266).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product library never links it.
 *
 * Pins (how each function is checked, tests/test_oracle_*.py):
 *   - every function below against the reference compiled here
 *     (oracle/_ref/libref_harness.so) and against the committed fixtures in
 *     tests/golden/;
 *   - d0 / TM-score / transform against closed forms;
 *   - kabsch / quaternion rotation against an SVD Kabsch (numpy) on rigid
 *     noise-free data (exact recovery, det = +1);
 *   - NW DP + traceback against brute-force enumeration of all alignments on
 *     tiny matrices (including ties and gap-open).
 */
#ifndef INDIALIGN_ORACLE_H
#define INDIALIGN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same memory layout as the reference api.h:10-36 structs. */
typedef struct {
    int backend_mode, use_fragment_search, max_iter, dp_iter;
    double max_mem_gb;
    int request_cuda;
    double rescue_score_1, rescue_score_2, early_term_score;
} OracleConfig;

typedef struct {
    int length;
    const double *pred_xyz, *native_xyz;
    const uint8_t *valid_mask, *pred_valid_mask, *native_valid_mask;
    double lnorm;
} OraclePair;

typedef struct { double score, rotation[9], translation[3]; } OracleResult;

/* Number of pipeline stages recorded by oracle_search_traced. */
#define ORACLE_N_STAGES 10

void oracle_default_config(OracleConfig *cfg);

/* Full search (indialign.cpp:29-266).  serial_bank=1 applies early_term_score
 * to the main seed bank (the reference's serial / batch semantics,
 * search.h:83); serial_bank=0 never terminates early (the reference's
 * multithreaded single-pair semantics, search.h:100-136).  stage_trace, if
 * not NULL, receives ORACLE_N_STAGES x 13 doubles (R[9], t[3], score) of the
 * running best after each stage (NaN score = stage skipped). */
int oracle_search_traced(const OraclePair *p, const OracleConfig *cfg, OracleResult *out,
                         int serial_bank, double *stage_trace);
int oracle_search_batch(int n, const OraclePair *p, const OracleConfig *cfg,
                        OracleResult *out, int serial_bank);

/* Mismatches of the FMA-corrected quotient (a * RN(1/b), one correction)
 * against a / b over n random cases (the identity the GPU kernels use for
 * d2 / d0^2). */
long long oracle_markstein_mismatches(long long n, unsigned long long seed);

/* Single functions (inputs AoS xyz like the reference). */
void oracle_d0(double L, double *out3);
void oracle_jacobi4(double *A16, double *V16);
void oracle_cov_to_rotation(const double *H9, double *R9);
void oracle_kabsch(const double *P, const double *Q, const uint8_t *m, int N, double *R,
                   double *t);
double oracle_tm_score(const double *p, const double *q, const uint8_t *v, const double *R,
                       const double *t, double d0, double Lnorm, double sd8, int N,
                       int use_d8);
int oracle_seed_masks(int N, const uint8_t *valid, int use_frag, int max_starts,
                      uint8_t *out, int cap);
int oracle_all_seeds(const double *pred, const double *native, const uint8_t *valid, int N,
                     double Lnorm, uint8_t *out, int cap, int *n_frag);
void oracle_topk_seed(const double *d2, const uint8_t *valid, int N, double frac,
                      uint8_t *mask);
void oracle_refine_seed(const double *p, const double *q, const uint8_t *v,
                        const uint8_t *seed, double d0s, int N, int max_iter, double *R,
                        double *t, int *n_passes);
void oracle_seed_bank(const double *p, const double *q, const uint8_t *v, const uint8_t *seeds,
                      int K, const double *d0c, int C, double d0, double sd8, double Lnorm,
                      int N, int max_iter, int use_d8, double early_term, double *out13);
void oracle_nw_dp(const double *smat, int Np, int Nn, double gap, double *H, int8_t *trace);
int oracle_nw_align(const double *smat, int Np, int Nn, double gap, int *ap, int *an);
void oracle_detailed_search(const double *p, const double *q, const int *ap, const int *an,
                            int na, double d0, double d0s, double sd8, double Lnorm,
                            int use_frag, int mi_override, int max_starts, double *out13);
void oracle_dp_refine(const double *p, const double *q, const uint8_t *v, const uint8_t *pv,
                      const uint8_t *nv, const double *R0, const double *t0, double d0,
                      double d0s, double sd8, double Lnorm, int N, int max_iter, int lite,
                      double *out13);
void oracle_local_fragment_dp(const double *p, const double *q, const uint8_t *v,
                              const uint8_t *pv, const uint8_t *nv, double d0, double d0s,
                              double Lnorm, int N, int dense, double *out13);
void oracle_threading_dp(const double *p, const double *q, const uint8_t *v, const uint8_t *pv,
                         const uint8_t *nv, double d0, double d0s, double Lnorm, int N,
                         double *out13);
void oracle_assign_ss(const double *xyz, const uint8_t *v, int N, char *ss);
void oracle_ss_seeds(const double *p, const double *q, const uint8_t *v, const uint8_t *pv,
                     const uint8_t *nv, const double *R0, const double *t0, double d0,
                     double d0s, double sd8, double Lnorm, int N, int dp_iter, double *out13);
void oracle_gapless(const double *p, const double *q, const uint8_t *pv, const uint8_t *nv,
                    const uint8_t *v, double d0, double d0s, double Lnorm, int N,
                    double *out13);
void oracle_weighted_refine(const double *p, const double *q, const uint8_t *v, double d0,
                            double Lnorm, int N, int iters, double *R, double *t, double *sc);

#ifdef __cplusplus
}
#endif
#endif
