// Test-infrastructure harness: exposes single functions of the reference
// C++ path (compiled from the reference headers where they lie, see
// oracle/Makefile) through a flat C ABI, so that per-function golden
// fixtures can be written by tests/golden/make_golden.py.  Nothing in the
// product links this file.
#include "core.h"
#include "seeds.h"
#include "search.h"
#include "nw.h"
#include "align.h"
#include "rescue.h"
#include "ss.h"
#include "threading.h"
#include <omp.h>
#include <cstring>
#include <vector>

static void put(const SeedResult &r, double *out13) {
    std::memcpy(out13, r.R, 72);
    std::memcpy(out13 + 9, r.t, 24);
    out13[12] = r.score;
}

extern "C" {

void ref_set_threads(int n) { omp_set_num_threads(n); }

void ref_d0(double L, double *out3) { d0_from_length_rna(L, &out3[0], &out3[1], &out3[2]); }

void ref_jacobi4(double *A16, double *V16) {
    double A[4][4], V[4][4];
    std::memcpy(A, A16, sizeof(A));
    jacobi4(A, V);
    std::memcpy(A16, A, sizeof(A));
    std::memcpy(V16, V, sizeof(V));
}

void ref_cov_to_rotation(const double *H9, double *R9) {
    double H[3][3];
    std::memcpy(H, H9, sizeof(H));
    cov_to_rotation(H, R9);
}

void ref_kabsch(const double *P, const double *Q, const uint8_t *m, int N, double *R, double *t) {
    kabsch(P, Q, m, N, R, t);
}

void ref_kabsch_weighted(const double *P, const double *Q, const uint8_t *v, const double *w,
                         int N, double *R, double *t) {
    kabsch_weighted(P, Q, v, w, N, R, t);
}

double ref_tm_score(const double *p, const double *q, const uint8_t *v, const double *R,
                    const double *t, double d0, double Lnorm, double sd8, int N) {
    return tm_score(p, q, v, R, t, d0, Lnorm, sd8, N);
}

double ref_tm_score_no_d8(const double *p, const double *q, const uint8_t *v, const double *R,
                          const double *t, double d0, double Lnorm, int N) {
    return tm_score_no_d8(p, q, v, R, t, d0, Lnorm, N);
}

void ref_weighted_tm_refine(const double *p, const double *q, const uint8_t *v, double d0,
                            double Lnorm, int N, int iters, double *R, double *t, double *sc) {
    weighted_tm_refine(p, q, v, d0, Lnorm, N, iters, R, t, sc);
}

// Frag/global seed masks (seeds.h build_seed_masks).  Returns K; writes up
// to cap masks of N bytes.
int ref_build_seed_masks(int N, const uint8_t *valid, int use_frag, int max_starts,
                         uint8_t *out, int cap) {
    std::vector<uint8_t> m;
    int K = 0;
    build_seed_masks(N, valid, use_frag != 0, m, K, max_starts);
    for (int k = 0; k < K && k < cap; k++) std::memcpy(out + (size_t)k * N, &m[(size_t)k * N], N);
    return K;
}

// Seeds exactly as the main search builds them: frag/global + anchors.
int ref_all_seeds(const double *pred, const double *native, const uint8_t *valid, int N,
                  double Lnorm, uint8_t *out, int cap, int *n_frag) {
    double d0, d0s, sd8;
    d0_from_length_rna(Lnorm, &d0, &d0s, &sd8);
    std::vector<uint8_t> m;
    int K = 0;
    build_seed_masks(N, valid, true, m, K);
    *n_frag = K;
    build_anchor_contact_seeds(pred, native, valid, N, d0s, d0, 5, m, K);
    for (int k = 0; k < K && k < cap; k++) std::memcpy(out + (size_t)k * N, &m[(size_t)k * N], N);
    return K;
}

void ref_topk_seed(const double *d2, const uint8_t *valid, int N, double frac, uint8_t *mask) {
    build_topk_seed(d2, valid, N, frac, mask);
}

void ref_iterative_seed_refine(const double *p, const double *q, const uint8_t *v,
                               const uint8_t *seed, double d0s, int N, int max_iter,
                               double *R, double *t) {
    std::vector<uint8_t> w(N), s(N);
    iterative_seed_refine(p, q, v, seed, d0s, N, max_iter, R, t, w.data(), s.data());
}

void ref_evaluate_seed_bank(const double *p, const double *q, const uint8_t *v,
                            const uint8_t *seeds, int K, const double *d0c, int C, double d0,
                            double sd8, double Lnorm, int N, int max_iter, int use_d8,
                            double early_term, double *out13) {
    put(evaluate_seed_bank(p, q, v, seeds, K, d0c, C, d0, sd8, Lnorm, N, max_iter,
                           use_d8 != 0, early_term), out13);
}

void ref_nw_dp(const double *smat, int Np, int Nn, double gap, double *H, int8_t *trace) {
    std::vector<double> h;
    std::vector<int8_t> tr;
    nw_dp(smat, Np, Nn, h, tr, gap);
    std::memcpy(H, h.data(), h.size() * sizeof(double));
    std::memcpy(trace, tr.data(), tr.size());
}

int ref_nw_align(const double *smat, int Np, int Nn, double gap, int *ap, int *an) {
    std::vector<double> h;
    std::vector<int8_t> tr;
    nw_dp(smat, Np, Nn, h, tr, gap);
    return nw_traceback(tr, Np, Nn, ap, an, std::min(Np, Nn));
}

void ref_alignment_detailed_search(const double *p, const double *q, const int *ap,
                                   const int *an, int na, double d0, double d0s, double sd8,
                                   double Lnorm, int N, int use_frag, int mi_override,
                                   int max_starts, double *out13) {
    put(alignment_detailed_search(p, q, ap, an, na, d0, d0s, sd8, Lnorm, N, use_frag != 0,
                                  mi_override, max_starts), out13);
}

void ref_dp_refine(const double *p, const double *q, const uint8_t *v, const uint8_t *pv,
                   const uint8_t *nv, const double *R0, const double *t0, double d0, double d0s,
                   double sd8, double Lnorm, int N, int max_iter, int lite, double *out13) {
    put(dp_refine(p, q, v, pv, nv, R0, t0, d0, d0s, sd8, Lnorm, N, max_iter, lite != 0), out13);
}

void ref_local_fragment_dp(const double *p, const double *q, const uint8_t *v, const uint8_t *pv,
                           const uint8_t *nv, double d0, double d0s, double Lnorm, int N,
                           int dense, double *out13) {
    put(evaluate_local_fragment_dp(p, q, v, pv, nv, d0, d0s, Lnorm, N, dense != 0), out13);
}

void ref_threading_dp(const double *p, const double *q, const uint8_t *v, const uint8_t *pv,
                      const uint8_t *nv, double d0, double d0s, double Lnorm, int N,
                      double *out13) {
    put(evaluate_threading_dp(p, q, v, pv, nv, d0, d0s, Lnorm, N), out13);
}

void ref_assign_rna_ss(const double *xyz, const uint8_t *v, int N, char *ss) {
    assign_rna_ss(xyz, v, N, ss);
}

void ref_evaluate_ss_seeds(const double *p, const double *q, const uint8_t *v, const uint8_t *pv,
                           const uint8_t *nv, const double *R0, const double *t0, double d0,
                           double d0s, double sd8, double Lnorm, int N, int dp_iter,
                           double *out13) {
    std::vector<char> sp(N), sn(N);
    assign_rna_ss(p, pv, N, sp.data());
    assign_rna_ss(q, nv, N, sn.data());
    put(evaluate_ss_seeds(p, q, v, pv, nv, sp.data(), sn.data(), R0, t0, d0, d0s, sd8, Lnorm,
                          N, dp_iter), out13);
}

void ref_gapless_threading(const double *p, const double *q, const uint8_t *pv,
                           const uint8_t *nv, const uint8_t *v, double d0, double d0s,
                           double Lnorm, int N, double *out13) {
    put(gapless_threading_search(p, q, pv, nv, v, d0, d0s, Lnorm, N), out13);
}

}  // extern "C"
