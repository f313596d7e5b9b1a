// Work-item ranker MLP: forward, backward and Adam as sequences of matrix-core
// GEMM launches (k_learn.hip).  Shapes, with X the n x 96 descriptor rows:
//   h1 = relu(X W1^T + b1)      n x 256
//   h2 = relu(h1 W2^T + b2)     n x 256
//   y  = h2 w3 + b3 + X[:, 31]  n   (residual on the refined same-index TM)
// Backward of L = mean (y - t)^2:
//   dy = 2 (y - t) / n,  dh2 = dy w3^T * [h2 > 0],  dh1 = (dh2 W2) * [h1 > 0]
//   gW2 = dh2^T h1, gW1 = dh1^T X, gw3 = h2^T dy, gb = column sums.
#include "mlp.h"

#include <cstring>

namespace ia {

namespace {

GemmArgs gemm(const float *A, long long sam, long long sak, const float *B, long long sbk,
              long long sbn, float *C, long long scm, long long scn, int M, int N, int K,
              int epi = EPI_STORE, const float *bias = nullptr, const float *mask = nullptr) {
    GemmArgs g;
    g.A = A; g.B = B; g.C = C;
    g.sam = sam; g.sak = sak; g.sbk = sbk; g.sbn = sbn; g.scm = scm; g.scn = scn;
    g.M = M; g.N = N; g.K = K; g.epi = epi; g.bias = bias; g.mask = mask;
    g.skip = nullptr; g.sskip = 0;
    return g;
}

}  // namespace

void Mlp::alloc() {
    if (P_) return;
    P_ = bP_.get<float>(kNParam);
    G_ = bG_.get<float>(kNParam);
    M_ = bM_.get<float>(kNParam);
    V_ = bV_.get<float>(kNParam);
    h1_ = bh1_.get<float>((size_t)kChunk * kHid);
    h2_ = bh2_.get<float>((size_t)kChunk * kHid);
    dh1_ = bdh1_.get<float>((size_t)kChunk * kHid);
    dh2_ = bdh2_.get<float>((size_t)kChunk * kHid);
    ones_ = bones_.get<float>(kChunk);
    y_ = by_.get<float>(kChunk);
    dy_ = bdy_.get<float>(kChunk);
    loss_ = bloss_.get<float>(4);
    W1b_ = bW1b_.get<short>((size_t)kHid * kFeat);
    W2b_ = bW2b_.get<short>((size_t)kHid * kHid);
    std::vector<float> one(kChunk, 1.0f);
    IA_CHECK(hipMemcpy(ones_, one.data(), sizeof(float) * kChunk, hipMemcpyHostToDevice));
}

void Mlp::set_params(hipStream_t s, const float *host) {
    alloc();
    IA_CHECK(hipMemcpyAsync(P_, host, sizeof(float) * kNParam, hipMemcpyHostToDevice, s));
    IA_CHECK(hipStreamSynchronize(s));
    loaded_ = true;
    wb_dirty_ = true;
}

void Mlp::get_params(hipStream_t s, float *host) {
    IA_CHECK(hipMemcpyAsync(host, P_, sizeof(float) * kNParam, hipMemcpyDeviceToHost, s));
    IA_CHECK(hipStreamSynchronize(s));
}

void Mlp::fwd_chunk(hipStream_t s, const float *X, int n, float *y) {
    launch_gemm(s, gemm(X, kFeat, 1, P_ + kOffW1, 1, kFeat, h1_, kHid, 1, n, kHid, kFeat,
                        EPI_BIAS_RELU, P_ + kOffB1));
    launch_gemm(s, gemm(h1_, kHid, 1, P_ + kOffW2, 1, kHid, h2_, kHid, 1, n, kHid, kHid,
                        EPI_BIAS_RELU, P_ + kOffB2));
    GemmArgs g3 = gemm(h2_, kHid, 1, P_ + kOffW3, 1, 0, y, 1, 0, n, 1, kHid, EPI_BIAS_SKIP,
                       P_ + kOffB3);
    g3.skip = X + kSkipCol;
    g3.sskip = kFeat;
    launch_gemm(s, g3);
}

void Mlp::forward(hipStream_t s, const float *X, long long n, float *y) {
    alloc();
    if (n <= 0) return;
    if (wb_dirty_) {
        launch_weights_bf16(s, P_, W1b_, W2b_);
        wb_dirty_ = false;
    }
    launch_mlp_infer(s, X, n, W1b_, W2b_, P_, y);
    IA_CHECK(hipGetLastError());
}

void Mlp::forward_gemm(hipStream_t s, const float *X, long long n, float *y) {
    alloc();
    for (long long r = 0; r < n; r += kChunk) {
        const int m = (int)std::min<long long>(kChunk, n - r);
        fwd_chunk(s, X + r * kFeat, m, y + r);
    }
    IA_CHECK(hipGetLastError());
}

void Mlp::backward(hipStream_t s, const float *X, const float *t, int n) {
    alloc();
    if (n <= 0 || n > kChunk) throw DevError("Mlp::backward: batch size out of range");
    fwd_chunk(s, X, n, y_);
    launch_mse(s, y_, t, n, dy_, loss_);
    // head: gw3 = h2^T dy, gb3 = sum dy
    launch_gemm(s, gemm(h2_, 1, kHid, dy_, 1, 0, G_ + kOffW3, 1, 0, kHid, 1, n));
    launch_gemm(s, gemm(dy_, 0, 1, ones_, 1, 0, G_ + kOffB3, 0, 0, 1, 1, n));
    launch_head_back(s, dy_, P_ + kOffW3, h2_, n, dh2_);
    // layer 2: gW2 = dh2^T h1, gb2 = dh2^T 1, dh1 = (dh2 W2) * [h1 > 0]
    launch_gemm(s, gemm(dh2_, 1, kHid, h1_, kHid, 1, G_ + kOffW2, kHid, 1, kHid, kHid, n));
    launch_gemm(s, gemm(dh2_, 1, kHid, ones_, 1, 0, G_ + kOffB2, 1, 0, kHid, 1, n));
    launch_gemm(s, gemm(dh2_, kHid, 1, P_ + kOffW2, kHid, 1, dh1_, kHid, 1, n, kHid, kHid,
                        EPI_MASK, nullptr, h1_));
    // layer 1: gW1 = dh1^T X, gb1 = dh1^T 1
    launch_gemm(s, gemm(dh1_, 1, kHid, X, kFeat, 1, G_ + kOffW1, kFeat, 1, kHid, kFeat, n));
    launch_gemm(s, gemm(dh1_, 1, kHid, ones_, 1, 0, G_ + kOffB1, 1, 0, kHid, 1, n));
    IA_CHECK(hipGetLastError());
}

void Mlp::reset_adam(hipStream_t s) {
    alloc();
    IA_CHECK(hipMemsetAsync(M_, 0, sizeof(float) * kNParam, s));
    IA_CHECK(hipMemsetAsync(V_, 0, sizeof(float) * kNParam, s));
    step_ = 0;
}

void Mlp::adam_step(hipStream_t s, float lr) {
    step_++;
    wb_dirty_ = true;
    launch_adam(s, P_, G_, M_, V_, kNParam, lr, 0.9f, 0.999f, 1e-8f, step_);
    IA_CHECK(hipGetLastError());
}

}  // namespace ia
