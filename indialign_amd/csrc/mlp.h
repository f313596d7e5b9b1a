// Host side of the work-item ranker: a 96-256-256-1 MLP (ReLU hidden layers)
// whose matrix products run on the matrix cores with BF16 operands and FP32
// accumulation, trained with MSE loss and Adam.  All buffers live on the
// device; every call is asynchronous on the given stream unless it says
// otherwise.
#pragma once
#include "host_util.h"
#include "ia_kernels.h"

#include <vector>

namespace ia {

class Mlp {
public:
    // Parameters (kNParam floats, layout of ia_kernels.h) from the host.
    void set_params(hipStream_t s, const float *host);
    void get_params(hipStream_t s, float *host);
    bool loaded() const { return loaded_; }
    void unload() { loaded_ = false; }
    // y[0..n) = f(X[0..n) rows); X and y on the device; any n.  One fused
    // matrix-core kernel (inference).
    void forward(hipStream_t s, const float *X, long long n, float *y);
    // The same through the training GEMM sequence (chunked).
    void forward_gemm(hipStream_t s, const float *X, long long n, float *y);
    // Gradient of mean (f(X) - t)^2 over n <= kChunk rows into grad(), and
    // the loss into loss_dev()[0].
    void backward(hipStream_t s, const float *X, const float *t, int n);
    void adam_step(hipStream_t s, float lr);
    void reset_adam(hipStream_t s);
    const float *grad() const { return G_; }
    const float *loss_dev() const { return loss_; }
    float *params_dev() { return P_; }

    static constexpr int kChunk = 1 << 18;

private:
    void alloc();
    void fwd_chunk(hipStream_t s, const float *X, int n, float *y);
    DBuf bP_, bG_, bM_, bV_, bh1_, bh2_, bdh1_, bdh2_, bones_, by_, bdy_, bloss_, bW1b_, bW2b_;
    short *W1b_ = nullptr, *W2b_ = nullptr;
    bool wb_dirty_ = true;
    float *P_ = nullptr, *G_ = nullptr, *M_ = nullptr, *V_ = nullptr;
    float *h1_ = nullptr, *h2_ = nullptr, *dh1_ = nullptr, *dh2_ = nullptr;
    float *ones_ = nullptr, *y_ = nullptr, *dy_ = nullptr, *loss_ = nullptr;
    bool loaded_ = false;
    int step_ = 0;
};

}  // namespace ia
