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

// Work buffers of the GEMM forward and backward passes for up to `rows` rows:
// h1, h2, dh1, dh2 (rows x kHid), y, dy (rows), ones (rows values of 1.0).
struct MlpWork {
    float *h1, *h2, *dh1, *dh2, *y, *dy;
    const float *ones;
};

// Forward through the GEMM sequence: y = f(X) for n rows (n <= work rows).
void mlp_forward_gemm(hipStream_t s, const MlpWork &w, const float *P, const float *X, int n,
                      float *y);
// Gradient of mean (f(X) - t)^2 over n rows into G (kNParam floats) and the
// loss into loss[0]; every output is stored, none accumulated.
void mlp_backward(hipStream_t s, const MlpWork &w, const float *P, const float *X, const float *t,
                  int n, float *G, float *loss);

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
    MlpWork work() const { return MlpWork{h1_, h2_, dh1_, dh2_, y_, dy_, ones_}; }
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
