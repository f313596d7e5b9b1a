// Work-item ranker MLP: forward, backward and Adam as sequences of matrix-core
// GEMM launches (k_learn.hip).  Shapes, with X the n x 96 descriptor rows:
//   h1 = relu(X W1^T + b1)      n x 256
//   h2 = relu(h1 W2^T + b2)     n x 256
//   y  = h2 w3 + b3 + X[:, 31]  n   (residual on the refined same-index TM)
// Backward of L = mean (y - t)^2:
//   dy = 2 (y - t) / n,  dh2 = dy w3^T * [h2 > 0],  dh1 = (dh2 W2) * [h1 > 0]
//   gW2 = dh2^T h1, gW1 = dh1^T X, gw3 = h2^T dy, gb = column sums.
#include "mlp.h"

#include <cstdio>
#include <cstring>
#include <mutex>

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

void mlp_forward_gemm(hipStream_t s, const MlpWork &w, const float *P, const float *X, int n,
                      float *y) {
    launch_gemm(s, gemm(X, kFeat, 1, P + kOffW1, 1, kFeat, w.h1, kHid, 1, n, kHid, kFeat,
                        EPI_BIAS_RELU, P + kOffB1));
    launch_gemm(s, gemm(w.h1, kHid, 1, P + kOffW2, 1, kHid, w.h2, kHid, 1, n, kHid, kHid,
                        EPI_BIAS_RELU, P + kOffB2));
    GemmArgs g3 = gemm(w.h2, kHid, 1, P + kOffW3, 1, 0, y, 1, 0, n, 1, kHid, EPI_BIAS_SKIP,
                       P + kOffB3);
    g3.skip = X + kSkipCol;
    g3.sskip = kFeat;
    launch_gemm(s, g3);
}

void mlp_backward(hipStream_t s, const MlpWork &w, const float *P, const float *X, const float *t,
                  int n, float *G, float *loss) {
    mlp_forward_gemm(s, w, P, X, n, w.y);
    launch_mse(s, w.y, t, n, w.dy, loss);
    // head: gw3 = h2^T dy, gb3 = sum dy
    launch_gemm(s, gemm(w.h2, 1, kHid, w.dy, 1, 0, G + kOffW3, 1, 0, kHid, 1, n));
    launch_gemm(s, gemm(w.dy, 0, 1, w.ones, 1, 0, G + kOffB3, 0, 0, 1, 1, n));
    launch_head_back(s, w.dy, P + kOffW3, w.h2, n, w.dh2);
    // layer 2: gW2 = dh2^T h1, gb2 = dh2^T 1, dh1 = (dh2 W2) * [h1 > 0]
    launch_gemm(s, gemm(w.dh2, 1, kHid, w.h1, kHid, 1, G + kOffW2, kHid, 1, kHid, kHid, n));
    launch_gemm(s, gemm(w.dh2, 1, kHid, w.ones, 1, 0, G + kOffB2, 1, 0, kHid, 1, n));
    launch_gemm(s, gemm(w.dh2, kHid, 1, P + kOffW2, kHid, 1, w.dh1, kHid, 1, n, kHid, kHid,
                        EPI_MASK, nullptr, w.h1));
    // layer 1: gW1 = dh1^T X, gb1 = dh1^T 1
    launch_gemm(s, gemm(w.dh1, 1, kHid, X, kFeat, 1, G + kOffW1, kFeat, 1, kHid, kFeat, n));
    launch_gemm(s, gemm(w.dh1, 1, kHid, w.ones, 1, 0, G + kOffB1, 1, 0, kHid, 1, n));
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
        mlp_forward_gemm(s, work(), P_, X + r * kFeat, m, y + r);
    }
    IA_CHECK(hipGetLastError());
}

void Mlp::backward(hipStream_t s, const float *X, const float *t, int n) {
    alloc();
    if (n <= 0 || n > kChunk) throw DevError("Mlp::backward: batch size out of range");
    mlp_backward(s, work(), P_, X, t, n, G_, loss_);
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

// ---- gradient entry without allocation ------------------------------------
// The work buffers of indialign_mlp_grad_noalloc live in one file-scope
// device array (`scratch`, kNoAllocBytes per GPU, held until the process
// ends).  Host code never passes the array's name to a kernel: at the first
// call on each device hipGetSymbolAddress gives that device's address, which
// is kept per device and used for every launch.  Every call on a GPU shares
// its array, so calls that overlap on one GPU race.
namespace {

constexpr int kNoAllocRows = 4096;
constexpr size_t kNoAllocFloats = 4 * (size_t)kNoAllocRows * ia::kHid + 3 * (size_t)kNoAllocRows;
constexpr int kMaxDevices = 64;

__device__ __align__(256) float scratch[kNoAllocFloats];

std::mutex g_na_mu;
float *g_na_ptr[kMaxDevices] = {};

__global__ void fill_value_kernel(float *p, int n, float v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// This device's scratch address; at the first call on the device, the
// `ones` slice (read by the bias-gradient products, never written) is
// filled.  nullptr on error (message on stderr).
float *device_scratch(hipStream_t s) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        std::fprintf(stderr, "indialign_amd: no usable HIP device for the no-allocation entry\n");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        float *p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            std::fprintf(stderr, "indialign_amd: hipGetSymbolAddress(scratch): %s\n",
                         hipGetErrorString(e));
            return nullptr;
        }
        float *ones = p + 4 * (size_t)kNoAllocRows * ia::kHid + 2 * (size_t)kNoAllocRows;
        hipLaunchKernelGGL(fill_value_kernel, dim3((kNoAllocRows + 255) / 256), dim3(256), 0, s,
                           ones, kNoAllocRows, 1.0f);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
            std::fprintf(stderr, "indialign_amd: scratch set-up failed\n");
            return nullptr;
        }
        g_na_ptr[dev] = p;
    }
    return g_na_ptr[dev];
}

}  // namespace

extern "C" int indialign_mlp_grad_noalloc(const float *params, const float *X, const float *t,
                                          int n, float *grad, float *loss, void *stream) {
    if (n < 1 || n > kNoAllocRows) {
        std::fprintf(stderr,
                     "indialign_amd: indialign_mlp_grad_noalloc: n = %d rows is outside 1..%d, "
                     "the size of its preallocated work buffers\n",
                     n, kNoAllocRows);
        return -1;
    }
    if (!params || !X || !t || !grad || !loss) return -1;
    const hipStream_t s = (hipStream_t)stream;
    float *p = device_scratch(s);
    if (!p) return -2;
    const size_t blk = (size_t)kNoAllocRows * ia::kHid;
    ia::MlpWork w;
    w.h1 = p;
    w.h2 = p + blk;
    w.dh1 = p + 2 * blk;
    w.dh2 = p + 3 * blk;
    w.y = p + 4 * blk;
    w.dy = w.y + kNoAllocRows;
    w.ones = w.dy + kNoAllocRows;
    ia::mlp_backward(s, w, params, X, t, n, grad, loss);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::fprintf(stderr, "indialign_amd: indialign_mlp_grad_noalloc: %s\n", hipGetErrorString(e));
        return -2;
    }
    return 0;
}
