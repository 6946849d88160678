// mlp.h — the schedule cost model: a small MLP trained and evaluated on the
// GPU with hand-written kernels (BF16 MFMA matrix products, FP32
// accumulation, FP32 master weights and Adam).
//
//   x (n_in) -> relu(x W1 + b1) (h1) -> relu(. W2 + b2) (h2) -> . w3 + b3 (1)
//
// It ranks candidate schedules (fusion partition + tile configuration) of a
// program by predicted log run time; it never touches program outputs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace bsz {

// ── kernels (mlp_kernels.hip) ──
// C[M,N] (+)= op(A)[M,K] . op(B)[K,N] in bf16 x bf16 -> fp32 on MFMA 16x16x16.
// A element (i, k) is at A[i * a_rs + k * a_cs] (so transposes are strides);
// likewise B (k, n) at B[k * b_rs + n * b_cs].  Epilogue: + bias[n] (if bias),
// relu (if relu), store fp32 C (row-major, ldc = N) and, if Cb, a bf16 copy.
void launch_gemm_bf16(hipStream_t s, const uint16_t* A, int64_t a_rs, int64_t a_cs, const uint16_t* B,
                      int64_t b_rs, int64_t b_cs, float* C, uint16_t* Cb, const float* bias, int M, int N, int K,
                      bool relu, bool accumulate);
void launch_f32_to_bf16(hipStream_t s, const float* x, uint16_t* y, int64_t n);
// dY = 2 (y_pred - y) / n  (fp32 and bf16), loss += sum (y_pred - y)^2 / n
void launch_mse_grad(hipStream_t s, const float* pred, const float* y, float* d, uint16_t* db, float* loss, int n);
// dH *= (H > 0), with a bf16 copy
void launch_relu_bwd(hipStream_t s, float* dH, const float* H, uint16_t* dHb, int64_t n);
// db[c] (+)= sum_r dY[r, c]
void launch_col_sum(hipStream_t s, const float* dY, float* db, int rows, int cols);
void launch_adam(hipStream_t s, float* w, uint16_t* wb, const float* g, float* m, float* v, int64_t n, float lr,
                 float b1, float b2, float eps, int t);

// ── host model ──
class Mlp {
public:
    Mlp(hipStream_t s, int n_in, int h1, int h2, uint64_t seed);
    ~Mlp();
    int n_in() const { return n_in_; }
    int h1() const { return h1_; }
    int h2() const { return h2_; }
    // Parameters as one fp32 vector: W1[n_in][h1], b1[h1], W2[h1][h2], b2[h2], w3[h2], b3.
    size_t n_params() const;
    void get_params(float* out) const;
    void set_params(const float* in);
    // Predictions for n rows of x (row-major n x n_in), synchronous.
    void predict(const float* x, int n, float* out);
    // Mean squared error and its gradient (same layout as the parameters) on
    // a batch, without updating; synchronous.
    float gradients(const float* x, const float* y, int n, float* grad);
    // One Adam step on a batch; returns the batch loss (synchronous).
    float train_step(const float* x, const float* y, int n, float lr);
    // `epochs` passes over (x, y) in minibatches of `batch` rows (order
    // shuffled per epoch with `seed`); data resident on the device for the
    // whole call.  Returns the loss of the last epoch.
    float train(const float* x, const float* y, int n, int epochs, int batch, float lr, uint64_t seed);

private:
    hipStream_t s_;
    int n_in_, h1_, h2_;
    int cap_ = 0;   // batch capacity of the activation buffers
    int t_ = 0;     // Adam step
    // fp32 master parameters, bf16 copies, gradients, Adam moments (one block each)
    float* p_ = nullptr;
    float* g_ = nullptr;
    float* m_ = nullptr;
    float* v_ = nullptr;
    uint16_t* pb_ = nullptr;
    // activations
    float *x_ = nullptr, *y_ = nullptr, *h1f_ = nullptr, *h2f_ = nullptr, *out_ = nullptr;
    float *dh1_ = nullptr, *dh2_ = nullptr, *dout_ = nullptr, *loss_ = nullptr;
    uint16_t *xb_ = nullptr, *h1b_ = nullptr, *h2b_ = nullptr, *dh1b_ = nullptr, *dh2b_ = nullptr, *doutb_ = nullptr;
    void ensure(int n);
    size_t off_W1() const { return 0; }
    size_t off_b1() const { return (size_t)n_in_ * h1_; }
    size_t off_W2() const { return off_b1() + h1_; }
    size_t off_b2() const { return off_W2() + (size_t)h1_ * h2_; }
    size_t off_w3() const { return off_b2() + h2_; }
    size_t off_b3() const { return off_w3() + h2_; }
    void forward(int n);           // from x_/xb_ (device), fills h1, h2, out
    void backward(int n);          // from dout_, fills g_
    void upload_xy(const float* x, const float* y, int n);
};

}  // namespace bsz
