// mlp_host.cpp — host side of the schedule cost model (see mlp.h).
#include "device.h"
#include "frontend.h"
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace bsz {

namespace {

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t& s) { return ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

double normal(uint64_t& s) {
    const double u = uniform01(s), v = uniform01(s);
    return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
}

template <class T>
T* dalloc(size_t n) {
    void* p = nullptr;
    BSZ_HIP(hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)));
    return static_cast<T*>(p);
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

}  // namespace

Mlp::Mlp(hipStream_t s, int n_in, int h1, int h2, uint64_t seed) : s_(s), n_in_(n_in), h1_(h1), h2_(h2) {
    if (n_in <= 0 || h1 <= 0 || h2 <= 0) throw RunError("mlp: layer sizes must be positive");
    const size_t np = n_params();
    p_ = dalloc<float>(np);
    g_ = dalloc<float>(np);
    m_ = dalloc<float>(np);
    v_ = dalloc<float>(np);
    pb_ = dalloc<uint16_t>(np);
    // He initialisation for the ReLU layers, small output layer, zero biases
    std::vector<float> p(np, 0.f);
    uint64_t st = seed * 0x2545f4914f6cdd1dull + 1;
    for (size_t i = 0; i < (size_t)n_in * h1; ++i) p[off_W1() + i] = (float)(normal(st) * std::sqrt(2.0 / n_in));
    for (size_t i = 0; i < (size_t)h1 * h2; ++i) p[off_W2() + i] = (float)(normal(st) * std::sqrt(2.0 / h1));
    for (int i = 0; i < h2; ++i) p[off_w3() + i] = (float)(normal(st) * std::sqrt(1.0 / h2));
    set_params(p.data());
    BSZ_HIP(hipMemsetAsync(m_, 0, np * sizeof(float), s_));
    BSZ_HIP(hipMemsetAsync(v_, 0, np * sizeof(float), s_));
    loss_ = dalloc<float>(1);
}

Mlp::~Mlp() {
    for (void* p : std::initializer_list<void*>{p_, g_, m_, v_, pb_, x_, y_, h1f_, h2f_, out_, dh1_, dh2_, dout_, loss_,
                                                xb_, h1b_, h2b_, dh1b_, dh2b_, doutb_})
        dfree(p);
}

size_t Mlp::n_params() const { return off_b3() + 1; }

void Mlp::get_params(float* out) const {
    BSZ_HIP(hipMemcpyAsync(out, p_, n_params() * sizeof(float), hipMemcpyDeviceToHost, s_));
    BSZ_HIP(hipStreamSynchronize(s_));
}

void Mlp::set_params(const float* in) {
    BSZ_HIP(hipMemcpyAsync(p_, in, n_params() * sizeof(float), hipMemcpyHostToDevice, s_));
    launch_f32_to_bf16(s_, p_, pb_, (int64_t)n_params());
    BSZ_HIP(hipStreamSynchronize(s_));
}

void Mlp::ensure(int n) {
    if (n <= cap_) return;
    for (void* p : std::initializer_list<void*>{x_, y_, h1f_, h2f_, out_, dh1_, dh2_, dout_, xb_, h1b_, h2b_, dh1b_,
                                                dh2b_, doutb_})
        dfree(p);
    cap_ = std::max(n, 256);
    x_ = dalloc<float>((size_t)cap_ * n_in_);
    y_ = dalloc<float>(cap_);
    h1f_ = dalloc<float>((size_t)cap_ * h1_);
    h2f_ = dalloc<float>((size_t)cap_ * h2_);
    out_ = dalloc<float>(cap_);
    dh1_ = dalloc<float>((size_t)cap_ * h1_);
    dh2_ = dalloc<float>((size_t)cap_ * h2_);
    dout_ = dalloc<float>(cap_);
    xb_ = dalloc<uint16_t>((size_t)cap_ * n_in_);
    h1b_ = dalloc<uint16_t>((size_t)cap_ * h1_);
    h2b_ = dalloc<uint16_t>((size_t)cap_ * h2_);
    dh1b_ = dalloc<uint16_t>((size_t)cap_ * h1_);
    dh2b_ = dalloc<uint16_t>((size_t)cap_ * h2_);
    doutb_ = dalloc<uint16_t>(cap_);
}

void Mlp::upload_xy(const float* x, const float* y, int n) {
    ensure(n);
    BSZ_HIP(hipMemcpyAsync(x_, x, (size_t)n * n_in_ * sizeof(float), hipMemcpyHostToDevice, s_));
    if (y) BSZ_HIP(hipMemcpyAsync(y_, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice, s_));
    launch_f32_to_bf16(s_, x_, xb_, (int64_t)n * n_in_);
}

void Mlp::forward(int n) {
    const uint16_t* W1 = pb_ + off_W1();
    const uint16_t* W2 = pb_ + off_W2();
    const uint16_t* w3 = pb_ + off_w3();
    // H1 = relu(X W1 + b1), H2 = relu(H1 W2 + b2), out = H2 w3 + b3
    launch_gemm_bf16(s_, xb_, n_in_, 1, W1, h1_, 1, h1f_, h1b_, p_ + off_b1(), n, h1_, n_in_, true, false);
    launch_gemm_bf16(s_, h1b_, h1_, 1, W2, h2_, 1, h2f_, h2b_, p_ + off_b2(), n, h2_, h1_, true, false);
    launch_gemm_bf16(s_, h2b_, h2_, 1, w3, 1, 1, out_, nullptr, p_ + off_b3(), n, 1, h2_, false, false);
}

void Mlp::backward(int n) {
    const uint16_t* W2 = pb_ + off_W2();
    const uint16_t* w3 = pb_ + off_w3();
    // output layer: dw3 = H2^T dout, db3 = sum dout
    launch_gemm_bf16(s_, h2b_, 1, h2_, doutb_, 1, 1, g_ + off_w3(), nullptr, nullptr, h2_, 1, n, false, false);
    launch_col_sum(s_, dout_, g_ + off_b3(), n, 1);
    // dH2 = (dout w3^T) * relu'(H2)
    launch_gemm_bf16(s_, doutb_, 1, 1, w3, 0, 1, dh2_, nullptr, nullptr, n, h2_, 1, false, false);
    launch_relu_bwd(s_, dh2_, h2f_, dh2b_, (int64_t)n * h2_);
    // dW2 = H1^T dH2, db2 = colsum dH2
    launch_gemm_bf16(s_, h1b_, 1, h1_, dh2b_, h2_, 1, g_ + off_W2(), nullptr, nullptr, h1_, h2_, n, false, false);
    launch_col_sum(s_, dh2_, g_ + off_b2(), n, h2_);
    // dH1 = (dH2 W2^T) * relu'(H1)
    launch_gemm_bf16(s_, dh2b_, h2_, 1, W2, 1, h2_, dh1_, nullptr, nullptr, n, h1_, h2_, false, false);
    launch_relu_bwd(s_, dh1_, h1f_, dh1b_, (int64_t)n * h1_);
    // dW1 = X^T dH1, db1 = colsum dH1
    launch_gemm_bf16(s_, xb_, 1, n_in_, dh1b_, h1_, 1, g_ + off_W1(), nullptr, nullptr, n_in_, h1_, n, false, false);
    launch_col_sum(s_, dh1_, g_ + off_b1(), n, h1_);
}

void Mlp::predict(const float* x, int n, float* out) {
    if (n <= 0) return;
    upload_xy(x, nullptr, n);
    forward(n);
    BSZ_HIP(hipMemcpyAsync(out, out_, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s_));
    BSZ_HIP(hipStreamSynchronize(s_));
}

float Mlp::gradients(const float* x, const float* y, int n, float* grad) {
    if (n <= 0) throw RunError("mlp: empty batch");
    upload_xy(x, y, n);
    forward(n);
    BSZ_HIP(hipMemsetAsync(loss_, 0, sizeof(float), s_));
    launch_mse_grad(s_, out_, y_, dout_, doutb_, loss_, n);
    backward(n);
    float loss = 0.f;
    BSZ_HIP(hipMemcpyAsync(grad, g_, n_params() * sizeof(float), hipMemcpyDeviceToHost, s_));
    BSZ_HIP(hipMemcpyAsync(&loss, loss_, sizeof(float), hipMemcpyDeviceToHost, s_));
    BSZ_HIP(hipStreamSynchronize(s_));
    return loss;
}

float Mlp::train_step(const float* x, const float* y, int n, float lr) {
    if (n <= 0) throw RunError("mlp: empty batch");
    upload_xy(x, y, n);
    forward(n);
    BSZ_HIP(hipMemsetAsync(loss_, 0, sizeof(float), s_));
    launch_mse_grad(s_, out_, y_, dout_, doutb_, loss_, n);
    backward(n);
    launch_adam(s_, p_, pb_, g_, m_, v_, (int64_t)n_params(), lr, 0.9f, 0.999f, 1e-8f, ++t_);
    float loss = 0.f;
    BSZ_HIP(hipMemcpyAsync(&loss, loss_, sizeof(float), hipMemcpyDeviceToHost, s_));
    BSZ_HIP(hipStreamSynchronize(s_));
    return loss;
}

float Mlp::train(const float* x, const float* y, int n, int epochs, int batch, float lr, uint64_t seed) {
    if (n <= 0 || batch <= 0) throw RunError("mlp: empty training set");
    batch = std::min(batch, n);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::vector<float> xs((size_t)batch * n_in_), ys(batch);
    uint64_t st = seed ^ 0x5851f42d4c957f2dull;
    float epoch_loss = 0.f;
    for (int e = 0; e < epochs; ++e) {
        for (int i = n - 1; i > 0; --i) std::swap(order[i], order[splitmix(st) % (uint64_t)(i + 1)]);
        double acc = 0.0;
        int nb = 0;
        for (int b0 = 0; b0 < n; b0 += batch) {
            const int m = std::min(batch, n - b0);
            for (int r = 0; r < m; ++r) {
                std::memcpy(&xs[(size_t)r * n_in_], x + (size_t)order[b0 + r] * n_in_, sizeof(float) * n_in_);
                ys[r] = y[order[b0 + r]];
            }
            acc += train_step(xs.data(), ys.data(), m, lr);
            ++nb;
        }
        epoch_loss = (float)(acc / std::max(1, nb));
    }
    return epoch_loss;
}

}  // namespace bsz
