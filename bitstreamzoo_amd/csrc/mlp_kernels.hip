// mlp_kernels.hip — training and inference kernels of the schedule cost
// model (gfx950): BF16 MFMA matrix products with FP32 accumulation, and the
// elementwise / reduction kernels of backpropagation and Adam.
#include <hip/hip_runtime.h>

#include "device.h"
#include "mlp.h"

namespace bsz {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

// One wavefront per 16x16 tile of C.  V_MFMA_F32_16X16X16_BF16 operand
// layout: lane l supplies A(row l%16, k 4*(l/16)+j) and B(k 4*(l/16)+j,
// col l%16) for j = 0..3, and receives D(row 4*(l/16)+j, col l%16).
__global__ __launch_bounds__(64) void k_gemm_bf16(const uint16_t* __restrict__ A, int64_t a_rs, int64_t a_cs,
                                                  const uint16_t* __restrict__ B, int64_t b_rs, int64_t b_cs,
                                                  float* __restrict__ C, uint16_t* __restrict__ Cb,
                                                  const float* __restrict__ bias, int M, int N, int K, int relu,
                                                  int accumulate) {
    const int lane = threadIdx.x;
    const int i0 = blockIdx.y * 16, n0 = blockIdx.x * 16;
    const int r = lane & 15, q = lane >> 4;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        v4s a, b;
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + 4 * q + j;
            const int ai = i0 + r, bn = n0 + r;
            a[j] = (ai < M && k < K) ? (short)A[ai * a_rs + k * a_cs] : (short)0;
            b[j] = (bn < N && k < K) ? (short)B[k * b_rs + bn * b_cs] : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int col = n0 + r;
    if (col >= N) return;
    for (int j = 0; j < 4; ++j) {
        const int row = i0 + 4 * q + j;
        if (row >= M) continue;
        float v = acc[j];
        const int64_t o = (int64_t)row * N + col;
        if (accumulate) v += C[o];
        if (bias) v += bias[col];
        if (relu) v = v > 0.f ? v : 0.f;
        C[o] = v;
        if (Cb) Cb[o] = f32_to_bf16(v);
    }
}

__global__ __launch_bounds__(256) void k_f32_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f32_to_bf16(x[i]);
}

__global__ __launch_bounds__(256) void k_mse_grad(const float* __restrict__ pred, const float* __restrict__ y,
                                                 float* __restrict__ d, uint16_t* __restrict__ db,
                                                 float* __restrict__ loss, int n) {
    __shared__ float part[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.f;
    if (i < n) {
        e = pred[i] - y[i];
        const float g = 2.f * e / (float)n;
        d[i] = g;
        db[i] = f32_to_bf16(g);
    }
    part[threadIdx.x] = e * e / (float)n;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0]);
}

__global__ __launch_bounds__(256) void k_relu_bwd(float* __restrict__ dH, const float* __restrict__ H,
                                                 uint16_t* __restrict__ dHb, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float g = H[i] > 0.f ? dH[i] : 0.f;
        dH[i] = g;
        dHb[i] = f32_to_bf16(g);
    }
}

// One block per column.
__global__ __launch_bounds__(256) void k_col_sum(const float* __restrict__ dY, float* __restrict__ db, int rows,
                                                int cols) {
    __shared__ float part[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) s += dY[(int64_t)r * cols + c];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = part[0];
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, uint16_t* __restrict__ wb,
                                             const float* __restrict__ g, float* __restrict__ m,
                                             float* __restrict__ v, int64_t n, float lr, float b1, float b2,
                                             float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    wb[i] = f32_to_bf16(wi);
}

unsigned blocks256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void launch_gemm_bf16(hipStream_t s, const uint16_t* A, int64_t a_rs, int64_t a_cs, const uint16_t* B,
                      int64_t b_rs, int64_t b_cs, float* C, uint16_t* Cb, const float* bias, int M, int N, int K,
                      bool relu, bool accumulate) {
    if (M <= 0 || N <= 0) return;
    dim3 grid((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16));
    k_gemm_bf16<<<grid, 64, 0, s>>>(A, a_rs, a_cs, B, b_rs, b_cs, C, Cb, bias, M, N, K, relu ? 1 : 0,
                                    accumulate ? 1 : 0);
    BSZ_HIP(hipGetLastError());
}

void launch_f32_to_bf16(hipStream_t s, const float* x, uint16_t* y, int64_t n) {
    if (n <= 0) return;
    k_f32_to_bf16<<<blocks256(n), 256, 0, s>>>(x, y, n);
    BSZ_HIP(hipGetLastError());
}

void launch_mse_grad(hipStream_t s, const float* pred, const float* y, float* d, uint16_t* db, float* loss, int n) {
    if (n <= 0) return;
    k_mse_grad<<<blocks256(n), 256, 0, s>>>(pred, y, d, db, loss, n);
    BSZ_HIP(hipGetLastError());
}

void launch_relu_bwd(hipStream_t s, float* dH, const float* H, uint16_t* dHb, int64_t n) {
    if (n <= 0) return;
    k_relu_bwd<<<blocks256(n), 256, 0, s>>>(dH, H, dHb, n);
    BSZ_HIP(hipGetLastError());
}

void launch_col_sum(hipStream_t s, const float* dY, float* db, int rows, int cols) {
    if (cols <= 0) return;
    k_col_sum<<<cols, 256, 0, s>>>(dY, db, rows, cols);
    BSZ_HIP(hipGetLastError());
}

void launch_adam(hipStream_t s, float* w, uint16_t* wb, const float* g, float* m, float* v, int64_t n, float lr,
                 float b1, float b2, float eps, int t) {
    if (n <= 0) return;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    k_adam<<<blocks256(n), 256, 0, s>>>(w, wb, g, m, v, n, lr, b1, b2, eps, c1, c2);
    BSZ_HIP(hipGetLastError());
}

}  // namespace bsz
