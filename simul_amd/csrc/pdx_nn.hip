// pdx_nn.hip -- BF16 MFMA kernels of the learned warm start (a per-particle
// MLP, 4 linear layers of width 256 with ReLU).  Training and inference run
// on these kernels; host code (Python) owns the device buffers and calls the
// C ABI below with device pointers on a given HIP stream.
//
// Matrices are row-major.  Every GEMM is one kernel, C = A B with A and B read
// through element strides, so transposed operands (W^T in the data gradient,
// X^T in the weight gradient) are views, not copies.  Each wave computes a
// 32x32 tile of C with v_mfma_f32_32x32x8_bf16 (FP32 accumulation); a block of
// 4 waves covers 128 rows x 32 columns.  The weight gradient reduces over the
// particle dimension in fixed-size slices (grid z) written as FP32 partials and
// summed in slice order by a second kernel: training is deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef short pdx_bf4 __attribute__((ext_vector_type(4)));
typedef float pdx_f16v __attribute__((ext_vector_type(16)));

#define NN_TM 128  // rows per block (4 waves x 32)
#define NN_TN 32   // columns per block

// FP32 -> BF16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

enum { NN_EPI_FWD = 0, NN_EPI_DGRAD = 1, NN_EPI_WGRAD = 2 };

// C[M x N] = A[M x K] B[K x N]; A(i, k) = A[i * sa0 + k * sa1], B(k, j) = B[k * sb0 + j * sb1].
// EPI_FWD:   y = acc + bias[j], ReLU if relu; writes ybf (ldy) and/or yf (ldy)
// EPI_DGRAD: y = acc, zeroed where act(i, j) <= 0 (ReLU derivative of the layer input) if act
// EPI_WGRAD: rows of the reduction over [z * kslice, (z + 1) * kslice); FP32 partial to yf + z * M * N
template <int EPI>
__global__ __launch_bounds__(256) void k_nn_gemm(const uint16_t *__restrict__ A, int64_t sa0, int64_t sa1,
                                                 const uint16_t *__restrict__ B, int64_t sb0, int64_t sb1, int M,
                                                 int N, int K, int kslice, const float *__restrict__ bias, int relu,
                                                 const uint16_t *__restrict__ act, int64_t lda_act,
                                                 uint16_t *__restrict__ ybf, float *__restrict__ yf, int64_t ldy)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int row0 = blockIdx.x * NN_TM + wv * 32, col0 = blockIdx.y * NN_TN;
    int k_lo = 0, k_hi = K;
    if (EPI == NN_EPI_WGRAD) {
        k_lo = blockIdx.z * kslice;
        k_hi = min(K, k_lo + kslice);
    }
    const int r = row0 + (lane & 31), c = col0 + (lane & 31), kq = (lane >> 5) * 4;
    pdx_f16v acc;
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    for (int k0 = k_lo; k0 < k_hi; k0 += 8) {
        pdx_bf4 a, b;
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + kq + q;
            const bool kin = k < k_hi;
            a[q] = (short)((kin && r < M) ? A[(int64_t)r * sa0 + (int64_t)k * sa1] : 0);
            b[q] = (short)((kin && c < N) ? B[(int64_t)k * sb0 + (int64_t)c * sb1] : 0);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    // D layout of the 32x32 MFMA: lane l, register v -> row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32
    const int j = col0 + (lane & 31);
    if (j >= N) return;
    for (int v = 0; v < 16; ++v) {
        const int i = row0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (i >= M) continue;
        float y = acc[v];
        if (EPI == NN_EPI_FWD) {
            if (bias) y += bias[j];
            if (relu) y = fmaxf(y, 0.f);
        } else if (EPI == NN_EPI_DGRAD) {
            if (act && !(bf2f(act[(int64_t)i * lda_act + j]) > 0.f)) y = 0.f;
        }
        if (EPI == NN_EPI_WGRAD) {
            yf[(int64_t)blockIdx.z * M * N + (int64_t)i * N + j] = y;
        } else {
            if (ybf) ybf[(int64_t)i * ldy + j] = f2bf(y);
            if (yf) yf[(int64_t)i * ldy + j] = y;
        }
    }
}

// out[i] = sum over slices s of part[s * n + i], in slice order
__global__ __launch_bounds__(256) void k_nn_sum_slices(const float *__restrict__ part, int nslice, int64_t n,
                                                       float *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < nslice; ++z) s += part[(int64_t)z * n + i];
    out[i] = s;
}

// column sums of a BF16 matrix (M x N) in row slices: part[z * N + j]
__global__ __launch_bounds__(256) void k_nn_colsum(const uint16_t *__restrict__ Y, int M, int N, int rslice,
                                                   float *__restrict__ part)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= N) return;
    const int r0 = blockIdx.y * rslice, r1 = min(M, r0 + rslice);
    float s = 0.f;
    for (int i = r0; i < r1; ++i) s += bf2f(Y[(int64_t)i * N + j]);
    part[(int64_t)blockIdx.y * N + j] = s;
}

// masked mean-squared error of Y (M x ldy, first 3 columns, FP32) against T
// (M x 3): per-slice loss partials and the BF16 gradient 2 (y - t) w_i / W
__global__ __launch_bounds__(256) void k_nn_mse(const float *__restrict__ Y, int64_t ldy, const float *__restrict__ T,
                                                const float *__restrict__ w, int M, float inv_wsum,
                                                uint16_t *__restrict__ dY, float *__restrict__ loss_part)
{
    __shared__ float red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float l = 0.f;
    if (i < M) {
        const float wi = w ? w[i] : 1.f;
        for (int q = 0; q < 3; ++q) {
            const float d = Y[(int64_t)i * ldy + q] - T[(int64_t)i * 3 + q];
            l += wi * d * d;
            dY[(int64_t)i * ldy + q] = f2bf(2.f * wi * d * inv_wsum);
        }
        for (int q = 3; q < ldy; ++q) dY[(int64_t)i * ldy + q] = 0;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] * inv_wsum;
}

// Adam on FP32 master weights; refreshes the BF16 copy used by the GEMMs
__global__ __launch_bounds__(256) void k_nn_adam(float *__restrict__ p, const float *__restrict__ g,
                                                 float *__restrict__ m, float *__restrict__ v,
                                                 uint16_t *__restrict__ pbf, int64_t n, float lr, float b1, float b2,
                                                 float eps, float c1, float c2, float wd)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (pbf) pbf[i] = f2bf(pi);
}

__global__ __launch_bounds__(256) void k_nn_to_bf16(const float *__restrict__ x, int64_t n, uint16_t *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// ---------------------------------------------------------------------------
// C ABI (declared in include/pdx_nn.h)
// ---------------------------------------------------------------------------
static int nn_check()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : 3;
}

static int nn_grid_rows(int M) { return (M + NN_TM - 1) / NN_TM; }

extern "C" int pdx_nn_linear_fwd(const uint16_t *X, const uint16_t *W, const float *bias, int M, int K, int N,
                                 int relu, uint16_t *Ybf, float *Yf, int ldy, void *stream)
{
    if (!X || !W || M < 0 || K <= 0 || N <= 0 || ldy < N || (!Ybf && !Yf)) return 1;
    if (M == 0) return 0;
    dim3 g(nn_grid_rows(M), (N + NN_TN - 1) / NN_TN);
    hipLaunchKernelGGL(k_nn_gemm<NN_EPI_FWD>, g, dim3(256), 0, (hipStream_t)stream, X, (int64_t)K, (int64_t)1, W,
                       (int64_t)N, (int64_t)1, M, N, K, K, bias, relu, nullptr, (int64_t)0, Ybf, Yf, (int64_t)ldy);
    return nn_check();
}

extern "C" int pdx_nn_linear_dgrad(const uint16_t *dY, const uint16_t *W, int M, int K, int N, int ldy,
                                   const uint16_t *act, uint16_t *dX, void *stream)
{
    if (!dY || !W || !dX || M < 0 || K <= 0 || N <= 0 || ldy < N) return 1;
    if (M == 0) return 0;
    // dX (M x K) = dY (M x N) W^T; B(k = n, j = k_in) = W[k_in * N + n]
    dim3 g(nn_grid_rows(M), (K + NN_TN - 1) / NN_TN);
    hipLaunchKernelGGL(k_nn_gemm<NN_EPI_DGRAD>, g, dim3(256), 0, (hipStream_t)stream, dY, (int64_t)ldy, (int64_t)1,
                       W, (int64_t)1, (int64_t)N, M, K, N, N, nullptr, 0, act, (int64_t)K, dX, nullptr, (int64_t)K);
    return nn_check();
}

extern "C" int pdx_nn_linear_wgrad(const uint16_t *X, const uint16_t *dY, int M, int K, int N, int ldy,
                                   int rows_per_slice, float *part, float *dW, float *db, void *stream)
{
    if (!X || !dY || !part || !dW || M <= 0 || K <= 0 || N <= 0 || ldy != N || rows_per_slice <= 0 ||
        rows_per_slice % 8)
        return 1;
    const int ns = (M + rows_per_slice - 1) / rows_per_slice;
    // dW (K x N) = X^T (K x M) dY (M x N): A(i = k_in, k = m) = X[m * K + k_in], B(k = m, j) = dY[m * ldy + j]
    dim3 g(nn_grid_rows(K), (N + NN_TN - 1) / NN_TN, ns);
    hipLaunchKernelGGL(k_nn_gemm<NN_EPI_WGRAD>, g, dim3(256), 0, (hipStream_t)stream, X, (int64_t)1, (int64_t)K, dY,
                       (int64_t)ldy, (int64_t)1, K, N, M, rows_per_slice, nullptr, 0, nullptr, (int64_t)0, nullptr,
                       part, (int64_t)N);
    const int64_t nw = (int64_t)K * N;
    hipLaunchKernelGGL(k_nn_sum_slices, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       part, ns, nw, dW);
    if (db) {
        // bias gradient: column sums of dY in the same row slices
        hipLaunchKernelGGL(k_nn_colsum, dim3((N + 255) / 256, ns), dim3(256), 0, (hipStream_t)stream, dY, M, ldy,
                           rows_per_slice, part);
        hipLaunchKernelGGL(k_nn_sum_slices, dim3((N + 255) / 256), dim3(256), 0, (hipStream_t)stream, part, ns,
                           (int64_t)ldy, db);
    }
    return nn_check();
}

extern "C" int pdx_nn_mse(const float *Y, int ldy, const float *T, const float *w, int M, float wsum, uint16_t *dY,
                          float *loss_part, void *stream)
{
    if (!Y || !T || !dY || !loss_part || M <= 0 || ldy < 3 || !(wsum > 0.f)) return 1;
    hipLaunchKernelGGL(k_nn_mse, dim3((M + 255) / 256), dim3(256), 0, (hipStream_t)stream, Y, (int64_t)ldy, T, w, M,
                       1.f / wsum, dY, loss_part);
    return nn_check();
}

extern "C" int pdx_nn_adam(float *p, const float *g, float *m, float *v, uint16_t *pbf, int64_t n, float lr,
                           float beta1, float beta2, float eps, int64_t step, float weight_decay, void *stream)
{
    if (!p || !g || !m || !v || n <= 0 || step < 1) return 1;
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_nn_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, p, g, m, v,
                       pbf, n, lr, beta1, beta2, eps, c1, c2, weight_decay);
    return nn_check();
}

extern "C" int pdx_nn_to_bf16(const float *x, int64_t n, uint16_t *y, void *stream)
{
    if (!x || !y || n < 0) return 1;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_nn_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x, n, y);
    return nn_check();
}
