// kernels_mlp.hip -- the per-aggregate parameter network: an MLP
// 12 -> 64 -> 64 -> 2 with GELU, FP32, forward and backward on the FP32
// matrix cores (v_mfma_f32_16x16x4f32), full-precision FP32 Adam.
//
// Parameters live in one FP32 vector (row-major):
//   W1 [16][64] (inputs 12..15 are zero padding), b1 [64],
//   W2 [64][64], b2 [64], W3 [64][16] (outputs 2..15 are padding), b3 [16].
// Batches are padded to a multiple of 64 rows; padded rows carry zero weight in
// the loss, so they contribute nothing to gradients.
// Loss: mean squared error over the two outputs and the valid rows.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <mutex>

#include "hz_internal.h"

// Scratch memory of hz_mlp_forward_noalloc: the padded inputs and the two
// hidden layers and padded outputs of up to kNoAllocRows rows (160 floats a
// row, 10 MiB).  Each GPU holds its own copy for the life of the process; host
// code reaches it only through the address hipGetSymbolAddress returns for the
// current device, never through this name.
constexpr int kNoAllocRows = HZ_MLP_NOALLOC_MAX_ROWS;
constexpr size_t kNoAllocFloats = (size_t)kNoAllocRows * (16 + 64 + 64 + 16);
static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace hz {

namespace {

constexpr int kIn = 16, kHid = 64, kOut = 16;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid * kOut, kParams = kOffB3 + kOut;
static_assert(kParams == HZ_MLP_PARAMS, "parameter layout");
constexpr int kRows = 64;   // batch rows per tile

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

enum Epi { kEpiNone = 0, kEpiGelu = 1, kEpiGeluGrad = 2 };

// C[M x N] = A[M x K] * op(B) (+ bias), K <= 64, N <= 64; op(B) = B [K x N] or
// B^T with B stored [N x K].  One workgroup (4 wavefronts) per 64-row tile;
// wavefront w owns rows 16w..16w+15 and all N columns as 16x16 MFMA blocks.
// Epilogues: kEpiGelu stores Z = acc + bias and C = gelu(Z); kEpiGeluGrad
// stores C = acc * gelu'(Zin).
// MFMA f32 16x16x4 lanes: A[lane % 16][k + lane / 16], B[k + lane / 16][lane % 16],
// accumulator element r = D[4 (lane / 16) + r][lane % 16].
template <bool kTransB, int kEpi>
__global__ __launch_bounds__(256) void k_mlp_gemm(int M, int N, int K, const float *__restrict__ A,
                                                  const float *__restrict__ B, const float *__restrict__ bias,
                                                  float *__restrict__ C, float *__restrict__ Z,
                                                  const float *__restrict__ Zin) {
    __shared__ float As[kRows][65];
    __shared__ float Bs[64][65];
    const int r0 = blockIdx.x * kRows;
    const int tid = threadIdx.x;
    for (int e = tid; e < kRows * 64; e += 256) {
        const int r = e / 64, k = e % 64;
        As[r][k] = (r0 + r < M && k < K) ? A[(size_t)(r0 + r) * K + k] : 0.0f;
    }
    for (int e = tid; e < 64 * 64; e += 256) {
        const int k = e / 64, c = e % 64;
        float v = 0.0f;
        if (k < K && c < N) v = kTransB ? B[(size_t)c * K + k] : B[(size_t)k * N + c];
        Bs[k][c] = v;
    }
    __syncthreads();
    const int w = tid / kWave, lane = tid % kWave;
    const int nb = (N + 15) / 16;
    f4 acc[4];
    for (int j = 0; j < 4; ++j) acc[j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    const int kpad = (K + 3) & ~3;
    for (int ks = 0; ks < kpad; ks += 4) {
        const float a = As[16 * w + lane % 16][ks + lane / 16];
        for (int j = 0; j < nb; ++j) {
            const float b = Bs[ks + lane / 16][16 * j + lane % 16];
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0);
        }
    }
    for (int j = 0; j < nb; ++j)
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + 16 * w + 4 * (lane / 16) + r;
            const int col = 16 * j + lane % 16;
            if (row >= M || col >= N) continue;
            const size_t o = (size_t)row * N + col;
            float v = acc[j][r];
            if (kEpi == kEpiGeluGrad) {
                C[o] = v * gelu_grad(Zin[o]);
            } else {
                if (bias) v += bias[col];
                if (kEpi == kEpiGelu) {
                    if (Z) Z[o] = v;   // kept for the backward pass; null in a forward-only pass
                    C[o] = gelu(v);
                } else {
                    C[o] = v;
                }
            }
        }
}

// Weight gradient partials: part[g] = sum over the rows of this workgroup's
// tiles of A^T D, A [M x KA], D [M x ND], KA, ND <= 64; plus column sums of D
// (bias gradient) in part[g][KA*ND ...].  The output tile (KA x ND) is split
// over the 4 wavefronts by 16-row bands of KA.
__global__ __launch_bounds__(256) void k_mlp_wgrad(int M, int KA, int ND, const float *__restrict__ A,
                                                   const float *__restrict__ D, float *__restrict__ part) {
    __shared__ float As[kRows][65];
    __shared__ float Ds[kRows][65];
    const int tid = threadIdx.x;
    const int w = tid / kWave, lane = tid % kWave;
    const int nb = (ND + 15) / 16;
    f4 acc[4];
    for (int j = 0; j < 4; ++j) acc[j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    float csum = 0.0f;   // column sums: thread t < ND owns column t
    const int ntiles = (M + kRows - 1) / kRows;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int r0 = t * kRows;
        __syncthreads();
        for (int e = tid; e < kRows * 64; e += 256) {
            const int r = e / 64, c = e % 64;
            const bool ok = r0 + r < M;
            As[r][c] = (ok && c < KA) ? A[(size_t)(r0 + r) * KA + c] : 0.0f;
            Ds[r][c] = (ok && c < ND) ? D[(size_t)(r0 + r) * ND + c] : 0.0f;
        }
        __syncthreads();
        if (16 * w < KA) {
            // C[i][j] += sum_b A[b][i] D[b][j]: MFMA A operand = A^T (rows i), K = batch rows
            for (int ks = 0; ks < kRows; ks += 4) {
                const float a = As[ks + lane / 16][16 * w + lane % 16];
                for (int j = 0; j < nb; ++j) {
                    const float b = Ds[ks + lane / 16][16 * j + lane % 16];
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0);
                }
            }
        }
        if (tid < ND)
            for (int r = 0; r < kRows; ++r) csum += Ds[r][tid];
    }
    float *out = part + (size_t)blockIdx.x * (KA * ND + ND);
    if (16 * w < KA)
        for (int j = 0; j < nb; ++j)
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * w + 4 * (lane / 16) + r;
                const int c = 16 * j + lane % 16;
                if (i < KA && c < ND) out[i * ND + c] = acc[j][r];
            }
    if (tid < ND) out[KA * ND + tid] = csum;
}

// grad[o + e] = sum over partials g (in order) of part[g][e]
__global__ void k_mlp_reduce(int G, int len, const float *__restrict__ part, float *__restrict__ out) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < len; e += gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int g = 0; g < G; ++g) s += part[(size_t)g * len + e];
        out[e] = s;
    }
}

// dY = 2 (Y - T) w / (2 * nvalid) on the two outputs, 0 elsewhere; per-block loss partials
__global__ __launch_bounds__(256) void k_mlp_dloss(int M, int nvalid, const float *__restrict__ Y,
                                                   const float *__restrict__ T, float *__restrict__ dY,
                                                   double *__restrict__ lpart) {
    __shared__ double sh[256 / kWave];
    double l = 0.0;
    const float scale = 1.0f / (2.0f * (float)nvalid);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M * kOut; i += gridDim.x * blockDim.x) {
        const int row = i / kOut, o = i % kOut;
        float g = 0.0f;
        if (row < nvalid && o < HZ_MLP_OUTPUTS) {
            const float d = Y[i] - T[row * HZ_MLP_OUTPUTS + o];
            g = 2.0f * d * scale;
            l += (double)d * (double)d;
        }
        dY[i] = g;
    }
    const double bs = block_sum<256>(l, sh);
    if (threadIdx.x == 0) lpart[blockIdx.x] = bs;
}

__global__ void k_mlp_loss_final(int G, int nvalid, const double *__restrict__ lpart, float *__restrict__ loss) {
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int g = 0; g < G; ++g) s += lpart[g];
        *loss = (float)(s / (2.0 * nvalid));
    }
}

__global__ void k_mlp_adam(int n, float *__restrict__ th, const float *__restrict__ g, float *__restrict__ m,
                           float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        th[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// padded input rows: x[b][0..11], zeros in 12..15 and in padded rows
__global__ void k_mlp_pad_inputs(int B, int M, const float *__restrict__ x, const int *__restrict__ idx,
                                 float *__restrict__ X) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M * kIn; i += gridDim.x * blockDim.x) {
        const int row = i / kIn, c = i % kIn;
        float v = 0.0f;
        if (row < B && c < HZ_MLP_INPUTS) v = x[(size_t)(idx ? idx[row] : row) * HZ_MLP_INPUTS + c];
        X[i] = v;
    }
}

__global__ void k_mlp_gather_targets(int B, const float *__restrict__ t, const int *__restrict__ idx,
                                     float *__restrict__ T) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < B * HZ_MLP_OUTPUTS; i += gridDim.x * blockDim.x) {
        const int row = i / HZ_MLP_OUTPUTS, o = i % HZ_MLP_OUTPUTS;
        T[i] = t[(size_t)(idx ? idx[row] : row) * HZ_MLP_OUTPUTS + o];
    }
}

__global__ void k_mlp_take_outputs(int B, const float *__restrict__ Y, float *__restrict__ y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < B * HZ_MLP_OUTPUTS; i += gridDim.x * blockDim.x)
        y[i] = Y[(i / HZ_MLP_OUTPUTS) * kOut + i % HZ_MLP_OUTPUTS];
}

inline int gsz(long n) { return (int)std::max(1L, std::min<long>((n + 255) / 256, 4096)); }

}  // namespace

MlpWork mlp_work(int capacity) {
    MlpWork w;
    w.cap = (capacity + kRows - 1) / kRows * kRows;
    const size_t M = (size_t)w.cap;
    w.X = arena().alloc_n<float>(M * kIn);
    w.Z1 = arena().alloc_n<float>(M * kHid);
    w.H1 = arena().alloc_n<float>(M * kHid);
    w.Z2 = arena().alloc_n<float>(M * kHid);
    w.H2 = arena().alloc_n<float>(M * kHid);
    w.Y = arena().alloc_n<float>(M * kOut);
    w.T = arena().alloc_n<float>(M * HZ_MLP_OUTPUTS + 1);
    w.dY = arena().alloc_n<float>(M * kOut);
    w.dH = arena().alloc_n<float>(M * kHid);
    w.dZ = arena().alloc_n<float>(M * kHid);
    w.part = arena().alloc_n<float>((size_t)kMlpWgradBlocks * (kHid * kHid + kHid));
    w.lpart = arena().alloc_n<double>(4096);
    return w;
}

void mlp_forward(const float *theta, const MlpWork &w, int B) {
    hipStream_t st = stream();
    const int M = (B + kRows - 1) / kRows * kRows;
    const int g = M / kRows;
    k_mlp_gemm<false, kEpiGelu><<<g, 256, 0, st>>>(M, kHid, kIn, w.X, theta + kOffW1, theta + kOffB1, w.H1, w.Z1,
                                                   nullptr);
    k_mlp_gemm<false, kEpiGelu><<<g, 256, 0, st>>>(M, kHid, kHid, w.H1, theta + kOffW2, theta + kOffB2, w.H2, w.Z2,
                                                   nullptr);
    k_mlp_gemm<false, kEpiNone><<<g, 256, 0, st>>>(M, kOut, kHid, w.H2, theta + kOffW3, theta + kOffB3, w.Y,
                                                   nullptr, nullptr);
    HZ_LAUNCH_CHECK();
}

void mlp_inputs(const float *x, const int *idx, int B, const MlpWork &w) {
    const int M = (B + kRows - 1) / kRows * kRows;
    k_mlp_pad_inputs<<<gsz((long)M * kIn), 256, 0, stream()>>>(B, M, x, idx, w.X);
    HZ_LAUNCH_CHECK();
}

void mlp_outputs(int B, const MlpWork &w, float *y) {
    k_mlp_take_outputs<<<gsz((long)B * 2), 256, 0, stream()>>>(B, w.Y, y);
    HZ_LAUNCH_CHECK();
}

void mlp_targets(const float *t, const int *idx, int B, const MlpWork &w) {
    k_mlp_gather_targets<<<gsz((long)B * 2), 256, 0, stream()>>>(B, t, idx, w.T);
    HZ_LAUNCH_CHECK();
}

void mlp_backward(const float *theta, const MlpWork &w, int B, float *grad, float *loss) {
    hipStream_t st = stream();
    const int M = (B + kRows - 1) / kRows * kRows;
    const int g = M / kRows;
    const int gl = gsz((long)M * kOut);
    k_mlp_dloss<<<gl, 256, 0, st>>>(M, B, w.Y, w.T, w.dY, w.lpart);
    k_mlp_loss_final<<<1, 64, 0, st>>>(gl, B, w.lpart, loss);
    const int gw = std::min(kMlpWgradBlocks, g);
    auto wgrad = [&](int KA, int ND, const float *A, const float *D, float *gW) {
        k_mlp_wgrad<<<gw, 256, 0, st>>>(M, KA, ND, A, D, w.part);
        const int len = KA * ND + ND;   // weights then bias (contiguous in theta)
        k_mlp_reduce<<<gsz(len), 256, 0, st>>>(gw, len, w.part, gW);
    };
    // layer 3
    wgrad(kHid, kOut, w.H2, w.dY, grad + kOffW3);
    k_mlp_gemm<true, kEpiGeluGrad><<<g, 256, 0, st>>>(M, kHid, kOut, w.dY, theta + kOffW3, nullptr, w.dZ,
                                                      nullptr, w.Z2);
    // layer 2
    wgrad(kHid, kHid, w.H1, w.dZ, grad + kOffW2);
    k_mlp_gemm<true, kEpiGeluGrad><<<g, 256, 0, st>>>(M, kHid, kHid, w.dZ, theta + kOffW2, nullptr, w.dH,
                                                      nullptr, w.Z1);
    // layer 1
    wgrad(kIn, kHid, w.X, w.dH, grad + kOffW1);
    HZ_LAUNCH_CHECK();
}

// The device address of `scratch` on the current device, looked up once per
// device (nullptr when the lookup fails).
static float *noalloc_scratch() {
    static std::mutex mu;
    static float *addr[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            std::fprintf(stderr, "hazmath_mi355x: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return nullptr;
        }
        addr[dev] = static_cast<float *>(p);
    }
    return addr[dev];
}

int mlp_forward_noalloc(const float *theta, const float *x, int B, float *y, hipStream_t st) {
    float *base = noalloc_scratch();
    if (!base) return HZ_ERROR_DEVICE;
    float *X = base;
    float *H1 = X + (size_t)kNoAllocRows * kIn;
    float *H2 = H1 + (size_t)kNoAllocRows * kHid;
    float *Y = H2 + (size_t)kNoAllocRows * kHid;
    const int M = (B + kRows - 1) / kRows * kRows;
    const int g = M / kRows;
    k_mlp_pad_inputs<<<gsz((long)M * kIn), 256, 0, st>>>(B, M, x, nullptr, X);
    k_mlp_gemm<false, kEpiGelu><<<g, 256, 0, st>>>(M, kHid, kIn, X, theta + kOffW1, theta + kOffB1, H1, nullptr,
                                                   nullptr);
    k_mlp_gemm<false, kEpiGelu><<<g, 256, 0, st>>>(M, kHid, kHid, H1, theta + kOffW2, theta + kOffB2, H2, nullptr,
                                                   nullptr);
    k_mlp_gemm<false, kEpiNone><<<g, 256, 0, st>>>(M, kOut, kHid, H2, theta + kOffW3, theta + kOffB3, Y, nullptr,
                                                   nullptr);
    k_mlp_take_outputs<<<gsz((long)B * 2), 256, 0, st>>>(B, Y, y);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::fprintf(stderr, "hazmath_mi355x: hz_mlp_forward_noalloc launch failed: %s\n", hipGetErrorString(e));
        return HZ_ERROR_DEVICE;
    }
    return HZ_SUCCESS;
}

void mlp_adam(float *theta, const float *grad, float *m, float *v, float lr, int step) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
    k_mlp_adam<<<gsz(kParams), 256, 0, stream()>>>(kParams, theta, grad, m, v, lr, b1, b2, eps, c1, c2);
    HZ_LAUNCH_CHECK();
}

}  // namespace hz
