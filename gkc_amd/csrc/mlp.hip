// Partition-size predictor: MLP 256 -> 128 -> 128 -> 2 (ReLU between),
// FP32 master weights trained with these forward / backward kernels and
// Adam, BF16 inference on the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16)
// with FP32 accumulation.  Inputs: a 256-entry 4-mer sketch per partition;
// outputs: normalised log total and log distinct k-mer counts.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gkc_hip.h"

namespace {

constexpr int NI = 256, NH = 128, NO = 2;
// flat parameter layout: W1[NI][NH], b1[NH], W2[NH][NH], b2[NH], W3[NH][NO], b3[NO]
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NI * NH;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH * NO;
constexpr int NPARAM = OFF_B3 + NO;

// C[m][n] = sum_k A[m*sam + k*sak] * B[k*sbk + n*sbn] (+ bias[n]) (relu)
// (* (mask[m*N+n] > 0) when mask is given); tiled 16x16 through LDS.
__global__ __launch_bounds__(256) void k_gemm(const float* __restrict__ A, long sam, long sak, const float* __restrict__ B,
                                              long sbk, long sbn, float* __restrict__ C, int M, int N, int K,
                                              const float* __restrict__ bias, int relu, const float* __restrict__ mask,
                                              int accumulate) {
    __shared__ float sa[16][17], sb[16][17];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int m = blockIdx.y * 16 + ty, n = blockIdx.x * 16 + tx;
    float acc = 0.f;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const int ka = k0 + tx, kb = k0 + ty;
        sa[ty][tx] = (m < M && ka < K) ? A[m * sam + ka * sak] : 0.f;
        sb[ty][tx] = (kb < K && n < N) ? B[kb * sbk + n * sbn] : 0.f;
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 16; t++) acc += sa[ty][t] * sb[t][tx];
        __syncthreads();
    }
    if (m < M && n < N) {
        if (bias) acc += bias[n];
        if (relu) acc = acc > 0.f ? acc : 0.f;
        if (mask) acc = mask[(long)m * N + n] > 0.f ? acc : 0.f;
        if (accumulate) C[(long)m * N + n] += acc;
        else C[(long)m * N + n] = acc;
    }
}

// column sums: out[n] = sum_m X[m][n]
__global__ void k_colsum(const float* __restrict__ X, int M, int N, float* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; m++) s += X[(long)m * N + n];
    out[n] = s;
}

// MSE on (y - t): loss = sum / (M*NO); dy = 2 (y - t) / (M*NO)
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, int M, float* __restrict__ dy,
                      float* __restrict__ loss) {
    __shared__ float s[256];
    float acc = 0.f;
    const float inv = 1.f / (float)(M * NO);
    for (int i = threadIdx.x; i < M * NO; i += blockDim.x) {
        const float d = y[i] - t[i];
        dy[i] = 2.f * d * inv;
        acc += d * d;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = s[0] * inv;
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2, float wd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a; m2[i] = b;
    w[i] -= lr * ((a / c1) / (sqrtf(b / c2) + eps) + wd * w[i]);
}

// ---- BF16 inference on the matrix cores ----
typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ unsigned short to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (unsigned short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__global__ void k_to_bf16(const float* __restrict__ x, unsigned short* __restrict__ y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// One 32 x 32 output tile of act[32][K] (bf16, LDS, row stride lda) times
// W[K][ldw] (bf16, global, columns n0..n0+31).  Operand layout of
// v_mfma_f32_32x32x16_bf16: lane l supplies A row (l % 32), k = 8*(l/32)..+7
// and B column (l % 32), k = 8*(l/32)..+7; the 16 accumulators of lane l hold
// column (l % 32), rows 8*(v/4) + 4*(l/32) + (v%4).
__device__ __forceinline__ f32x16 tile32(const unsigned short* act, int lda, const unsigned short* __restrict__ W, int ldw,
                                         int n0, int ncols, int K) {
    const int lane = threadIdx.x % 64;
    const int r = lane % 32, kq = (lane / 32) * 8;
    f32x16 acc = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            a[t] = (short)act[r * lda + k0 + kq + t];
            const int col = n0 + r;
            b[t] = col < ncols ? (short)W[(long)(k0 + kq + t) * ldw + col] : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// One wave per 32 input rows; activations live in LDS as bf16.
__global__ __launch_bounds__(64) void k_mlp_infer_bf16(const float* __restrict__ X, int M, const unsigned short* __restrict__ W1,
                                                      const float* __restrict__ b1, const unsigned short* __restrict__ W2,
                                                      const float* __restrict__ b2, const unsigned short* __restrict__ W3,
                                                      const float* __restrict__ b3, float* __restrict__ Y) {
    __shared__ unsigned short sx[32 * NI];
    __shared__ unsigned short sh[32 * NH];
    const int lane = threadIdx.x;
    const int row0 = blockIdx.x * 32;
    for (int i = lane; i < 32 * NI; i += 64) {
        const int r = i / NI, c = i % NI;
        sx[i] = (row0 + r < M) ? to_bf16(X[(long)(row0 + r) * NI + c]) : (unsigned short)0;
    }
    __syncthreads();
    const int col = lane % 32, rq = (lane / 32) * 4;
    // layer 1: 256 -> 128, relu
    f32x16 h1[NH / 32];
#pragma unroll
    for (int t = 0; t < NH / 32; t++) h1[t] = tile32(sx, NI, W1, NH, t * 32, NH, NI);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NH / 32; t++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int r = 8 * (v / 4) + rq + (v % 4), cc = t * 32 + col;
            const float z = h1[t][v] + b1[cc];
            sh[r * NH + cc] = to_bf16(z > 0.f ? z : 0.f);
        }
    __syncthreads();
    // layer 2: 128 -> 128, relu (result back into sx as the next activations)
    f32x16 h2[NH / 32];
#pragma unroll
    for (int t = 0; t < NH / 32; t++) h2[t] = tile32(sh, NH, W2, NH, t * 32, NH, NH);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NH / 32; t++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int r = 8 * (v / 4) + rq + (v % 4), cc = t * 32 + col;
            const float z = h2[t][v] + b2[cc];
            sx[r * NH + cc] = to_bf16(z > 0.f ? z : 0.f);
        }
    __syncthreads();
    // layer 3: 128 -> 2 (columns 2..31 of the tile are zero padding)
    const f32x16 y = tile32(sx, NH, W3, NO, 0, NO, NH);
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int r = 8 * (v / 4) + rq + (v % 4);
        if (col < NO && row0 + r < M) Y[(long)(row0 + r) * NO + col] = y[v] + b3[col];
    }
}

// Features of the device predictor: x[p][i] = (log1p(sketch[p][i]) - mean[i]) / sd[i]
// (the normalisation the predictor was trained with), one workgroup per partition.
__global__ __launch_bounds__(NI) void k_presize_features(const uint32_t* __restrict__ sketch, const float* __restrict__ norm,
                                                         float* __restrict__ x) {
    const int p = blockIdx.x, i = threadIdx.x;
    x[(long)p * NI + i] = (log1pf((float)sketch[(long)p * NI + i]) - norm[i]) / norm[NI + i];
}

// Table capacities from the predictions: partition p holds total[p] k-mers;
// its distinct count is total * exp(min(log ratio, 0)) with the log ratio
// de-normalised from output 1; capacity = min(ceil(distinct * margin) + 16,
// total).  Also their sum (one workgroup of 1024 threads, P <= 1024).
// norm: [0, NI) mean, [NI, 2 NI) sd, then t_mean[2], t_sd[2], margin.
__global__ __launch_bounds__(1024) void k_presize_caps(const float* __restrict__ y, const uint32_t* __restrict__ total,
                                                       uint32_t P, const float* __restrict__ norm,
                                                       uint32_t* __restrict__ caps, uint64_t* __restrict__ cap_sum) {
    __shared__ unsigned long long s_sum;
    const uint32_t p = threadIdx.x;
    if (p == 0) s_sum = 0;
    __syncthreads();
    if (p < P) {
        const float* tn = norm + 2 * NI;
        const float lr = y[p * NO + 1] * tn[3] + tn[1];
        const double t = (double)total[p];
        const double d = t * (double)expf(fminf(lr, 0.f));
        const double c = fmin(ceil(d * (double)tn[4]) + 16.0, t);
        caps[p] = (uint32_t)c;
        atomicAdd(&s_sum, (unsigned long long)c);
    }
    __syncthreads();
    if (p == 0) *cap_sum = s_sum;
}

// Plain matrix product through the same tile routine, for checking the
// operand layout: C[32][N] = A[32][K] * B[K][N] (bf16 in, fp32 out).
__global__ __launch_bounds__(64) void k_bf16_matmul32(const float* __restrict__ A, const unsigned short* __restrict__ B, int K,
                                                     int N, float* __restrict__ C) {
    __shared__ unsigned short sa[32 * 256];
    const int lane = threadIdx.x;
    for (int i = lane; i < 32 * K; i += 64) sa[i] = to_bf16(A[i]);
    __syncthreads();
    const int col = lane % 32, rq = (lane / 32) * 4;
    for (int n0 = 0; n0 < N; n0 += 32) {
        const f32x16 acc = tile32(sa, K, B, N, n0, N, K);
        for (int v = 0; v < 16; v++) {
            const int r = 8 * (v / 4) + rq + (v % 4);
            if (n0 + col < N) C[r * N + n0 + col] = acc[v];
        }
    }
}

struct Buf {
    float* p = nullptr;
    size_t n = 0;
    ~Buf() { if (p) (void)hipFree(p); }
    hipError_t ensure(size_t m) {
        if (m <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
        hipError_t e = hipMalloc(&p, m * sizeof(float));
        if (e == hipSuccess) n = m;
        return e;
    }
};

std::string& mlp_err() { static thread_local std::string e; return e; }

}  // namespace

struct gkc_mlp {
    std::mutex mu;
    int device = 0;
    hipStream_t st = nullptr;
    Buf w, g, m1, m2, tmp;         // params, grads, adam moments, scratch
    Buf x, t, z1, a1, z2, a2, y, dy, d2, d1, loss;
    unsigned short* wb = nullptr;  // bf16 copy of the weights
    bool wb_valid = false;
    int step = 0;
};

#define MLPX(call)                                                                            \
    do {                                                                                      \
        hipError_t _e = (call);                                                               \
        if (_e != hipSuccess) { mlp_err() = std::string(#call) + ": " + hipGetErrorString(_e); \
                                (void)hipGetLastError(); return GKC_EHIP; }                   \
    } while (0)

static int mlp_fail(const std::string& m, int code) { mlp_err() = m; return code; }

static dim3 g2(int M, int N) { return dim3((N + 15) / 16, (M + 15) / 16); }

// forward over M rows in x (device) -> y; keeps activations for backward
static int forward_f32(gkc_mlp* c, int M) {
    float* W = c->w.p;
    MLPX(c->z1.ensure((size_t)M * NH)); MLPX(c->a1.ensure((size_t)M * NH));
    MLPX(c->z2.ensure((size_t)M * NH)); MLPX(c->a2.ensure((size_t)M * NH)); MLPX(c->y.ensure((size_t)M * NO));
    hipLaunchKernelGGL(k_gemm, g2(M, NH), dim3(256), 0, c->st, c->x.p, (long)NI, 1L, W + OFF_W1, (long)NH, 1L, c->z1.p, M, NH, NI,
                       W + OFF_B1, 0, nullptr, 0);
    hipLaunchKernelGGL(k_gemm, g2(M, NH), dim3(256), 0, c->st, c->x.p, (long)NI, 1L, W + OFF_W1, (long)NH, 1L, c->a1.p, M, NH, NI,
                       W + OFF_B1, 1, nullptr, 0);
    hipLaunchKernelGGL(k_gemm, g2(M, NH), dim3(256), 0, c->st, c->a1.p, (long)NH, 1L, W + OFF_W2, (long)NH, 1L, c->z2.p, M, NH, NH,
                       W + OFF_B2, 0, nullptr, 0);
    hipLaunchKernelGGL(k_gemm, g2(M, NH), dim3(256), 0, c->st, c->a1.p, (long)NH, 1L, W + OFF_W2, (long)NH, 1L, c->a2.p, M, NH, NH,
                       W + OFF_B2, 1, nullptr, 0);
    hipLaunchKernelGGL(k_gemm, g2(M, NO), dim3(256), 0, c->st, c->a2.p, (long)NH, 1L, W + OFF_W3, (long)NO, 1L, c->y.p, M, NO, NH,
                       W + OFF_B3, 0, nullptr, 0);
    MLPX(hipGetLastError());
    return 0;
}

// backward of the MSE loss against t (device) -> grads in g
static int backward_f32(gkc_mlp* c, int M) {
    float* W = c->w.p;
    float* G = c->g.p;
    MLPX(c->dy.ensure((size_t)M * NO)); MLPX(c->d2.ensure((size_t)M * NH)); MLPX(c->d1.ensure((size_t)M * NH));
    MLPX(c->loss.ensure(4));
    hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, c->st, c->y.p, c->t.p, M, c->dy.p, c->loss.p);
    // layer 3: dW3 = a2^T dy, db3 = colsum(dy), d2 = (dy W3^T) * (z2 > 0)
    hipLaunchKernelGGL(k_gemm, g2(NH, NO), dim3(256), 0, c->st, c->a2.p, 1L, (long)NH, c->dy.p, (long)NO, 1L, G + OFF_W3, NH, NO, M,
                       nullptr, 0, nullptr, 0);
    hipLaunchKernelGGL(k_colsum, dim3(1), dim3(NO), 0, c->st, c->dy.p, M, NO, G + OFF_B3);
    hipLaunchKernelGGL(k_gemm, g2(M, NH), dim3(256), 0, c->st, c->dy.p, (long)NO, 1L, W + OFF_W3, 1L, (long)NO, c->d2.p, M, NH, NO,
                       nullptr, 0, c->z2.p, 0);
    // layer 2
    hipLaunchKernelGGL(k_gemm, g2(NH, NH), dim3(256), 0, c->st, c->a1.p, 1L, (long)NH, c->d2.p, (long)NH, 1L, G + OFF_W2, NH, NH, M,
                       nullptr, 0, nullptr, 0);
    hipLaunchKernelGGL(k_colsum, dim3(1), dim3(NH), 0, c->st, c->d2.p, M, NH, G + OFF_B2);
    hipLaunchKernelGGL(k_gemm, g2(M, NH), dim3(256), 0, c->st, c->d2.p, (long)NH, 1L, W + OFF_W2, 1L, (long)NH, c->d1.p, M, NH, NH,
                       nullptr, 0, c->z1.p, 0);
    // layer 1
    hipLaunchKernelGGL(k_gemm, g2(NI, NH), dim3(256), 0, c->st, c->x.p, 1L, (long)NI, c->d1.p, (long)NH, 1L, G + OFF_W1, NI, NH, M,
                       nullptr, 0, nullptr, 0);
    hipLaunchKernelGGL(k_colsum, dim3(1), dim3(NH), 0, c->st, c->d1.p, M, NH, G + OFF_B1);
    MLPX(hipGetLastError());
    return 0;
}

extern "C" {

uint64_t gkc_mlp_n_params(void) { return NPARAM; }
const char* gkc_mlp_last_error(void) { return mlp_err().c_str(); }

gkc_mlp* gkc_mlp_create(int device, uint64_t seed) {
    gkc_mlp* c = new gkc_mlp();
    c->device = device;
    auto bad = [&](const char* w) { mlp_err() = w; delete c; return (gkc_mlp*)nullptr; };
    if (hipSetDevice(device) != hipSuccess) return bad("hipSetDevice failed");
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) return bad("stream");
    if (c->w.ensure(NPARAM) || c->g.ensure(NPARAM) || c->m1.ensure(NPARAM) || c->m2.ensure(NPARAM)) return bad("alloc");
    if (hipMalloc(&c->wb, NPARAM * sizeof(unsigned short)) != hipSuccess) return bad("alloc");
    // He-uniform init from a splitmix64 stream, biases zero
    std::vector<float> h(NPARAM, 0.f);
    uint64_t s = seed ^ 0x5DEECE66DULL;
    auto nxt = [&]() { uint64_t z = (s += 0x9E3779B97F4A7C15ULL); z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
                       z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL; return (double)((z ^ (z >> 31)) >> 11) / 9007199254740992.0; };
    auto fill = [&](int off, int fan_in, int n) { const double a = std::sqrt(6.0 / fan_in);
                                                  for (int i = 0; i < n; i++) h[off + i] = (float)((2 * nxt() - 1) * a); };
    fill(OFF_W1, NI, NI * NH); fill(OFF_W2, NH, NH * NH); fill(OFF_W3, NH, NH * NO);
    if (hipMemcpy(c->w.p, h.data(), NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess) return bad("copy");
    if (hipMemset(c->m1.p, 0, NPARAM * 4) != hipSuccess || hipMemset(c->m2.p, 0, NPARAM * 4) != hipSuccess) return bad("memset");
    return c;
}

void gkc_mlp_destroy(gkc_mlp* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->st) (void)hipStreamSynchronize(c->st);
    if (c->wb) (void)hipFree(c->wb);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

int gkc_mlp_get_params(gkc_mlp* c, float* out) {
    if (!c) return mlp_fail("null", GKC_EINVAL);
    std::lock_guard<std::mutex> lk(c->mu);
    MLPX(hipSetDevice(c->device));
    MLPX(hipStreamSynchronize(c->st));
    MLPX(hipMemcpy(out, c->w.p, NPARAM * 4, hipMemcpyDeviceToHost));
    return 0;
}

int gkc_mlp_set_params(gkc_mlp* c, const float* in) {
    if (!c) return mlp_fail("null", GKC_EINVAL);
    std::lock_guard<std::mutex> lk(c->mu);
    MLPX(hipSetDevice(c->device));
    MLPX(hipStreamSynchronize(c->st));
    MLPX(hipMemcpy(c->w.p, in, NPARAM * 4, hipMemcpyHostToDevice));
    MLPX(hipMemset(c->m1.p, 0, NPARAM * 4)); MLPX(hipMemset(c->m2.p, 0, NPARAM * 4));
    c->step = 0;
    c->wb_valid = false;
    return 0;
}

// FP32 forward (the training path) for n rows of host features.
int gkc_mlp_forward(gkc_mlp* c, const float* x, int64_t n, float* y) {
    if (!c || n < 0) return mlp_fail("bad args", GKC_EINVAL);
    if (n == 0) return 0;
    std::lock_guard<std::mutex> lk(c->mu);
    MLPX(hipSetDevice(c->device));
    MLPX(c->x.ensure((size_t)n * NI));
    MLPX(hipMemcpyAsync(c->x.p, x, (size_t)n * NI * 4, hipMemcpyHostToDevice, c->st));
    int r = forward_f32(c, (int)n);
    if (r) return r;
    MLPX(hipMemcpyAsync(y, c->y.p, (size_t)n * NO * 4, hipMemcpyDeviceToHost, c->st));
    MLPX(hipStreamSynchronize(c->st));
    return 0;
}

// Loss and the full gradient (flat, NPARAM floats) on n rows: for checks.
int gkc_mlp_loss_grad(gkc_mlp* c, const float* x, const float* t, int64_t n, float* loss, float* grad) {
    if (!c || n <= 0) return mlp_fail("bad args", GKC_EINVAL);
    std::lock_guard<std::mutex> lk(c->mu);
    MLPX(hipSetDevice(c->device));
    MLPX(c->x.ensure((size_t)n * NI)); MLPX(c->t.ensure((size_t)n * NO));
    MLPX(hipMemcpyAsync(c->x.p, x, (size_t)n * NI * 4, hipMemcpyHostToDevice, c->st));
    MLPX(hipMemcpyAsync(c->t.p, t, (size_t)n * NO * 4, hipMemcpyHostToDevice, c->st));
    int r = forward_f32(c, (int)n);
    if (!r) r = backward_f32(c, (int)n);
    if (r) return r;
    if (loss) MLPX(hipMemcpyAsync(loss, c->loss.p, 4, hipMemcpyDeviceToHost, c->st));
    if (grad) MLPX(hipMemcpyAsync(grad, c->g.p, NPARAM * 4, hipMemcpyDeviceToHost, c->st));
    MLPX(hipStreamSynchronize(c->st));
    return 0;
}

// Mini-batch Adam over (x, t) rows for `epochs` passes (rows shuffled per
// epoch from `seed`).  losses[epoch] receives the mean batch loss.
int gkc_mlp_train(gkc_mlp* c, const float* x, const float* t, int64_t n, int epochs, int batch, float lr, float wd,
                  uint64_t seed, float* losses) {
    if (!c || n <= 0 || epochs < 0 || batch <= 0) return mlp_fail("bad args", GKC_EINVAL);
    std::lock_guard<std::mutex> lk(c->mu);
    MLPX(hipSetDevice(c->device));
    std::vector<int64_t> perm(n);
    for (int64_t i = 0; i < n; i++) perm[i] = i;
    std::vector<float> bx((size_t)batch * NI), bt((size_t)batch * NO);
    MLPX(c->x.ensure((size_t)batch * NI)); MLPX(c->t.ensure((size_t)batch * NO));
    uint64_t s = seed;
    auto rnd = [&]() { uint64_t z = (s += 0x9E3779B97F4A7C15ULL); z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
                       z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL; return z ^ (z >> 31); };
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; e++) {
        for (int64_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
        double lsum = 0;
        int nb = 0;
        for (int64_t b0 = 0; b0 < n; b0 += batch) {
            const int m = (int)std::min<int64_t>(batch, n - b0);
            for (int i = 0; i < m; i++) {
                std::memcpy(&bx[(size_t)i * NI], x + perm[b0 + i] * NI, NI * 4);
                std::memcpy(&bt[(size_t)i * NO], t + perm[b0 + i] * NO, NO * 4);
            }
            MLPX(hipMemcpyAsync(c->x.p, bx.data(), (size_t)m * NI * 4, hipMemcpyHostToDevice, c->st));
            MLPX(hipMemcpyAsync(c->t.p, bt.data(), (size_t)m * NO * 4, hipMemcpyHostToDevice, c->st));
            int r = forward_f32(c, m);
            if (!r) r = backward_f32(c, m);
            if (r) return r;
            c->step++;
            const float c1 = 1.f - std::pow(b1, (float)c->step), c2 = 1.f - std::pow(b2, (float)c->step);
            hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, c->st, c->w.p, c->g.p, c->m1.p, c->m2.p,
                               NPARAM, lr, b1, b2, eps, c1, c2, wd);
            MLPX(hipGetLastError());
            float l = 0;
            MLPX(hipMemcpyAsync(&l, c->loss.p, 4, hipMemcpyDeviceToHost, c->st));
            MLPX(hipStreamSynchronize(c->st));   // bx/bt are reused by the next batch
            lsum += l; nb++;
        }
        if (losses) losses[e] = (float)(lsum / std::max(nb, 1));
    }
    c->wb_valid = false;
    return 0;
}

// BF16 matrix-core inference for n rows of host features.
int gkc_mlp_forward_bf16(gkc_mlp* c, const float* x, int64_t n, float* y) {
    if (!c || n < 0) return mlp_fail("bad args", GKC_EINVAL);
    if (n == 0) return 0;
    std::lock_guard<std::mutex> lk(c->mu);
    MLPX(hipSetDevice(c->device));
    if (!c->wb_valid) {
        hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, c->st, c->w.p, c->wb, NPARAM);
        MLPX(hipGetLastError());
        c->wb_valid = true;
    }
    MLPX(c->x.ensure((size_t)n * NI)); MLPX(c->y.ensure((size_t)n * NO));
    MLPX(hipMemcpyAsync(c->x.p, x, (size_t)n * NI * 4, hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_mlp_infer_bf16, dim3((unsigned)((n + 31) / 32)), dim3(64), 0, c->st, c->x.p, (int)n, c->wb + OFF_W1,
                       c->w.p + OFF_B1, c->wb + OFF_W2, c->w.p + OFF_B2, c->wb + OFF_W3, c->w.p + OFF_B3, c->y.p);
    MLPX(hipGetLastError());
    MLPX(hipMemcpyAsync(y, c->y.p, (size_t)n * NO * 4, hipMemcpyDeviceToHost, c->st));
    MLPX(hipStreamSynchronize(c->st));
    return 0;
}

}  // extern "C"

// Learned table presizing on the device, queued on `st` (no host round trip):
// features from the partition sketch into x[P][256], BF16 MFMA inference into
// y[P][2], capacities and their sum (x, y, caps, cap_sum: the caller's device
// buffers).  `norm` is device memory in k_presize_caps' layout.  Only the
// MLP's weights are read (their BF16 copy is refreshed first when stale).
namespace gkc {
int gkc_mlp_presize_device(gkc_mlp* c, const uint32_t* sketch, const uint32_t* total, uint32_t P, const float* norm,
                           float* x, float* y, uint32_t* caps, uint64_t* cap_sum, hipStream_t st) {
    if (!c || P == 0 || P > 1024) return mlp_fail("presize: bad arguments", GKC_EINVAL);
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->wb_valid) {
        hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, st, c->w.p, c->wb, NPARAM);
        MLPX(hipGetLastError());
        c->wb_valid = true;
    }
    hipLaunchKernelGGL(k_presize_features, dim3(P), dim3(NI), 0, st, sketch, norm, x);
    hipLaunchKernelGGL(k_mlp_infer_bf16, dim3((P + 31) / 32), dim3(64), 0, st, (const float*)x, (int)P, c->wb + OFF_W1,
                       c->w.p + OFF_B1, c->wb + OFF_W2, c->w.p + OFF_B2, c->wb + OFF_W3, c->w.p + OFF_B3, y);
    hipLaunchKernelGGL(k_presize_caps, dim3(1), dim3(1024), 0, st, (const float*)y, total, P, norm, caps, cap_sum);
    MLPX(hipGetLastError());
    return 0;
}
}  // namespace gkc

extern "C" {

// C[32][N] = A[32][K] * B[K][N] through the inference kernel's MFMA tile
// routine (inputs rounded to bf16).  K <= 256, K % 16 == 0.
int gkc_bf16_matmul32(int device, const float* A, const float* B, int K, int N, float* C) {
    if (K <= 0 || K > 256 || K % 16 || N <= 0) return mlp_fail("bad shape", GKC_EINVAL);
    MLPX(hipSetDevice(device));
    float *dA = nullptr, *dB = nullptr, *dC = nullptr;
    unsigned short* dBb = nullptr;
    MLPX(hipMalloc(&dA, 32 * K * 4)); MLPX(hipMalloc(&dB, (size_t)K * N * 4)); MLPX(hipMalloc(&dC, 32 * (size_t)N * 4));
    MLPX(hipMalloc(&dBb, (size_t)K * N * 2));
    MLPX(hipMemcpy(dA, A, 32 * K * 4, hipMemcpyHostToDevice));
    MLPX(hipMemcpy(dB, B, (size_t)K * N * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, dim3((K * N + 255) / 256), dim3(256), 0, 0, dB, dBb, K * N);
    hipLaunchKernelGGL(k_bf16_matmul32, dim3(1), dim3(64), 0, 0, dA, dBb, K, N, dC);
    MLPX(hipGetLastError());
    MLPX(hipMemcpy(C, dC, 32 * (size_t)N * 4, hipMemcpyDeviceToHost));
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC); (void)hipFree(dBb);
    return 0;
}

}  // extern "C"
