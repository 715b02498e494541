// The learned proposal's MLP on the CDNA4 matrix cores.
//
// 8 -> 128 -> 128 -> 128 -> 128 -> 2, exact GELU after each hidden layer.
// Parameters are kept as FP32 masters (Adam state beside them) and as a
// bfloat16 copy that the matrix products read; activations are FP32 in HBM
// and rounded to bfloat16 when a product stages them; every product
// accumulates in FP32 (v_mfma_f32_32x32x16_bf16).
//
// All products run through one tiled kernel, C = A . B with free strides on
// both operands (so transposes cost nothing) and a fused epilogue:
//   forward      Z_l = A_{l-1} W_l + b_l, A_l = gelu(Z_l)
//   backward     dZ_{l-1} = (dZ_l W_l^T) * gelu'(Z_{l-1})
//   weights      dW_l = A_{l-1}^T dZ_l (split over the batch, FP32 atomics)
// The packed parameter layout (W row-major [in][out] then b, per layer) is
// the one the oracle and the host API use.
#include "common.h"
#include "../../include/atomorph_hip.h"

#include <cmath>
#include <cstring>
#include <vector>

using amh::check_hip;
using amh::set_error;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kIn = 8, kH = 128, kOut = 2;
constexpr int kDims[6] = {kIn, kH, kH, kH, kH, kOut};
constexpr int kParams = 50946;
constexpr int kLayers = 5;

__host__ __device__ constexpr size_t w_off(int l) {
    return l == 0 ? 0 : w_off(l - 1) + size_t(kDims[l - 1]) * kDims[l] + kDims[l];
}
__host__ __device__ constexpr size_t b_off(int l) { return w_off(l) + size_t(kDims[l]) * kDims[l + 1]; }
static_assert(b_off(4) + kOut == kParams, "parameter layout");

// FP32 -> bfloat16 bits, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t to_bf16(float f) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f)); }
// erf by Abramowitz & Stegun, Handbook of Mathematical Functions (1964),
// formula 7.1.26: for x >= 0, erf x = 1 - (a1 t + a2 t^2 + a3 t^3 + a4 t^4 +
// a5 t^5) exp(-x^2) with t = 1 / (1 + p x), |error| <= 1.5e-7; odd in x.
// About a third of the instructions of erff, and the GELUs (4 x 128 per
// row) are what bounds the fused forward; the activations are rounded to
// bfloat16 (2^-8 relative) right after.
__device__ __forceinline__ float erf_as(float x) {
    const float ax = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, ax, 1.0f));
    float poly = fmaf(t, 1.061405429f, -1.453152027f);
    poly = fmaf(t, poly, 1.421413741f);
    poly = fmaf(t, poly, -0.284496736f);
    poly = fmaf(t, poly, 0.254829592f);
    poly *= t;
    const float y = fmaf(-poly, __expf(-ax * ax), 1.0f);
    return copysignf(y, x);
}
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erf_as(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
    const float cdf = 0.5f * (1.0f + erf_as(z * 0.70710678118654752f));
    const float pdf = __expf(-0.5f * z * z) * 0.39894228040143268f;
    return cdf + z * pdf;
}

template <class T> __device__ __forceinline__ uint16_t load_bf16(const T *p);
template <> __device__ __forceinline__ uint16_t load_bf16<float>(const float *p) { return to_bf16(*p); }
template <> __device__ __forceinline__ uint16_t load_bf16<uint16_t>(const uint16_t *p) { return *p; }

// ------------------------------------------------------------------ GEMM
// Block tile 64 x 64 (2 x 2 waves of 32 x 32), K staged 32 at a time in LDS
// as bfloat16: As[m][k] and Bs[n][k] so each lane's MFMA operand is 16
// contiguous bytes.  MFMA 32x32x16 operands: lane l holds row/column l % 32
// and k = 8 * (l / 32) + 0..7; the accumulator's register r of lane l is
// row 8 * (r / 4) + 4 * (l / 32) + r % 4, column l % 32.
constexpr int kBM = 64, kBN = 64, kBK = 32, kPadK = kBK + 8;

struct Gemm {
    int M, N, K;
    int64_t sam, sak;  // A(m, k) = a[m * sam + k * sak]
    int64_t sbk, sbn;  // B(k, n) = b[k * sbk + n * sbn]
    int ksplit;        // K per blockIdx.z
};

template <class TB, class Epi>
__global__ __launch_bounds__(256) void k_gemm(const float *__restrict__ a, const TB *__restrict__ b, Gemm g, Epi epi) {
    __shared__ __attribute__((aligned(16))) uint16_t As[kBM][kPadK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kBN][kPadK];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.x * kBM, n0 = blockIdx.y * kBN;
    const int kb = blockIdx.z * g.ksplit, ke = min(g.K, kb + g.ksplit);
    floatx16 acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    const int sr = t >> 2, sk = (t & 3) * 8;  // staging: row sr, k sk..sk+7
    for (int k0 = kb; k0 < ke; k0 += kBK) {
        {
            const int m = m0 + sr;
            alignas(16) uint16_t va[8], vb[8];
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + sk + j;
                va[j] = (m < g.M && k < ke) ? load_bf16(a + m * g.sam + k * g.sak) : uint16_t(0);
                const int n = n0 + sr;
                vb[j] = (n < g.N && k < ke) ? load_bf16(b + k * g.sbk + n * g.sbn) : uint16_t(0);
            }
            *reinterpret_cast<uint4 *>(&As[sr][sk]) = *reinterpret_cast<const uint4 *>(va);
            *reinterpret_cast<uint4 *>(&Bs[sr][sk]) = *reinterpret_cast<const uint4 *>(vb);
        }
        __syncthreads();
        for (int ks = 0; ks < kBK; ks += 16) {
            const int kk = ks + 8 * (lane >> 5);
            const bf16x8 fa = *reinterpret_cast<const bf16x8 *>(&As[wm * 32 + (lane & 31)][kk]);
            const bf16x8 fb = *reinterpret_cast<const bf16x8 *>(&Bs[wn * 32 + (lane & 31)][kk]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int col = n0 + wn * 32 + (lane & 31);
    for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (row < g.M && col < g.N) epi(row, col, acc[r]);
    }
}

// epilogues
struct EpiStore {  // c[row][col] = v
    float *c;
    int ld;
    __device__ void operator()(int row, int col, float v) const { c[int64_t(row) * ld + col] = v; }
};
struct EpiAdd {  // c[row][col] += v (split-K partial sums)
    float *c;
    int ld;
    __device__ void operator()(int row, int col, float v) const { atomicAdd(&c[int64_t(row) * ld + col], v); }
};
struct EpiBiasGelu {  // z = v + bias; act = gelu(z) when act != NULL
    float *z, *act;
    const float *bias;
    int ld;
    __device__ void operator()(int row, int col, float v) const {
        const float zz = v + bias[col];
        z[int64_t(row) * ld + col] = zz;
        if (act) act[int64_t(row) * ld + col] = gelu(zz);
    }
};
struct EpiGeluD {  // dz = v * gelu'(zref)
    float *dz;
    const float *zref;
    int ld;
    __device__ void operator()(int row, int col, float v) const {
        const int64_t k = int64_t(row) * ld + col;
        dz[k] = v * gelu_d(zref[k]);
    }
};

template <class TB, class Epi>
void gemm(const float *a, const TB *b, const Gemm &g, const Epi &epi, int splits, hipStream_t s) {
    dim3 grid(uint32_t((g.M + kBM - 1) / kBM), uint32_t((g.N + kBN - 1) / kBN), uint32_t(splits));
    hipLaunchKernelGGL((k_gemm<TB, Epi>), grid, dim3(256), 0, s, a, b, g, epi);
}

// --------------------------------------------------------- elementwise
__global__ void k_to_bf16(const float *__restrict__ p, uint16_t *__restrict__ q, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) q[i] = to_bf16(p[i]);
}

__global__ void k_fill(float *p, int64_t n, float v) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// dout = 2 (out - t) / n; loss += sum (out - t)^2 / n
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ out, const float *__restrict__ t, int n,
                                            float *__restrict__ dout, float *__restrict__ loss) {
    __shared__ float part[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (i < n * kOut) {
        const float d = out[i] - t[i];
        dout[i] = 2.0f * d / float(n);
        l = d * d / float(n);
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0 && loss) atomicAdd(loss, part[0] + part[1] + part[2] + part[3]);
}

// db[c] += column sums of dz[n][ld] over a chunk of rows; grid (chunks, cols)
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ dz, int n, int ld, float *__restrict__ db) {
    __shared__ float part[4];
    const int c = blockIdx.y;
    float s = 0.0f;
    for (int r = blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256) s += dz[int64_t(r) * ld + c];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&db[c], part[0] + part[1] + part[2] + part[3]);
}

// Adam (Kingma & Ba 2015) on the FP32 masters; refreshes the bfloat16 copy.
__global__ void k_adam(float *__restrict__ p, uint16_t *__restrict__ pb, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    pb[i] = to_bf16(pi);
}

// ------------------------------------------------------- fused forward
// Inference in one kernel with the activations kept in registers: each wave
// computes 32 rows at a time as Z^T = W^T A^T, so the accumulator of one
// layer (lane = row l % 32, registers = output features) is already the
// next layer's B operand once its K order is permuted to match:
//   MFMA k = 16 ks + 8 (l / 32) + j  <->  feature feat(ks, l / 32, j)
// and the weights are staged into LDS in that order (A operands, 16 bytes
// per lane).  Layer 1 pads K 8 -> 16 with zeros, the output layer pads its
// 2 rows to 32.  LDS: 4 + 3 x 32 + 8 KB of weights, 2 KB of biases.
__host__ __device__ constexpr int feat(int ks, int h, int j) {
    return 32 * (ks / 2) + 16 * (ks % 2) + 8 * (j / 4) + 4 * h + (j % 4);
}
struct FusedLds {
    bf16x8 w1[4][64];
    bf16x8 wh[3][4][8][64];
    bf16x8 w5[8][64];
    float b[4][kH];
    float b5[kOut];
};
constexpr size_t kFusedLds = sizeof(FusedLds);
static_assert(kFusedLds % 16 == 0, "LDS image copied in 16-byte units");

__device__ __forceinline__ bf16x8 pack8(const uint16_t *v) {
    bf16x8 r;
    __builtin_memcpy(&r, v, 16);
    return r;
}

// The LDS image of the weights in MFMA operand order (built once per
// parameter change into global memory; the fused kernel copies it).
__global__ __launch_bounds__(256) void k_fused_stage(const uint16_t *__restrict__ pb, const float *__restrict__ p,
                                                    FusedLds *__restrict__ S) {
    const int t = blockIdx.x * 256 + threadIdx.x, nt = gridDim.x * 256;
    for (int e = t; e < 4 * 64; e += nt) {
        const int mb = e / 64, l = e % 64, lh = l >> 5, lm = l & 31;
        uint16_t v[8];
        for (int j = 0; j < 8; ++j) v[j] = lh == 0 ? pb[w_off(0) + j * kH + mb * 32 + lm] : uint16_t(0);
        S->w1[mb][l] = pack8(v);
    }
    for (int e = t; e < 3 * 4 * 8 * 64; e += nt) {
        const int li = e / (4 * 8 * 64), mb = (e / (8 * 64)) % 4, ks = (e / 64) % 8, l = e % 64;
        const int lh = l >> 5, lm = l & 31;
        uint16_t v[8];
        for (int j = 0; j < 8; ++j) v[j] = pb[w_off(li + 1) + size_t(feat(ks, lh, j)) * kH + mb * 32 + lm];
        S->wh[li][mb][ks][l] = pack8(v);
    }
    for (int e = t; e < 8 * 64; e += nt) {
        const int ks = e / 64, l = e % 64, lh = l >> 5, lm = l & 31;
        uint16_t v[8];
        for (int j = 0; j < 8; ++j) v[j] = lm < kOut ? pb[w_off(4) + size_t(feat(ks, lh, j)) * kOut + lm] : uint16_t(0);
        S->w5[ks][l] = pack8(v);
    }
    for (int e = t; e < 4 * kH; e += nt) S->b[e / kH][e % kH] = p[b_off(e / kH) + e % kH];
    if (t < kOut) S->b5[t] = p[b_off(4) + t];
}

constexpr int kFusedThreads = 1024;  // 16 waves (4 per SIMD, <= 128 registers each) share one copy of the weights

// Bias + GELU of one 32-feature output block (accumulator layout: feature
// 8*(r/4) + 4*h + r%4 of block mb) -> the two B operands it feeds.
__device__ __forceinline__ void gelu_block(const floatx16 &acc, const float *__restrict__ bias, int mb, int h,
                                           bf16x8 &lo, bf16x8 &hi) {
    uint16_t v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = to_bf16(gelu(acc[r] + bias[mb * 32 + 8 * (r >> 2) + 4 * h + (r & 3)]));
    lo = pack8(v);
    hi = pack8(v + 8);
}

__global__ __launch_bounds__(kFusedThreads) void k_mlp_fused(const float *__restrict__ x, int64_t n,
                                                  const FusedLds *__restrict__ image, float *__restrict__ out) {
    extern __shared__ __attribute__((aligned(16))) char lds_raw[];
    FusedLds &S = *reinterpret_cast<FusedLds *>(lds_raw);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, h = lane >> 5, m = lane & 31;
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(image);
        uint4 *dst = reinterpret_cast<uint4 *>(lds_raw);
        for (size_t e = t; e < kFusedLds / 16; e += kFusedThreads) dst[e] = src[e];
    }
    __syncthreads();

    const int64_t tiles = (n + 31) / 32;
    constexpr int kWaves = kFusedThreads / 64;
    for (int64_t tile = int64_t(blockIdx.x) * kWaves + wave; tile < tiles; tile += int64_t(gridDim.x) * kWaves) {
        const int64_t row = tile * 32 + m;
        uint16_t xin[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) xin[j] = (h == 0 && row < n) ? to_bf16(x[row * kIn + j]) : uint16_t(0);
        const bf16x8 xa = pack8(xin);
        // Layer 1 (K = 8 in one MFMA), two output blocks at a time: each
        // pair is finished (bias, GELU) before the next starts, so only two
        // accumulators are live.
        bf16x8 B[8], Bn[8];
#pragma unroll
        for (int mp = 0; mp < 2; ++mp) {
            floatx16 a0, a1;
#pragma unroll
            for (int r = 0; r < 16; ++r) a0[r] = a1[r] = 0.0f;
            a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(S.w1[2 * mp][lane], xa, a0, 0, 0, 0);
            a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(S.w1[2 * mp + 1][lane], xa, a1, 0, 0, 0);
            gelu_block(a0, S.b[0], 2 * mp, h, B[4 * mp], B[4 * mp + 1]);
            gelu_block(a1, S.b[0], 2 * mp + 1, h, B[4 * mp + 2], B[4 * mp + 3]);
            __builtin_amdgcn_sched_barrier(0);  // keep the pairs apart: bounded register use
        }
        // Hidden layers 2-4: B (this layer's input) -> Bn, then swap.
#pragma unroll 1
        for (int li = 0; li < 3; ++li) {
#pragma unroll
            for (int mp = 0; mp < 2; ++mp) {
                floatx16 a0, a1;
#pragma unroll
                for (int r = 0; r < 16; ++r) a0[r] = a1[r] = 0.0f;
#pragma unroll
                for (int ks = 0; ks < 8; ++ks) {
                    a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(S.wh[li][2 * mp][ks][lane], B[ks], a0, 0, 0, 0);
                    a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(S.wh[li][2 * mp + 1][ks][lane], B[ks], a1, 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
                gelu_block(a0, S.b[li + 1], 2 * mp, h, Bn[4 * mp], Bn[4 * mp + 1]);
                gelu_block(a1, S.b[li + 1], 2 * mp + 1, h, Bn[4 * mp + 2], Bn[4 * mp + 3]);
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int ks = 0; ks < 8; ++ks) B[ks] = Bn[ks];
        }
        floatx16 o;
#pragma unroll
        for (int r = 0; r < 16; ++r) o[r] = 0.0f;
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) o = __builtin_amdgcn_mfma_f32_32x32x16_bf16(S.w5[ks][lane], B[ks], o, 0, 0, 0);
        if (h == 0 && row < n) {  // rows 0 and 1 of the output block: registers 0 and 1
            out[row * kOut + 0] = o[0] + S.b5[0];
            out[row * kOut + 1] = o[1] + S.b5[1];
        }
    }
}

}  // namespace

// ------------------------------------------------------------- handle
struct amh_mlp_s {
    int32_t batch = 0;
    void *buf = nullptr;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // [kParams]
    uint16_t *pb = nullptr;                                         // bfloat16 copy
    FusedLds *image = nullptr;  // weights in the fused forward's LDS order
    bool image_stale = true;
    float *z[kLayers] = {};    // [batch][dims[l+1]] pre-activations
    float *act[kLayers] = {};  // [batch][dims[l]] inputs of layer l (act[0] unused: the caller's x)
    float *dz[2] = {};         // [batch][128] ping-pong gradients
    float *loss = nullptr;     // [1]
    int64_t step = 0;
};

namespace {
int g_cus = 0;
int g_fused = -1;  // -1 unknown, 0 no, 1 yes
// The fused forward needs kFusedLds bytes of LDS in one workgroup.
bool fused_ok() {
    if (g_fused >= 0) return g_fused == 1;
    int dev = 0, lds = 0, cus = 0;
    g_fused = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return false;
    if (size_t(lds) < kFusedLds) return false;
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mlp_fused), hipFuncAttributeMaxDynamicSharedMemorySize,
                            int(kFusedLds)) != hipSuccess)
        return false;
    g_cus = cus;
    g_fused = 1;
    return true;
}

size_t mlp_bytes(int32_t batch) {
    const size_t B = size_t(batch);
    return 4 * kParams * 4 + kParams * 2 + kFusedLds + B * kH * 4 * (kLayers + kLayers + 2) + 64 * 256;
}

// forward of rows [0, n) of x (n <= batch) into M->z[4] (n x 2)
void forward_batch(amh_mlp_s *M, const float *x, int n, hipStream_t s) {
    const float *in = x;
    for (int l = 0; l < kLayers; ++l) {
        const int din = kDims[l], dout = kDims[l + 1];
        Gemm g{n, dout, din, din, 1, dout, 1, din};
        EpiBiasGelu e{M->z[l], l < kLayers - 1 ? M->act[l + 1] : nullptr, M->p + b_off(l), dout};
        gemm(in, M->pb + w_off(l), g, e, 1, s);
        if (l < kLayers - 1) in = M->act[l + 1];
    }
}
}  // namespace

extern "C" int amh_mlp_create(void **mlp, int32_t batch_max, void *stream) {
    if (!mlp || batch_max <= 0 || batch_max > (1 << 22)) {
        set_error("amh_mlp_create: invalid argument");
        return AMH_E_ARG;
    }
    amh_mlp_s *M = new amh_mlp_s;
    M->batch = batch_max;
    const size_t bytes = mlp_bytes(batch_max);
    if (int r = check_hip(hipMalloc(&M->buf, bytes), "amh_mlp_create")) {
        delete M;
        return r;
    }
    char *q = static_cast<char *>(M->buf);
    auto take = [&](size_t b) { char *r = q; q += (b + 255) & ~size_t(255); return r; };
    M->p = reinterpret_cast<float *>(take(kParams * 4));
    M->g = reinterpret_cast<float *>(take(kParams * 4));
    M->m = reinterpret_cast<float *>(take(kParams * 4));
    M->v = reinterpret_cast<float *>(take(kParams * 4));
    M->pb = reinterpret_cast<uint16_t *>(take(kParams * 2));
    M->image = reinterpret_cast<FusedLds *>(take(kFusedLds));
    const size_t B = size_t(batch_max);
    for (int l = 0; l < kLayers; ++l) M->z[l] = reinterpret_cast<float *>(take(B * kH * 4));
    for (int l = 0; l < kLayers; ++l) M->act[l] = reinterpret_cast<float *>(take(B * kH * 4));
    M->dz[0] = reinterpret_cast<float *>(take(B * kH * 4));
    M->dz[1] = reinterpret_cast<float *>(take(B * kH * 4));
    M->loss = reinterpret_cast<float *>(take(256));
    if (int r = check_hip(hipMemsetAsync(M->buf, 0, bytes, static_cast<hipStream_t>(stream)), "amh_mlp_create")) {
        (void)hipFree(M->buf);
        delete M;
        return r;
    }
    *mlp = M;
    return AMH_OK;
}

extern "C" int amh_mlp_destroy(void *mlp) {
    amh_mlp_s *M = static_cast<amh_mlp_s *>(mlp);
    if (!M) return AMH_OK;
    if (M->buf) (void)hipFree(M->buf);
    delete M;
    return AMH_OK;
}

extern "C" int amh_mlp_set_params(void *mlp, const float *host, void *stream) {
    amh_mlp_s *M = static_cast<amh_mlp_s *>(mlp);
    if (!M || !host) { set_error("amh_mlp_set_params: NULL"); return AMH_E_ARG; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (int r = check_hip(hipMemcpyAsync(M->p, host, kParams * 4, hipMemcpyHostToDevice, s), "set_params")) return r;
    hipLaunchKernelGGL(k_to_bf16, dim3((kParams + 255) / 256), dim3(256), 0, s, M->p, M->pb, kParams);
    hipLaunchKernelGGL(k_fill, dim3((kParams + 255) / 256), dim3(256), 0, s, M->m, int64_t(kParams), 0.0f);
    hipLaunchKernelGGL(k_fill, dim3((kParams + 255) / 256), dim3(256), 0, s, M->v, int64_t(kParams), 0.0f);
    M->step = 0;
    M->image_stale = true;
    if (fused_ok()) {  // the fused forward's weight image, now rather than at the first forward
        hipLaunchKernelGGL(k_fused_stage, dim3(64), dim3(256), 0, s, M->pb, M->p, M->image);
        M->image_stale = false;
    }
    AMH_LAUNCH_CHECK("amh_mlp_set_params");
    return check_hip(hipStreamSynchronize(s), "set_params");
}

extern "C" int amh_mlp_get_params(void *mlp, float *host, void *stream) {
    amh_mlp_s *M = static_cast<amh_mlp_s *>(mlp);
    if (!M || !host) { set_error("amh_mlp_get_params: NULL"); return AMH_E_ARG; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (int r = check_hip(hipMemcpyAsync(host, M->p, kParams * 4, hipMemcpyDeviceToHost, s), "get_params")) return r;
    return check_hip(hipStreamSynchronize(s), "get_params");
}

// Inference: out[n][2] for x[n][8], in batches of the handle's size.
extern "C" int amh_mlp_forward(void *mlp, const float *x, int64_t n, float *out, void *stream) {
    amh_mlp_s *M = static_cast<amh_mlp_s *>(mlp);
    if (!M || (n > 0 && (!x || !out)) || n < 0) { set_error("amh_mlp_forward: invalid argument"); return AMH_E_ARG; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    amh::ProfScope ps_("mlp_forward", s);
    if (n == 0) return AMH_OK;
    if (fused_ok()) {
        const int64_t tiles = (n + 31) / 32;
        const uint32_t grid = uint32_t(std::min<int64_t>(g_cus, (tiles + kFusedThreads / 64 - 1) / (kFusedThreads / 64)));
        if (M->image_stale) {
            hipLaunchKernelGGL(k_fused_stage, dim3(64), dim3(256), 0, s, M->pb, M->p, M->image);
            M->image_stale = false;
        }
        hipLaunchKernelGGL(k_mlp_fused, dim3(grid), dim3(kFusedThreads), kFusedLds, s, x, n, M->image, out);
        AMH_LAUNCH_CHECK("amh_mlp_forward");
        return AMH_OK;
    }
    for (int64_t i = 0; i < n; i += M->batch) {
        const int c = int(std::min<int64_t>(M->batch, n - i));
        forward_batch(M, x + i * kIn, c, s);
        if (int r = check_hip(hipMemcpyAsync(out + i * kOut, M->z[kLayers - 1], size_t(c) * kOut * 4,
                                             hipMemcpyDeviceToDevice, s),
                              "amh_mlp_forward"))
            return r;
    }
    AMH_LAUNCH_CHECK("amh_mlp_forward");
    return AMH_OK;
}

// Gradient of mean_s sum_o (out - t)^2 over n <= batch samples into the
// handle's gradient buffer (and, if grad != NULL, a copy there); the loss
// is added to *loss_dev when it is not NULL.
extern "C" int amh_mlp_gradients(void *mlp, const float *x, const float *t, int32_t n, float *grad, float *loss_dev,
                                 void *stream) {
    amh_mlp_s *M = static_cast<amh_mlp_s *>(mlp);
    if (!M || !x || !t || n <= 0 || n > M->batch) { set_error("amh_mlp_gradients: invalid argument"); return AMH_E_ARG; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    amh::ProfScope ps_("mlp_gradients", s);
    forward_batch(M, x, n, s);
    hipLaunchKernelGGL(k_fill, dim3((kParams + 255) / 256), dim3(256), 0, s, M->g, int64_t(kParams), 0.0f);
    float *dcur = M->dz[0], *dnext = M->dz[1];
    hipLaunchKernelGGL(k_mse, dim3((n * kOut + 255) / 256), dim3(256), 0, s, M->z[kLayers - 1], t, n, dcur, loss_dev);
    // the weight gradients sum over the batch: split it over blocks (~512
    // rows each) that add their partial products
    const int splits = std::max(1, std::min(64, n / 512));
    const int klen = ((n + splits - 1) / splits + kBK - 1) / kBK * kBK;
    for (int l = kLayers - 1; l >= 0; --l) {
        const int din = kDims[l], dout = kDims[l + 1];
        const float *ain = l == 0 ? x : M->act[l];
        // dW_l = A_{l-1}^T dz_l: M = din, N = dout, K = n
        Gemm gw{din, dout, n, 1, din, dout, 1, klen};
        gemm(ain, static_cast<const float *>(dcur), gw, EpiAdd{M->g + w_off(l), dout}, (n + klen - 1) / klen, s);
        hipLaunchKernelGGL(k_colsum, dim3(uint32_t(std::min(64, (n + 255) / 256)), uint32_t(dout)), dim3(256), 0, s,
                           dcur, n, dout, M->g + b_off(l));
        if (l == 0) break;
        // dz_{l-1} = (dz_l W_l^T) * gelu'(z_{l-1}): M = n, N = din, K = dout
        Gemm gd{n, din, dout, dout, 1, 1, dout, dout};
        gemm(dcur, M->pb + w_off(l), gd, EpiGeluD{dnext, M->z[l - 1], din}, 1, s);
        std::swap(dcur, dnext);
    }
    if (grad && grad != M->g) {
        if (int r = check_hip(hipMemcpyAsync(grad, M->g, kParams * 4, hipMemcpyDeviceToDevice, s), "gradients"))
            return r;
    }
    AMH_LAUNCH_CHECK("amh_mlp_gradients");
    return AMH_OK;
}

// One Adam step with the handle's gradient buffer (from amh_mlp_gradients).
extern "C" int amh_mlp_adam_step(void *mlp, float lr, void *stream) {
    amh_mlp_s *M = static_cast<amh_mlp_s *>(mlp);
    if (!M || !(lr > 0.0f)) { set_error("amh_mlp_adam_step: invalid argument"); return AMH_E_ARG; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    M->step += 1;
    M->image_stale = true;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, float(M->step)), c2 = 1.0f - std::pow(b2, float(M->step));
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, M->p, M->pb, M->g, M->m, M->v, kParams,
                       lr, b1, b2, 1e-8f, c1, c2);
    AMH_LAUNCH_CHECK("amh_mlp_adam_step");
    return AMH_OK;
}

// Plain product for tests of the matrix-core kernel: c[M][N] = a[M][K] b[K][N]
// (FP32 row-major in, bfloat16 rounding of both operands, FP32 out).
extern "C" int amh_mlp_gemm_test(const float *a, const float *b, float *c, int32_t M, int32_t N, int32_t K,
                                 void *stream) {
    if (!a || !b || !c || M <= 0 || N <= 0 || K <= 0) { set_error("amh_mlp_gemm_test: invalid argument"); return AMH_E_ARG; }
    Gemm g{M, N, K, K, 1, N, 1, ((K + kBK - 1) / kBK) * kBK};
    gemm(a, b, g, EpiStore{c, N}, 1, static_cast<hipStream_t>(stream));
    AMH_LAUNCH_CHECK("amh_mlp_gemm_test");
    return AMH_OK;
}
