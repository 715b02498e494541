// Shared device helpers for the atomorph MI355X kernels (gfx950, wave64).
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amh {

// ---------------------------------------------------------------- points
// am::point word (atomorph.h:275-284): x u16 | y u16 | x_fract u8 | y_fract u8
// | flags u8 | 0.  Fixed-point coordinates are 256*x + fract (< 2^24).
constexpr uint32_t HAS_PIXEL = 1;
constexpr uint32_t HAS_FLUID = 2;

__host__ __device__ __forceinline__ uint32_t pt_x(uint64_t w) { return uint32_t(w) & 0xFFFFu; }
__host__ __device__ __forceinline__ uint32_t pt_y(uint64_t w) { return uint32_t(w >> 16) & 0xFFFFu; }
__host__ __device__ __forceinline__ uint32_t pt_xf(uint64_t w) { return uint32_t(w >> 32) & 0xFFu; }
__host__ __device__ __forceinline__ uint32_t pt_yf(uint64_t w) { return uint32_t(w >> 40) & 0xFFu; }
__host__ __device__ __forceinline__ uint32_t pt_flags(uint64_t w) { return uint32_t(w >> 48) & 0xFFu; }
__host__ __device__ __forceinline__ int32_t pt_fx(uint64_t w) { return int32_t((pt_x(w) << 8) | pt_xf(w)); }
__host__ __device__ __forceinline__ int32_t pt_fy(uint64_t w) { return int32_t((pt_y(w) << 8) | pt_yf(w)); }
__host__ __device__ __forceinline__ uint64_t make_pt(uint32_t x, uint32_t y, uint32_t xf, uint32_t yf,
                                                     uint32_t fl) {
    return uint64_t(x & 0xFFFFu) | (uint64_t(y & 0xFFFFu) << 16) | (uint64_t(xf & 0xFFu) << 32) |
           (uint64_t(yf & 0xFFu) << 40) | (uint64_t(fl & 0xFFu) << 48);
}

// atomorph.h:334-339, exact.
__host__ __device__ __forceinline__ uint64_t pt_dist(uint64_t a, uint64_t b) {
    int64_t dx = int64_t(pt_fx(a)) - pt_fx(b);
    int64_t dy = int64_t(pt_fy(a)) - pt_fy(b);
    return uint64_t(dx * dx) + uint64_t(dy * dy);
}

// ------------------------------------------------------------ randomness
// SplitMix64 finaliser: the counter-based generator behind every random
// choice the kernels make (pairings, rows, duplicate fractions).
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__host__ __device__ __forceinline__ uint32_t ceil_log2(uint32_t n) {
    uint32_t k = 0;
    while ((1u << k) < n) ++k;
    return k;
}

// Balanced 4-round Feistel permutation of [0, 2^k) (k even), cycle-walked
// onto [0, n).
__host__ __device__ inline uint32_t feistel_perm(uint32_t i, uint32_t n, uint64_t key) {
    uint32_t k = ceil_log2(n < 2 ? 2 : n);
    if (k & 1) ++k;
    const uint32_t half = k / 2, mask = (1u << half) - 1u;
    uint32_t v = i;
    do {
        uint32_t L = v >> half, R = v & mask;
        for (uint32_t r = 0; r < 4; ++r) {
            uint32_t f = uint32_t(mix64(key ^ (uint64_t(R) << 8) ^ r)) & mask;
            uint32_t nl = R;
            R = L ^ f;
            L = nl;
        }
        v = (L << half) | R;
    } while (v >= n);
    return v;
}

// One round of the conflict-free schedule over n slots: the "fold" matching
// {i, c - i mod n} for a random c in [0, n).  It pairs floor(n/2) slots when
// c is odd and floor((n-1)/2) when c is even (slots with 2i = c mod n stay
// single).  Every pair {i, j} belongs to exactly one c, so a pair is tried
// with probability 1/n per round and no lane is wasted on padding.
struct RoundPar {
    uint32_t base;    // floor(c / 2)
    uint32_t odd;     // c & 1
    uint32_t npairs;  // pairs in this round
};
__host__ __device__ __forceinline__ RoundPar round_params(uint64_t seed, uint64_t stream, uint64_t round,
                                                           uint32_t n) {
    const uint64_t h = mix64(seed ^ mix64(stream ^ (round * 0xD1B54A32D192ED03ull)));
    const uint32_t c = uint32_t(((h >> 32) * uint64_t(n)) >> 32);
    RoundPar p;
    p.base = c >> 1;
    p.odd = c & 1u;
    p.npairs = p.odd ? n / 2 : (n - 1) / 2;
    return p;
}
// Pair t (0 <= t < npairs) of the round: slots (i, j), i + j = c mod n.
__host__ __device__ __forceinline__ void round_pair(const RoundPar &p, uint32_t n, uint32_t t, uint32_t &i,
                                                   uint32_t &j) {
    uint32_t a = p.base + 1u + t;
    if (a >= n) a -= n;
    const uint32_t d = t + 1u - p.odd;  // base - j
    const uint32_t b = p.base >= d ? p.base - d : p.base + n - d;
    i = a;
    j = b;
}

// Stream id of (chain, pass, group) and the pass's row choice; shared with
// the oracle's replay of the batched schedule.
__host__ __device__ __forceinline__ uint64_t pass_stream(uint64_t chain, uint64_t pass, uint64_t group) {
    return mix64((chain << 32) ^ pass) ^ group;
}
__host__ __device__ __forceinline__ uint32_t pass_row(uint64_t seed, uint64_t chain, uint64_t pass, uint32_t h) {
    return uint32_t(mix64(seed ^ 0x5851F42D4C957F2Dull ^ mix64((chain << 32) ^ pass)) % h);
}

// ------------------------------------------------- spatial keys (kNN)
// Morton (Z-order) key of integer pixel coordinates (x, y < 2^16), and the
// coordinates back from a key.
__host__ __device__ __forceinline__ uint32_t spread16(uint32_t v) {
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
__host__ __device__ __forceinline__ uint32_t squeeze16(uint32_t v) {
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0F0F0F0Fu;
    v = (v | (v >> 4)) & 0x00FF00FFu;
    v = (v | (v >> 8)) & 0x0000FFFFu;
    return v;
}
__host__ __device__ __forceinline__ uint32_t morton2(uint32_t x, uint32_t y) {
    return spread16(x) | (spread16(y) << 1);
}
// The draw of slot l in round r of a kNN-guided pass (stream = pass_stream):
// bits 12..31 its claim priority, bits 32..47 a random other slot, bits
// 48..63 the pick among its candidates.
__host__ __device__ __forceinline__ uint64_t knn_draw(uint64_t seed, uint64_t stream, uint32_t round, uint32_t l) {
    return mix64(seed ^ 0x2545F4914F6CDD1Dull ^ mix64(stream ^ (uint64_t(round) << 32) ^ l));
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// 128-bit accumulate into (lo, hi) with two 64-bit atomics.
__device__ __forceinline__ void atomic_add_u128(unsigned long long *lo, unsigned long long *hi, uint64_t v) {
    unsigned long long old = atomicAdd(lo, (unsigned long long)v);
    if (old + v < old) atomicAdd(hi, 1ull);
}

}  // namespace amh

// Error plumbing shared by the C-ABI entry points (api.hip).
namespace amh {
void set_error(const char *fmt, ...);
int check_hip(hipError_t e, const char *what);

// Optional per-kernel device timing (amh_profile_enable): when on, the
// entry points bracket their kernels with HIP events on the launch stream.
bool profiling();
void prof_begin(const char *kernel, hipStream_t s);
void prof_end(const char *kernel, hipStream_t s);
struct ProfScope {
    const char *k;
    hipStream_t s;
    ProfScope(const char *kernel, hipStream_t st) : k(kernel), s(st) {  // kernel NULL: not timed
        if (k && profiling()) prof_begin(k, s);
    }
    ~ProfScope() {
        if (k && profiling()) prof_end(k, s);
    }
};
// An event pair for one timed launch of `kernel` (false: none available),
// and its hand-back once the launch is enqueued.
bool prof_take(hipEvent_t &start, hipEvent_t &stop);
void prof_put(const char *kernel, hipEvent_t start, hipEvent_t stop);

// Launches one kernel.  With per-kernel timing on, the dispatch itself
// records the start and stop events (hipExtLaunchKernelGGL), so the time is
// the kernel's own execution, as a kernel trace reports it, without the
// gaps that events recorded around the launch add.
template <typename F, typename... Args>
void launch_timed(const char *name, F kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t s, Args... args) {
    hipEvent_t a = nullptr, b = nullptr;
    if (profiling() && prof_take(a, b)) {
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, a, b, 0, args...);
        prof_put(name, a, b);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
    }
}
}  // namespace amh

#define AMH_LAUNCH_CHECK(what)                                              \
    do {                                                                    \
        hipError_t e_ = hipGetLastError();                                  \
        if (e_ != hipSuccess) return amh::check_hip(e_, what);              \
    } while (0)
