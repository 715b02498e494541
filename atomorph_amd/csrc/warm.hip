// Learned warm start of the transport refinement.
//
// The reference starts every chain row in the row-major pixel order its
// blobs were built in (thread.cpp:740-891) and leaves the whole assignment
// to the pair-swap descent.  Here, before the descent, row y+1 of each chain
// is reordered so that slot i receives the point of key frame y+1 whose
// rank in raster order (y, then x, at 1/16 px) equals the rank of the
// MLP's prediction for slot i's row-y atom:
//   features  per atom of row y: x/W, y/H, RGBA/255 of its key-frame pixel,
//             (x - cx)/s, (y - cy)/s with (cx, cy) the chain row's centroid
//             and s its RMS radius + 1 px;
//   predict   MLP -> (u, v), target position (cx' + s' u, cy' + s' v) in the
//             statistics of row y+1;
//   match     sort (chain, raster key of the prediction) and (chain, raster
//             key of the row y+1 point) with an LSD radix sort and pair them
//             rank by rank, so every row stays a permutation of its points.
// The reference's start pairs the blobs' pixels rank by rank in the same
// raster order; a prediction that keeps each atom's position relative to
// its blob (scaled) reproduces that start, so what the network learns is
// what changes it.  Rows are processed y = 0 .. h-2 in order; row 0 keeps
// its order.  The same feature kernel writes training targets (the row y+1
// position of each slot, normalised the same way) from converged chains.
#include "common.h"
#include "../../include/atomorph_hip.h"
#include "scan.h"
#include "shade.h"

#include <algorithm>

using namespace amh;
using namespace amh::shade;

namespace {

constexpr int kT = 256;

struct RowStat {
    double cx, cy, s;
};

__device__ __forceinline__ double pt_xd(uint64_t w) { return double(pt_fx(w)) * (1.0 / 256.0); }
__device__ __forceinline__ double pt_yd(uint64_t w) { return double(pt_fy(w)) * (1.0 / 256.0); }

// Centroid and RMS radius (+1 px) of every (row, chain): one block per pair.
__global__ __launch_bounds__(kT) void k_row_stats(amh_render R, RowStat *__restrict__ st) {
    __shared__ double part[4][4];
    const int c = blockIdx.x, y = blockIdx.y;
    const int64_t off = R.chain_off_dev[c];
    const int w = R.chain_w_dev[c];
    const uint64_t *row = R.pts_dev + int64_t(y) * R.wtot + off;
    double sx = 0, sy = 0, sxx = 0, syy = 0;
    for (int i = threadIdx.x; i < w; i += kT) {
        const double x = pt_xd(row[i]), v = pt_yd(row[i]);
        sx += x;
        sy += v;
        sxx += x * x;
        syy += v * v;
    }
    double q[4] = {sx, sy, sxx, syy};
    for (int k = 0; k < 4; ++k)
        for (int o = 32; o > 0; o >>= 1) q[k] += __shfl_down(q[k], o, 64);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 4; ++k) part[threadIdx.x >> 6][k] = q[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[4] = {0, 0, 0, 0};
        for (int wv = 0; wv < kT / 64; ++wv)
            for (int k = 0; k < 4; ++k) t[k] += part[wv][k];
        const double n = w > 0 ? double(w) : 1.0;
        const double cx = t[0] / n, cy = t[1] / n;
        const double var = fmax(t[2] / n - cx * cx + t[3] / n - cy * cy, 0.0);
        st[int64_t(y) * R.nchains + c] = RowStat{cx, cy, sqrt(var) + 1.0};
    }
}

// Features of row y (x[wtot][8]) and, when t != NULL, the targets: the
// normalised positions the same slots hold in row y+1 (t[wtot][2]).
__global__ __launch_bounds__(kT) void k_features(amh_render R, int y, const RowStat *__restrict__ st,
                                                float *__restrict__ x, float *__restrict__ t) {
    const int64_t a = int64_t(blockIdx.x) * kT + threadIdx.x;
    const int c = block_chain(R.chain_off_dev, R.nchains, a);
    if (a >= R.wtot) return;
    const int yn = (y + 1) % R.nkeys;
    const uint64_t p = R.pts_dev[int64_t(y) * R.wtot + a];
    const RowStat s = st[int64_t(y) * R.nchains + c];
    const double px = pt_xd(p), py = pt_yd(p);
    uint32_t col = 0;
    if (pt_flags(p) & HAS_PIXEL) col = key_px(R.keys_dev, y, R.width, R.height, pt_x(p), pt_y(p));
    float *f = x + a * 8;
    f[0] = float(px / R.width);
    f[1] = float(py / R.height);
    for (int k = 0; k < 4; ++k) f[2 + k] = float(ch(col, k)) * (1.0f / 255.0f);
    f[6] = float((px - s.cx) / s.s);
    f[7] = float((py - s.cy) / s.s);
    if (t) {
        const uint64_t q = R.pts_dev[int64_t(yn) * R.wtot + a];
        const RowStat sn = st[int64_t(yn) * R.nchains + c];
        t[a * 2 + 0] = float((pt_xd(q) - sn.cx) / sn.s);
        t[a * 2 + 1] = float((pt_yd(q) - sn.cy) / sn.s);
    }
}

// Raster key of a position at 1/16 px: set bit, chain, y, x, with cb chain
// bits and cx bits per coordinate (positions clamped to the frame's range).
__device__ __forceinline__ uint64_t raster_key(uint64_t set, uint64_t chain, double x, double y, int cb, int cx) {
    const double lim = double((1ull << cx) - 1);
    const uint64_t fx = uint64_t(fmin(fmax(rint(x * 16.0), 0.0), lim));
    const uint64_t fy = uint64_t(fmin(fmax(rint(y * 16.0), 0.0), lim));
    return (set << (cb + 2 * cx)) | (chain << (2 * cx)) | (fy << cx) | fx;
}

// One sort for both lists: keys [0, n) are the predictions for row y+1
// (set 0), keys [n, 2n) the points row y+1 holds now (set 1); values are
// the slots.  Sorted, set 0 and set 1 each list chain by chain (each chain
// has its width in both), so rank r of chain c sits at the same offset in
// both halves.
__global__ __launch_bounds__(kT) void k_match_keys(amh_render R, int yn, const RowStat *__restrict__ st,
                                                  const float *__restrict__ pred, int cb, int cx,
                                                  uint64_t *__restrict__ key, uint32_t *__restrict__ val) {
    const int64_t a = int64_t(blockIdx.x) * kT + threadIdx.x;
    const int c = block_chain(R.chain_off_dev, R.nchains, a);
    if (a >= R.wtot) return;
    const RowStat s = st[int64_t(yn) * R.nchains + c];
    const double ux = s.cx + s.s * double(pred[a * 2 + 0]), uy = s.cy + s.s * double(pred[a * 2 + 1]);
    const uint64_t q = R.pts_dev[int64_t(yn) * R.wtot + a];
    key[a] = raster_key(0, uint64_t(c), ux, uy, cb, cx);
    const uint64_t lim = (1ull << cx) - 1;
    key[R.wtot + a] = (1ull << (cb + 2 * cx)) | (uint64_t(c) << (2 * cx)) |
                      (min(uint64_t(pt_fy(q) >> 4), lim) << cx) | min(uint64_t(pt_fx(q) >> 4), lim);
    val[a] = uint32_t(a);
    val[R.wtot + a] = uint32_t(a);
}

// Rank r of both sorted halves pairs up inside one chain: the slot with the
// r-th prediction takes the r-th point.
__global__ void k_match_apply(int64_t n, const uint32_t *__restrict__ sorted, const uint64_t *__restrict__ row,
                              uint64_t *__restrict__ out) {
    const int64_t r = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (r < n) out[sorted[r]] = row[sorted[n + r]];
}

// ------------------------------------------------------- radix sort
// LSD radix sort of (u64 key, u32 value) pairs, 8 bits per pass; a block
// covers kRsTile elements; the scatter is stable (rank within a round from
// per-digit lane masks built from eight ballots).
constexpr int kRsTile = kT * 8;

__global__ __launch_bounds__(kT) void k_rs_hist(const uint64_t *__restrict__ key, int64_t n, int shift,
                                               int nblk, int *__restrict__ hist) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = int64_t(blockIdx.x) * kRsTile;
    for (int j = 0; j < 8; ++j) {
        const int64_t e = base + j * kT + threadIdx.x;
        if (e < n) atomicAdd(&h[int((key[e] >> shift) & 255u)], 1);
    }
    __syncthreads();
    hist[int64_t(threadIdx.x) * nblk + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kT) void k_rs_scatter(const uint64_t *__restrict__ key, const uint32_t *__restrict__ val,
                                                  int64_t n, int shift, int nblk, const int *__restrict__ off,
                                                  uint64_t *__restrict__ okey, uint32_t *__restrict__ oval) {
    __shared__ int base[256];
    __shared__ int cnt[kT / 64][256];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    base[t] = off[int64_t(t) * nblk + blockIdx.x];
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t b0 = int64_t(blockIdx.x) * kRsTile;
    for (int j = 0; j < 8; ++j) {
        for (int w = 0; w < kT / 64; ++w) cnt[w][t] = 0;
        __syncthreads();
        const int64_t e = b0 + j * kT + t;
        const bool ok = e < n;
        const uint64_t k = ok ? key[e] : 0;
        const uint32_t v = ok ? val[e] : 0;
        const int d = int((k >> shift) & 255u);
        uint64_t same = __ballot(ok);
        for (int bit = 0; bit < 8; ++bit) {
            const uint64_t bb = __ballot((d >> bit) & 1);
            same &= ((d >> bit) & 1) ? bb : ~bb;
        }
        const int rank = int(__popcll(same & below));
        if (ok && rank == 0) cnt[wave][d] = int(__popcll(same));
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; ++w) before += cnt[w][d];
        if (ok) {
            const int pos = base[d] + before + rank;
            okey[pos] = k;
            oval[pos] = v;
        }
        __syncthreads();
        int tot = 0;
        for (int w = 0; w < kT / 64; ++w) tot += cnt[w][t];
        base[t] += tot;
        __syncthreads();
    }
}

struct WarmLayout {
    size_t stats, x, pred, kk, kv, k2, i2, hist, hoff, sums, offs, row, total;
    int nblk;
};

WarmLayout warm_layout(int64_t wtot, int32_t nchains, int32_t nkeys) {
    WarmLayout L{};
    size_t o = 0;
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t n = size_t(wtot);
    L.nblk = int((2 * wtot + kRsTile - 1) / kRsTile);
    const size_t hn = size_t(256) * size_t(std::max(L.nblk, 1));
    const size_t sb = size_t(scan::blocks(int64_t(hn)));
    L.stats = o; o = al(o + sizeof(RowStat) * size_t(nkeys) * size_t(std::max(nchains, 1)));
    L.x = o; o = al(o + n * 8 * 4);
    L.pred = o; o = al(o + n * 2 * 4);
    L.kk = o; o = al(o + 2 * n * 8);
    L.kv = o; o = al(o + 2 * n * 4);
    L.k2 = o; o = al(o + 2 * n * 8);
    L.i2 = o; o = al(o + 2 * n * 4);
    L.hist = o; o = al(o + hn * 4);
    L.hoff = o; o = al(o + (hn + 1) * 4);
    L.sums = o; o = al(o + sb * 4);
    L.offs = o; o = al(o + (sb + 1) * 4);
    L.row = o; o = al(o + n * 8);
    L.total = o;
    return L;
}

// Sorts (k, v) by the low `bits` bits of the keys, ping-ponging through
// (k2, v2); returns the values array that holds the result.
uint32_t *radix_sort(uint64_t *k, uint32_t *v, uint64_t *k2, uint32_t *v2, int64_t n, int bits, char *base,
                     const WarmLayout &L, hipStream_t s) {
    int *hist = reinterpret_cast<int *>(base + L.hist);
    int *hoff = reinterpret_cast<int *>(base + L.hoff);
    int *sums = reinterpret_cast<int *>(base + L.sums);
    int *offs = reinterpret_cast<int *>(base + L.offs);
    const int64_t hn = int64_t(256) * L.nblk;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(uint32_t(L.nblk)), dim3(kT), 0, s, k, n, shift, L.nblk, hist);
        scan::exclusive(hist, hn, sums, offs, hoff, s);
        hipLaunchKernelGGL(k_rs_scatter, dim3(uint32_t(L.nblk)), dim3(kT), 0, s, k, v, n, shift, L.nblk, hoff, k2,
                           v2);
        std::swap(k, k2);
        std::swap(v, v2);
    }
    return v;
}

}  // namespace

extern "C" size_t amh_warm_workspace_bytes(int64_t wtot, int32_t nchains, int32_t nkeys) {
    return warm_layout(wtot, nchains, nkeys).total;
}

// Features (x_dev[wtot][8]) of key frame y and, if t_dev != NULL, the
// targets (t_dev[wtot][2]) from row y+1; uses the workspace's statistics.
extern "C" int amh_learn_features(const amh_render *R, int32_t y, float *x_dev, float *t_dev, void *ws,
                                  size_t ws_bytes, void *stream) {
    if (!R || !x_dev || !ws || y < 0 || y >= R->nkeys || R->wtot <= 0 || R->nchains <= 0 || !R->pts_dev ||
        !R->keys_dev || !R->chain_off_dev || !R->chain_w_dev) {
        set_error("amh_learn_features: invalid argument");
        return AMH_E_ARG;
    }
    const WarmLayout L = warm_layout(R->wtot, R->nchains, R->nkeys);
    if (ws_bytes < L.total) {
        set_error("amh_learn_features: workspace %zu < %zu bytes", ws_bytes, L.total);
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    RowStat *st = reinterpret_cast<RowStat *>(static_cast<char *>(ws) + L.stats);
    hipLaunchKernelGGL(k_row_stats, dim3(uint32_t(R->nchains), uint32_t(R->nkeys)), dim3(kT), 0, s, *R, st);
    hipLaunchKernelGGL(k_features, dim3(uint32_t((R->wtot + kT - 1) / kT)), dim3(kT), 0, s, *R, y, st, x_dev, t_dev);
    AMH_LAUNCH_CHECK("amh_learn_features");
    return AMH_OK;
}

// Reorders rows 1 .. nkeys-1 of pts_dev (the chains R describes; R->pts_dev
// must be pts_dev) by the MLP's predictions, in place.
extern "C" int amh_warm_start(void *mlp, const amh_render *R, uint64_t *pts_dev, void *ws, size_t ws_bytes,
                              void *stream) {
    if (!mlp || !R || !pts_dev || R->pts_dev != pts_dev || !ws || R->wtot <= 0 || R->nchains <= 0 ||
        R->nchains > 65535 || R->wtot > int64_t(0x3FFFFFFF) || !R->keys_dev || !R->chain_off_dev ||
        !R->chain_w_dev) {
        set_error("amh_warm_start: invalid argument");
        return AMH_E_ARG;
    }
    const WarmLayout L = warm_layout(R->wtot, R->nchains, R->nkeys);
    if (ws_bytes < L.total) {
        set_error("amh_warm_start: workspace %zu < %zu bytes", ws_bytes, L.total);
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    amh::ProfScope ps_("warm_start", s);
    char *b = static_cast<char *>(ws);
    RowStat *st = reinterpret_cast<RowStat *>(b + L.stats);
    float *x = reinterpret_cast<float *>(b + L.x), *pred = reinterpret_cast<float *>(b + L.pred);
    uint64_t *kk = reinterpret_cast<uint64_t *>(b + L.kk), *k2 = reinterpret_cast<uint64_t *>(b + L.k2);
    uint32_t *kv = reinterpret_cast<uint32_t *>(b + L.kv), *i2 = reinterpret_cast<uint32_t *>(b + L.i2);
    uint64_t *row = reinterpret_cast<uint64_t *>(b + L.row);
    const uint32_t ga = uint32_t((R->wtot + kT - 1) / kT);
    hipLaunchKernelGGL(k_row_stats, dim3(uint32_t(R->nchains), uint32_t(R->nkeys)), dim3(kT), 0, s, *R, st);
    // key bits: set (1), chain, and y, x at 1/16 px over the frame
    int cb = 0, cx = 0;
    while ((1 << cb) < R->nchains) ++cb;
    while ((int64_t(1) << cx) < int64_t(std::max(std::max(R->width, R->height), 1)) * 16) ++cx;
    const int bits = 1 + cb + 2 * cx;
    for (int y = 0; y + 1 < R->nkeys; ++y) {
        hipLaunchKernelGGL(k_features, dim3(ga), dim3(kT), 0, s, *R, y, st, x, static_cast<float *>(nullptr));
        if (int r = amh_mlp_forward(mlp, x, R->wtot, pred, stream)) return r;
        hipLaunchKernelGGL(k_match_keys, dim3(ga), dim3(kT), 0, s, *R, y + 1, st, pred, cb, cx, kk, kv);
        const uint32_t *sorted = radix_sort(kk, kv, k2, i2, 2 * R->wtot, bits, b, L, s);
        uint64_t *dst = pts_dev + int64_t(y + 1) * R->wtot;
        if (int r = check_hip(hipMemcpyAsync(row, dst, size_t(R->wtot) * 8, hipMemcpyDeviceToDevice, s), "warm"))
            return r;
        hipLaunchKernelGGL(k_match_apply, dim3(ga), dim3(kT), 0, s, R->wtot, sorted, row, dst);
    }
    AMH_LAUNCH_CHECK("amh_warm_start");
    return AMH_OK;
}
