// Key-frame ingest, blob segmentation, blob statistics and chain build.
//
// Segmentation replaces thread::blobify_frame (thread.cpp:225-412).  At the
// library default blob_threshold 1.0 every 4-neighbour pair of present pixels
// passes the colour test (color_distance <= 1, color.h:17-24), so the greedy
// merge ends with exactly one blob per 4-connected component.  Here it is a
// union-find: LDS-local within 64x16 tiles, then a global merge across tile
// borders (atomicMin links, root = smallest row-major index), then flatten.
#include "colour.h"
#include "common.h"
#include "../../include/atomorph_hip.h"
#include "scan.h"

using namespace amh;

namespace {

constexpr int TW = 64, TH = 16, TPX = TW * TH;  // CCL tile
constexpr int kCclThreads = 256;
constexpr int kScanBlock = 1024;               // pixels per compaction block

// ----------------------------------------------------------------- ingest
__global__ void k_ingest(const uint32_t *__restrict__ rgba, int64_t n, int hsp, uint32_t *__restrict__ stored,
                         uint32_t *__restrict__ shown) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = rgba[i];
    uint32_t st = 0, sh = 0;
    if ((c >> 24) != 0) {
        st = hsp ? rgb_to_hsp(c) : c;
        sh = hsp ? hsp_to_rgb(st) : st;
    }
    if (stored) stored[i] = st;
    if (shown) shown[i] = sh;
}

// -------------------------------------------------------------- union-find
__device__ __forceinline__ int lds_find(int *L, int x) {
    int p = L[x];
    while (p != x) {
        x = p;
        p = L[x];
    }
    return x;
}
__device__ void lds_unite(int *L, int a, int b) {
    for (;;) {
        a = lds_find(L, a);
        b = lds_find(L, b);
        if (a == b) return;
        if (a < b) {
            const int old = atomicMin(&L[b], a);
            if (old == b) return;
            b = old;
        } else {
            const int old = atomicMin(&L[a], b);
            if (old == a) return;
            a = old;
        }
    }
}
__device__ __forceinline__ int g_load(const int *L, int x) {
    return __hip_atomic_load(L + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int g_find(int *L, int x) {
    int p = g_load(L, x);
    while (p != x) {
        x = p;
        p = g_load(L, x);
    }
    return x;
}
__device__ void g_unite(int *L, int a, int b) {
    for (;;) {
        a = g_find(L, a);
        b = g_find(L, b);
        if (a == b) return;
        if (a < b) {
            const int old = atomicMin(&L[b], a);
            if (old == b) return;
            b = old;
        } else {
            const int old = atomicMin(&L[a], b);
            if (old == a) return;
            a = old;
        }
    }
}

__global__ __launch_bounds__(kCclThreads) void k_ccl_local(const uint32_t *__restrict__ rgba, int W, int H,
                                                           int *__restrict__ labels) {
    __shared__ int L[TPX];
    const int tx0 = blockIdx.x * TW, ty0 = blockIdx.y * TH;
    for (int l = threadIdx.x; l < TPX; l += kCclThreads) {
        const int x = tx0 + (l % TW), y = ty0 + (l / TW);
        L[l] = (x < W && y < H && (rgba[int64_t(y) * W + x] >> 24) != 0) ? l : -1;
    }
    __syncthreads();
    for (int l = threadIdx.x; l < TPX; l += kCclThreads) {
        if (L[l] < 0) continue;
        const int lx = l % TW, ly = l / TW;
        if (lx + 1 < TW && L[l + 1] >= 0) lds_unite(L, l, l + 1);
        if (ly + 1 < TH && L[l + TW] >= 0) lds_unite(L, l, l + TW);
    }
    __syncthreads();
    for (int l = threadIdx.x; l < TPX; l += kCclThreads) {
        const int x = tx0 + (l % TW), y = ty0 + (l / TW);
        if (x >= W || y >= H) continue;
        int lab = -1;
        if (L[l] >= 0) {
            const int r = lds_find(L, l);
            lab = (ty0 + r / TW) * W + tx0 + (r % TW);
        }
        labels[int64_t(y) * W + x] = lab;
    }
}

// Merge across the right and bottom edge of every tile.
__global__ void k_ccl_merge(int W, int H, int *labels) {
    const int tiles_x = (W + TW - 1) / TW;
    const int tile = blockIdx.x;
    const int tx0 = (tile % tiles_x) * TW, ty0 = (tile / tiles_x) * TH;
    // threads 0..TH-1: right edge rows; TH..TH+TW-1: bottom edge columns
    const int t = threadIdx.x;
    if (t < TH) {
        const int x = tx0 + TW - 1, y = ty0 + t;
        if (x + 1 < W && y < H) {
            const int a = y * W + x;
            if (g_load(labels, a) >= 0 && g_load(labels, a + 1) >= 0) g_unite(labels, a, a + 1);
        }
    } else if (t < TH + TW) {
        const int x = tx0 + (t - TH), y = ty0 + TH - 1;
        if (x < W && y + 1 < H) {
            const int a = y * W + x;
            if (g_load(labels, a) >= 0 && g_load(labels, a + W) >= 0) g_unite(labels, a, a + W);
        }
    }
}

__global__ void k_ccl_flatten(int64_t n, int *labels) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (labels[i] >= 0) labels[i] = g_find(labels, int(i));
}

// ------------------------------------------------------------ compaction
// Blob numbers follow the order of the components' smallest pixel index:
// count roots per block, scan, rank inside the block.
__global__ __launch_bounds__(256) void k_root_count(const int *__restrict__ labels, int64_t n,
                                                    int *__restrict__ bcount) {
    __shared__ int part[4];
    const int64_t base = int64_t(blockIdx.x) * kScanBlock;
    int cnt = 0;
    for (int l = threadIdx.x; l < kScanBlock; l += 256) {
        const int64_t i = base + l;
        if (i < n && labels[i] == int(i)) ++cnt;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Single-workgroup exclusive scan of nb ints into boff; total into boff[nb].
__global__ __launch_bounds__(1024) void k_scan_small(const int *__restrict__ in, int nb, int *__restrict__ out) {
    __shared__ int buf[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        int v = i < nb ? in[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) out[i] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[nb] = carry;
}

__global__ __launch_bounds__(256) void k_root_ids(const int *__restrict__ labels, int64_t n,
                                                  const int *__restrict__ boff, int *__restrict__ root_id) {
    __shared__ int wcount[4 * (kScanBlock / 256)];
    const int64_t base = int64_t(blockIdx.x) * kScanBlock;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // pass over the block in 4 chunks of 256; ranks in row-major order
    int running = boff[blockIdx.x];
    for (int chunk = 0; chunk < kScanBlock / 256; ++chunk) {
        const int64_t i = base + chunk * 256 + threadIdx.x;
        const bool root = i < n && labels[i] == int(i);
        const uint64_t ball = __ballot(root);
        const int below = __popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) wcount[chunk * 4 + wave] = __popcll(ball);
        __syncthreads();
        int before = running;
        for (int w = 0; w < wave; ++w) before += wcount[chunk * 4 + w];
        if (root) root_id[i] = before + below;
        running += wcount[chunk * 4 + 0] + wcount[chunk * 4 + 1] + wcount[chunk * 4 + 2] + wcount[chunk * 4 + 3];
        __syncthreads();
    }
}

// Per-blob sums in two levels: each block covers kStatChunk consecutive
// pixels (a few rows), folds wave-uniform runs with shuffles and everything
// else with LDS atomics into a small open-addressed table of the blobs it
// meets, then flushes one set of global atomics per (block, blob).  Device
// atomics on this GPU are performed past the L2, so the global traffic is
// kept to the table flushes.
constexpr int kStatThreads = 256;
constexpr int kStatChunk = 4096;
constexpr int kStatSlots = 256;

struct StatTable {
    int key[kStatSlots];
    unsigned long long sum[kStatSlots][7];  // count, x, y, c0..c3
    int box[kStatSlots][4];                 // x1 y1 x2 y2
};

// Slot of blob b in the table (inserting it), or -1 when the table is full.
__device__ __forceinline__ int stat_slot(StatTable &T, int b) {
    int h = int((uint32_t(b) * 2654435761u) >> 24);  // 8 bits: [0, 256)
    for (int probe = 0; probe < kStatSlots; ++probe) {
        const int k = atomicCAS(&T.key[h], -1, b);
        if (k == -1 || k == b) return h;
        h = (h + 1) & (kStatSlots - 1);
    }
    return -1;
}

__device__ __forceinline__ void stat_add_global(amh_blob_stat *p, const unsigned long long *v, int x1, int y1,
                                                int x2, int y2) {
    atomicAdd((unsigned long long *)&p->count, v[0]);
    atomicAdd((unsigned long long *)&p->sum_x, v[1]);
    atomicAdd((unsigned long long *)&p->sum_y, v[2]);
    for (int k = 0; k < 4; ++k) atomicAdd((unsigned long long *)&p->sum_c[k], v[3 + k]);
    atomicMin(&p->x1, x1);
    atomicMin(&p->y1, y1);
    atomicMax(&p->x2, x2);
    atomicMax(&p->y2, y2);
}

__device__ __forceinline__ void stat_add(StatTable &T, amh_blob_stat *st, int b, const unsigned long long *v, int x1,
                                         int y1, int x2, int y2) {
    const int h = stat_slot(T, b);
    if (h < 0) {  // more blobs in one chunk than slots: straight to memory
        stat_add_global(st + b, v, x1, y1, x2, y2);
        return;
    }
    for (int k = 0; k < 7; ++k) atomicAdd(&T.sum[h][k], v[k]);
    atomicMin(&T.box[h][0], x1);
    atomicMin(&T.box[h][1], y1);
    atomicMax(&T.box[h][2], x2);
    atomicMax(&T.box[h][3], y2);
}

__global__ __launch_bounds__(kStatThreads) void k_blob_stats(const int *__restrict__ labels,
                                                             const uint32_t *__restrict__ stored, int W, int64_t n,
                                                             const int *__restrict__ root_id, int *__restrict__ bop,
                                                             amh_blob_stat *__restrict__ st) {
    __shared__ StatTable T;
    for (int h = threadIdx.x; h < kStatSlots; h += kStatThreads) {
        T.key[h] = -1;
        for (int k = 0; k < 7; ++k) T.sum[h][k] = 0ull;
        T.box[h][0] = T.box[h][1] = 0x7FFFFFFF;
        T.box[h][2] = T.box[h][3] = -1;
    }
    __syncthreads();
    const int64_t base = int64_t(blockIdx.x) * kStatChunk;
    for (int l = threadIdx.x; l < kStatChunk; l += kStatThreads) {
        const int64_t i = base + l;
        const bool fg = i < n && labels[i] >= 0;
        int b = -1;
        if (i < n) {
            if (fg) b = root_id[labels[i]];
            bop[i] = b;
        }
        const int x = fg ? int(i % W) : 0, y = fg ? int(i / W) : 0;
        const uint32_t c = fg ? stored[i] : 0u;
        const int b0 = __shfl(b, 0, 64);
        const bool uniform = __all(b == b0) && b0 >= 0;
        unsigned long long v[7] = {1ull, (unsigned long long)x, (unsigned long long)y, c & 255u, (c >> 8) & 255u,
                                   (c >> 16) & 255u, c >> 24};
        if (uniform) {  // whole wave in one blob: reduce first
            int mnx = x, mny = y, mxx = x, mxy = y;
            for (int off = 32; off > 0; off >>= 1) {
                for (int k = 0; k < 7; ++k) v[k] += __shfl_down(v[k], off, 64);
                mnx = min(mnx, __shfl_down(mnx, off, 64));
                mny = min(mny, __shfl_down(mny, off, 64));
                mxx = max(mxx, __shfl_down(mxx, off, 64));
                mxy = max(mxy, __shfl_down(mxy, off, 64));
            }
            if ((threadIdx.x & 63) == 0) stat_add(T, st, b0, v, mnx, mny, mxx, mxy);
        } else if (fg) {
            stat_add(T, st, b, v, x, y, x, y);
        }
        if (fg && labels[i] == int(i)) st[b].first = i;
    }
    __syncthreads();
    for (int h = threadIdx.x; h < kStatSlots; h += kStatThreads) {
        const int b = T.key[h];
        if (b >= 0) stat_add_global(st + b, T.sum[h], T.box[h][0], T.box[h][1], T.box[h][2], T.box[h][3]);
    }
}

__global__ void k_stats_init(amh_blob_stat *st, int nb) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    amh_blob_stat z;
    z.count = z.sum_x = z.sum_y = 0;
    z.sum_c[0] = z.sum_c[1] = z.sum_c[2] = z.sum_c[3] = 0;
    z.x1 = z.y1 = 0x7FFFFFFF;
    z.x2 = z.y2 = -1;
    z.first = -1;
    st[b] = z;
}

// ----------------------------------------------------- grouping by blob
// Stable counting sort of pixel indices by blob number.  Per block of
// kScanBlock pixels: histogram of blob numbers; a column scan over blocks
// gives each (block, blob) its start; inside a block, waves place their
// pixels in order using a per-blob running offset kept in LDS.
// Sizes: nb blocks x nblobs counters.
__global__ __launch_bounds__(256) void k_group_hist(const int *__restrict__ bop, int64_t n, int nblobs,
                                                    int *__restrict__ hist) {
    const int64_t base = int64_t(blockIdx.x) * kScanBlock;
    int *h = hist + int64_t(blockIdx.x) * nblobs;
    for (int l = threadIdx.x; l < kScanBlock; l += 256) {
        const int64_t i = base + l;
        if (i < n) {
            const int b = bop[i];
            if (b >= 0) atomicAdd(&h[b], 1);
        }
    }
}

// Column scan, one workgroup per blob b: hist[blk][b] becomes the number of
// b's pixels in earlier blocks, and count[b] the blob's total.
__global__ __launch_bounds__(256) void k_group_colscan(int *__restrict__ hist, int nblk, int nblobs,
                                                       int64_t *__restrict__ count) {
    __shared__ int wsum[4];
    __shared__ int carry;
    const int b = int(blockIdx.x), lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nblk; base += 256) {
        const int k = base + int(threadIdx.x);
        int *p = hist + int64_t(k) * nblobs + b;
        const int v = k < nblk ? *p : 0;
        int incl = v;
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int before = carry;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        if (k < nblk) *p = before + incl - v;
        __syncthreads();
        if (threadIdx.x == 255) carry = before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) count[b] = carry;
}

// Blob starts: exclusive scan of the counts in place, first[nblobs] = total.
__global__ __launch_bounds__(1024) void k_first_scan(int nblobs, int64_t *__restrict__ first) {
    __shared__ int64_t carry;
    __shared__ int64_t buf[1024];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nblobs; base += 1024) {
        const int b = base + threadIdx.x;
        const int64_t cnt = b < nblobs ? first[b] : 0;
        buf[threadIdx.x] = cnt;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int64_t t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (b < nblobs) first[b] = carry + buf[threadIdx.x] - cnt;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) first[nblobs] = carry;
}

__global__ __launch_bounds__(64) void k_group_scatter(const int *__restrict__ bop, int64_t n, int nblobs,
                                                      int *__restrict__ hist, const int64_t *__restrict__ first,
                                                      int *__restrict__ list) {
    // one wave per block: walk the block's pixels in order, 64 at a time;
    // lanes with the same blob are ranked by a match over the wave.
    const int64_t base = int64_t(blockIdx.x) * kScanBlock;
    int *h = hist + int64_t(blockIdx.x) * nblobs;
    const int lane = threadIdx.x;
    for (int chunk = 0; chunk < kScanBlock; chunk += 64) {
        const int64_t i = base + chunk + lane;
        const int b = i < n ? bop[i] : -1;
        // rank among earlier lanes with the same blob, the run's size and the
        // last lane holding that blob (all lanes run the full loop)
        int rank = 0, total = 0, lastlane = -1;
        for (int s = 0; s < 64; ++s) {
            const int bs = __shfl(b, s, 64);
            if (bs == b) {
                if (s < lane) ++rank;
                ++total;
                lastlane = s;
            }
        }
        const int start = b >= 0 ? h[b] : 0;
        __syncthreads();
        if (b >= 0) list[first[b] + start + rank] = int(i);
        if (b >= 0 && lastlane == lane) h[b] = start + total;
        __syncthreads();
    }
}

// Grouping with many blobs: a stable LSD radix sort of the pixels by blob
// number (8-bit digits, as many passes as the blob numbers need), so the
// memory is O(pixels + blobs) whatever the blob count.  Pixels without a
// blob get key nblobs and sort last.  Per pass: a digit histogram per block
// of kGrpTile pixels, an exclusive scan of the digit-major histograms, and a
// scatter that ranks equal digits in pixel order with ballots (stable).
constexpr int kGrpTile = 256 * 8;

__device__ __forceinline__ uint32_t grp_key(bool from_bop, const int *__restrict__ bop,
                                            const uint32_t *__restrict__ kin, int64_t e, int nblobs) {
    if (!from_bop) return kin[e];
    const int b = bop[e];
    return b >= 0 ? uint32_t(b) : uint32_t(nblobs);
}

__global__ __launch_bounds__(256) void k_grp_hist(bool from_bop, const int *__restrict__ bop,
                                                  const uint32_t *__restrict__ kin, int64_t n, int nblobs, int shift,
                                                  int nblk, int *__restrict__ hist) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t b0 = int64_t(blockIdx.x) * kGrpTile;
    for (int j = 0; j < kGrpTile / 256; ++j) {
        const int64_t e = b0 + j * 256 + threadIdx.x;
        if (e < n) atomicAdd(&h[(grp_key(from_bop, bop, kin, e, nblobs) >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[int64_t(threadIdx.x) * nblk + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one pass.  `last`: values go to `list` and only those of
// real blobs (their places are below the number of blob pixels).
__global__ __launch_bounds__(256) void k_grp_scatter(bool from_bop, const int *__restrict__ bop,
                                                     const uint32_t *__restrict__ kin,
                                                     const int32_t *__restrict__ vin, int64_t n, int nblobs,
                                                     int shift, int nblk, const int *__restrict__ off, bool last,
                                                     uint32_t *__restrict__ kout, int32_t *__restrict__ vout) {
    __shared__ int base[256];
    __shared__ int cnt[4][256];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    base[t] = off[int64_t(t) * nblk + blockIdx.x];
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t b0 = int64_t(blockIdx.x) * kGrpTile;
    for (int j = 0; j < kGrpTile / 256; ++j) {
        for (int w = 0; w < 4; ++w) cnt[w][t] = 0;
        __syncthreads();
        const int64_t e = b0 + j * 256 + t;
        const bool ok = e < n;
        const uint32_t k = ok ? grp_key(from_bop, bop, kin, e, nblobs) : 0u;
        const int32_t v = ok ? (from_bop ? int32_t(e) : vin[e]) : 0;
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
            if (!last) {
                kout[pos] = k;
                vout[pos] = v;
            } else if (k < uint32_t(nblobs)) {
                kout[pos] = k;
                vout[pos] = v;
            }
        }
        __syncthreads();
        int tot = 0;
        for (int w = 0; w < 4; ++w) tot += cnt[w][t];
        base[t] += tot;
        __syncthreads();
    }
}

// Per blob, from the sorted keys of its pixels (places [0, *total)): the
// start and end of its run; absent blobs keep start = end = 0.
__global__ void k_grp_bounds_total(const uint32_t *__restrict__ key, const unsigned long long *__restrict__ ptotal,
                                   int *__restrict__ start, int *__restrict__ end) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t total = int64_t(*ptotal);
    if (p >= total) return;
    const uint32_t k = key[p];
    if (p == 0 || key[p - 1] != k) start[k] = int(p);
    if (p == total - 1 || key[p + 1] != k) end[k] = int(p + 1);
}

__global__ void k_grp_counts(const int *__restrict__ start, const int *__restrict__ end, int nblobs,
                             int64_t *__restrict__ count) {
    const int b = int(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b < nblobs) count[b] = end[b] - start[b];
}

// Counts the pixels that have a blob.
__global__ void k_grp_present(const int *__restrict__ bop, int64_t n, unsigned long long *__restrict__ total) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool has = i < n && bop[i] >= 0;
    const uint64_t m = __ballot(has);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(total, (unsigned long long)__popcll(m));
}

struct GroupLayout {
    bool radix;
    int nblk;                   // table: pixel blocks; radix: sort tiles
    int passes;
    size_t hist, k0, v0, k1, v1, hoff, sums, offs, start, end, total, bytes;
};

GroupLayout group_layout(int64_t n, int64_t nblobs) {
    GroupLayout L{};
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    const int64_t tblk = (n + kScanBlock - 1) / kScanBlock;
    const int64_t nb = nblobs > 0 ? nblobs : 1;
    // the per-(block, blob) table when it is at most 2 ints per pixel
    if (tblk * nb <= 2 * std::max<int64_t>(n, 1)) {
        L.radix = false;
        L.nblk = int(tblk);
        L.bytes = size_t(tblk) * size_t(nb) * sizeof(int);
        return L;
    }
    L.radix = true;
    L.nblk = int((n + kGrpTile - 1) / kGrpTile);
    int bits = 1;
    while ((int64_t(1) << bits) <= nblobs) ++bits;  // keys 0..nblobs
    L.passes = (bits + 7) / 8;
    const size_t hn = size_t(256) * size_t(std::max(L.nblk, 1));
    size_t o = 0;
    L.k0 = o; o = al(o + size_t(n) * 4);
    L.v0 = o; o = al(o + size_t(n) * 4);
    L.k1 = o; o = al(o + size_t(n) * 4);
    L.v1 = o; o = al(o + size_t(n) * 4);
    L.hist = o; o = al(o + hn * 4);
    L.hoff = o; o = al(o + (hn + 1) * 4);
    L.sums = o; o = al(o + size_t(scan::blocks(int64_t(hn))) * 4);
    L.offs = o; o = al(o + (size_t(scan::blocks(int64_t(hn))) + 1) * 4);
    L.start = o; o = al(o + size_t(nb) * 4);
    L.end = o; o = al(o + size_t(nb) * 4);
    L.total = o; o = al(o + 8);
    L.bytes = o;
    return L;
}

int group_radix(const GroupLayout &L, const int32_t *bop, int64_t n, int nblobs, char *ws, int64_t *first,
                int32_t *list, hipStream_t s) {
    uint32_t *k[2] = {reinterpret_cast<uint32_t *>(ws + L.k0), reinterpret_cast<uint32_t *>(ws + L.k1)};
    int32_t *v[2] = {reinterpret_cast<int32_t *>(ws + L.v0), reinterpret_cast<int32_t *>(ws + L.v1)};
    int *hist = reinterpret_cast<int *>(ws + L.hist), *hoff = reinterpret_cast<int *>(ws + L.hoff);
    int *sums = reinterpret_cast<int *>(ws + L.sums), *offs = reinterpret_cast<int *>(ws + L.offs);
    int *start = reinterpret_cast<int *>(ws + L.start), *end = reinterpret_cast<int *>(ws + L.end);
    unsigned long long *total = reinterpret_cast<unsigned long long *>(ws + L.total);
    const int64_t hn = int64_t(256) * L.nblk;
    int cur = 0;
    for (int p = 0; p < L.passes; ++p) {
        const bool from_bop = p == 0, last = p + 1 == L.passes;
        hipLaunchKernelGGL(k_grp_hist, dim3(uint32_t(L.nblk)), dim3(256), 0, s, from_bop, bop, k[cur], n, nblobs,
                           8 * p, L.nblk, hist);
        scan::exclusive(hist, hn, sums, offs, hoff, s);
        hipLaunchKernelGGL(k_grp_scatter, dim3(uint32_t(L.nblk)), dim3(256), 0, s, from_bop, bop, k[cur], v[cur], n,
                           nblobs, 8 * p, L.nblk, hoff, last, k[cur ^ 1], last ? list : v[cur ^ 1]);
        cur ^= 1;
    }
    // runs of the sorted keys -> counts -> first[] (exclusive scan)
    if (int r = check_hip(hipMemsetAsync(ws + L.start, 0, L.total + 8 - L.start, s), "amh_group_pixels memset"))
        return r;
    const uint32_t gn = uint32_t((n + 255) / 256);
    hipLaunchKernelGGL(k_grp_present, dim3(gn), dim3(256), 0, s, bop, n, total);
    // the blob pixels are the first `total` places; bounds over all n places
    // would read the keys of blobless pixels, which the last pass skipped
    hipLaunchKernelGGL(k_grp_bounds_total, dim3(gn), dim3(256), 0, s, k[cur], total, start, end);
    hipLaunchKernelGGL(k_grp_counts, dim3(uint32_t((nblobs + 255) / 256)), dim3(256), 0, s, start, end, nblobs,
                       first);
    return AMH_OK;
}

// ------------------------------------------------------------ chain build
// Row y of every chain (thread.cpp:803-848).
__global__ void k_build_row(int y, int h, int64_t wtot, int nchains, const int64_t *__restrict__ off,
                            const int32_t *__restrict__ cw, const int32_t *__restrict__ cblob,
                            const uint32_t *__restrict__ exy, const int32_t *__restrict__ plist,
                            const int64_t *__restrict__ bfirst, int W, uint64_t seed,
                            uint64_t *__restrict__ pts) {
    const int64_t a = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (a >= wtot) return;
    int lo = 0, hi = nchains - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= a) lo = mid; else hi = mid - 1;
    }
    const int c = lo;
    const int64_t x = a - off[c];
    const int b = cblob[int64_t(c) * h + y];
    int64_t n;
    int64_t first = 0;
    if (b >= 0) {
        first = bfirst[b];
        n = bfirst[b + 1] - first;
    } else {
        n = 1;
    }
    uint64_t pt;
    if (x < n) {
        if (b >= 0) {
            const int p = plist[first + x];
            pt = make_pt(uint32_t(p % W), uint32_t(p / W), 0, 0, HAS_PIXEL | HAS_FLUID);
        } else {
            const uint32_t e = exy[int64_t(c) * h + y];
            pt = make_pt(e & 0xFFFFu, e >> 16, 0, 0, HAS_FLUID);
        }
    } else {
        // duplicate of a shuffled copy (thread.cpp:837-848)
        const uint64_t key = mix64(seed ^ (uint64_t(c) << 32) ^ uint64_t(y));
        const uint32_t q = feistel_perm(uint32_t(x % n), uint32_t(n), key);
        uint32_t px, py, fl;
        if (b >= 0) {
            const int p = plist[first + q];
            px = uint32_t(p % W);
            py = uint32_t(p / W);
            fl = HAS_PIXEL;
        } else {
            const uint32_t e = exy[int64_t(c) * h + y];
            px = e & 0xFFFFu;
            py = e >> 16;
            fl = 0;
        }
        const uint64_t r = mix64(key ^ (uint64_t(x) * 0x9E3779B97F4A7C15ull));
        pt = make_pt(px, py, uint32_t(r & 255u), uint32_t((r >> 8) & 255u), fl);
    }
    pts[int64_t(y) * wtot + a] = pt;
}

int check_frame(int32_t W, int32_t H, const char *what) {
    if (W <= 0 || H <= 0 || W > 65535 || H > 65535 || int64_t(W) * H >= (int64_t(1) << 31)) {
        set_error("%s: invalid frame size %dx%d", what, W, H);
        return AMH_E_ARG;
    }
    return AMH_OK;
}

}  // namespace

extern "C" int amh_ingest_frame(const uint32_t *rgba, int32_t W, int32_t H, int32_t hsp, uint32_t *stored,
                                uint32_t *shown, void *stream) {
    if (int r = check_frame(W, H, "amh_ingest_frame")) return r;
    if (!rgba || (!stored && !shown)) {
        set_error("amh_ingest_frame: NULL buffer");
        return AMH_E_ARG;
    }
    const int64_t n = int64_t(W) * H;
    hipLaunchKernelGGL(k_ingest, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream), rgba,
                       n, hsp, stored, shown);
    AMH_LAUNCH_CHECK("amh_ingest_frame");
    return AMH_OK;
}

extern "C" int amh_segment(const uint32_t *rgba, int32_t W, int32_t H, int32_t *labels, void *stream) {
    if (int r = check_frame(W, H, "amh_segment")) return r;
    if (!rgba || !labels) {
        set_error("amh_segment: NULL buffer");
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int tx = (W + TW - 1) / TW, ty = (H + TH - 1) / TH;
    hipLaunchKernelGGL(k_ccl_local, dim3(tx, ty), dim3(kCclThreads), 0, s, rgba, W, H, labels);
    hipLaunchKernelGGL(k_ccl_merge, dim3(tx * ty), dim3(128), 0, s, W, H, labels);
    const int64_t n = int64_t(W) * H;
    hipLaunchKernelGGL(k_ccl_flatten, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, n, labels);
    AMH_LAUNCH_CHECK("amh_segment");
    return AMH_OK;
}

extern "C" size_t amh_blob_workspace_bytes(int32_t W, int32_t H) {
    const int64_t n = int64_t(W) * H;
    const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
    return size_t(n) * sizeof(int) + size_t(2 * nb + 2) * sizeof(int);
}

extern "C" int amh_blob_count(const int32_t *labels, int32_t W, int32_t H, void *ws, int32_t *nblobs, void *stream) {
    if (int r = check_frame(W, H, "amh_blob_count")) return r;
    if (!labels || !ws || !nblobs) {
        set_error("amh_blob_count: NULL argument");
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t n = int64_t(W) * H;
    const int nb = int((n + kScanBlock - 1) / kScanBlock);
    int *root_id = static_cast<int *>(ws);
    int *bcount = root_id + n;
    int *boff = bcount + nb;
    hipLaunchKernelGGL(k_root_count, dim3(nb), dim3(256), 0, s, labels, n, bcount);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, bcount, nb, boff);
    hipLaunchKernelGGL(k_root_ids, dim3(nb), dim3(256), 0, s, labels, n, boff, root_id);
    AMH_LAUNCH_CHECK("amh_blob_count");
    int total = 0;
    if (int r = check_hip(hipMemcpyAsync(&total, boff + nb, sizeof(int), hipMemcpyDeviceToHost, s), "D2H count"))
        return r;
    if (int r = check_hip(hipStreamSynchronize(s), "amh_blob_count sync")) return r;
    *nblobs = total;
    return AMH_OK;
}

extern "C" int amh_blob_stats(const int32_t *labels, const uint32_t *stored, int32_t W, int32_t H, void *ws,
                              int32_t nblobs, int32_t *bop, amh_blob_stat *st, void *stream) {
    if (int r = check_frame(W, H, "amh_blob_stats")) return r;
    if (!labels || !stored || !ws || !bop || (nblobs > 0 && !st) || nblobs < 0) {
        set_error("amh_blob_stats: invalid argument");
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t n = int64_t(W) * H;
    if (nblobs > 0)
        hipLaunchKernelGGL(k_stats_init, dim3((nblobs + 255) / 256), dim3(256), 0, s, st, int(nblobs));
    amh::launch_timed("k_blob_stats", k_blob_stats, dim3(uint32_t((n + kStatChunk - 1) / kStatChunk)),
                      dim3(kStatThreads), 0, s, labels, stored, int(W), n, static_cast<const int *>(ws), bop, st);
    AMH_LAUNCH_CHECK("amh_blob_stats");
    return AMH_OK;
}

extern "C" size_t amh_group_workspace_bytes(int32_t npixels, int32_t nblobs) {
    return group_layout(npixels > 0 ? npixels : 0, nblobs > 0 ? nblobs : 0).bytes;
}

extern "C" int amh_group_pixels(const int32_t *bop, int32_t npixels, int32_t nblobs, void *ws, size_t ws_bytes,
                                int64_t *first, int32_t *list, void *stream) {
    if (!bop || !first || !list || npixels < 0 || nblobs < 0 ||
        ws_bytes < amh_group_workspace_bytes(npixels, nblobs) || (!ws && ws_bytes)) {
        set_error("amh_group_pixels: invalid argument");
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (nblobs == 0) return check_hip(hipMemsetAsync(first, 0, sizeof(int64_t), s), "amh_group_pixels");
    const int64_t n = npixels;
    const GroupLayout L = group_layout(n, nblobs);
    if (L.radix) {
        if (int r = group_radix(L, bop, n, int(nblobs), static_cast<char *>(ws), first, list, s)) return r;
        hipLaunchKernelGGL(k_first_scan, dim3(1), dim3(1024), 0, s, int(nblobs), first);
        AMH_LAUNCH_CHECK("amh_group_pixels");
        return AMH_OK;
    }
    const int nblk = L.nblk;
    int *hist = static_cast<int *>(ws);
    if (int r = check_hip(hipMemsetAsync(hist, 0, size_t(nblk) * nblobs * sizeof(int), s), "amh_group_pixels memset"))
        return r;
    hipLaunchKernelGGL(k_group_hist, dim3(nblk), dim3(256), 0, s, bop, n, int(nblobs), hist);
    hipLaunchKernelGGL(k_group_colscan, dim3(uint32_t(nblobs)), dim3(256), 0, s, hist, nblk, int(nblobs), first);
    hipLaunchKernelGGL(k_first_scan, dim3(1), dim3(1024), 0, s, int(nblobs), first);
    hipLaunchKernelGGL(k_group_scatter, dim3(nblk), dim3(64), 0, s, bop, n, int(nblobs), hist, first, list);
    AMH_LAUNCH_CHECK("amh_group_pixels");
    return AMH_OK;
}

extern "C" int amh_build_chains(const amh_chain_build *a, uint64_t *pts, void *stream) {
    if (!a || !pts || a->nkeys <= 0 || a->nchains <= 0 || a->wtot <= 0 || !a->chain_off_dev || !a->chain_w_dev ||
        !a->chain_blob_dev || !a->empty_xy_dev || !a->pixel_list_dev || !a->blob_first_dev || a->frame_width <= 0) {
        set_error("amh_build_chains: invalid argument");
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t blocks = uint32_t((a->wtot + 255) / 256);
    for (int y = 0; y < a->nkeys; ++y) {
        hipLaunchKernelGGL(k_build_row, dim3(blocks), dim3(256), 0, s, y, a->nkeys, a->wtot, a->nchains,
                           a->chain_off_dev, a->chain_w_dev, a->chain_blob_dev, a->empty_xy_dev,
                           a->pixel_list_dev[y], a->blob_first_dev[y], a->frame_width, a->seed, pts);
    }
    AMH_LAUNCH_CHECK("amh_build_chains");
    return AMH_OK;
}

// Host-callable colour conversions with the kernels' arithmetic.
extern "C" uint32_t amh_rgb_to_hsp(uint32_t rgba) { return rgb_to_hsp(rgba); }
extern "C" uint32_t amh_hsp_to_rgb(uint32_t hspa) { return hsp_to_rgb(hspa); }
extern "C" void amh_rgb2hsp(double r, double g, double b, double *h, double *s, double *p) { rgb2hsp(r, g, b, h, s, p); }
extern "C" void amh_hsp2rgb(double h, double s, double p, double *r, double *g, double *b) { hsp2rgb(h, s, p, r, g, b); }
