// Frame rendering on the atom path: morph::get_pixels(t) -> draw_atoms
// (morph.cpp:1302-1421) over get_pixels(blob, t) (morph.cpp:452-678).
//
// Per frame:
//   prepare  per atom: trajectory point and colour at t (FP64, the
//            reference's operation order), bilinear taps; per tap pixel the
//            smallest blob index that touches it (atomicMin).
//   splat    taps of a pixel's smallest blob go to dense per-pixel integer
//            accumulators (count, sum n, sum c*n with n = the integer
//            bilinear weight numerator); taps of any other blob are counted
//            per pixel.
//   scan     exclusive scan of those counts -> per-pixel overflow segments.
//   overflow the other taps scattered into their pixel's segment (the count
//            is decremented back to zero, so the workspace cleans itself).
//   composite per pixel: resolve the smallest blob, then the overflow blobs
//            in increasing blob order, "over" in that order, feather, the
//            background blend, RGBA8 out.
// Composite order equals the reference's (blob order); sums are exact, so
// nothing depends on the order the atoms land in.
#include <mutex>

#include "colour.h"
#include "common.h"
#include "../../include/atomorph_hip.h"
#include "scan.h"
#include "shade.h"

using namespace amh;
using namespace amh::shade;

namespace {

constexpr uint32_t kNoPos = 0xFFFFFFFFu;
// --------------------------------------------------------------- prepare
struct Rec {       // one atom's splat, 16 bytes
    uint32_t pos;  // y * W + x of the base tap, kNoPos if not drawn
    uint32_t fm;   // x_fract | y_fract << 8 | tap mask << 16
    uint32_t col;  // RGBA8
    int32_t blob;  // composite index
};

// What changes from frame to frame when frames replay from a captured
// hipGraph: the kernels read it from device memory instead of their
// arguments.
struct RenderParams {
    Frame F;
    uint32_t *out;
};

__global__ void k_set_render_params(RenderParams *dst, RenderParams v) { *dst = v; }

__global__ __launch_bounds__(256) void k_prepare(amh_render R, Frame F, Rec *__restrict__ recs, int *__restrict__ bmin,
                                                 const RenderParams *__restrict__ fp) {
    if (fp) F = fp->F;
    __shared__ int lagp[512], slopep[512];
    if (R.fading == kPerlin) {
        for (int i = threadIdx.x; i < 512; i += 256) {
            lagp[i] = R.lag_perm_dev[i];
            slopep[i] = R.slope_perm_dev[i];
        }
    }
    __syncthreads();
    const int64_t a = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const int c = block_chain(R.chain_off_dev, R.nchains, a);
    if (a >= R.wtot) return;
    Rec rec = {kNoPos, 0u, 0u, -1};
    const int W = R.width, H = R.height;
    const int b = R.rank_dev[int64_t(F.fk) * R.nchains + c];
    const uint64_t p1 = R.pts_dev[int64_t(F.fk) * R.wtot + a];
    const uint64_t p2 = R.pts_dev[int64_t(F.fnext) * R.wtot + a];
    const bool h1 = pt_flags(p1) & HAS_PIXEL, h2 = pt_flags(p2) & HAS_PIXEL;
    if (b < 0 || (!h1 && !h2)) {
        recs[a] = rec;
        return;
    }
    // colours at both ends (morph.cpp:495-522); a missing end fades from /
    // to the other end's colour with alpha 0
    uint32_t c1, c2;
    if (h1 && !h2) {
        c1 = key_px(R.keys_dev, F.fk, W, H, pt_x(p1), pt_y(p1));
        c2 = c1 & 0x00FFFFFFu;
    } else if (h2 && !h1) {
        c2 = key_px(R.keys_dev, F.fnext, W, H, pt_x(p2), pt_y(p2));
        c1 = c2 & 0x00FFFFFFu;
    } else {
        c1 = key_px(R.keys_dev, F.fk, W, H, pt_x(p1), pt_y(p1));
        c2 = key_px(R.keys_dev, F.fnext, W, H, pt_x(p2), pt_y(p2));
    }
    uint32_t qx, qy, qxf, qyf;
    traj_point(R, F, a, p1, p2, qx, qy, qxf, qyf);
    const uint32_t col = atom_colour(R, F, p1, c1, c2, lagp, slopep);
    // guards and taps (morph.cpp:552-589)
    const int X = int(qx), Y = int(qy);
    if (X >= W || Y >= H) {
        if (X > R.bbox_x2 || X < R.bbox_x1 || Y > R.bbox_y2 || Y < R.bbox_y1) {
            recs[a] = rec;
            return;
        }
    }
    const int xf = int(qxf), yf = int(qyf);
    uint32_t mask = 0;
    if ((255 - xf) * (255 - yf) > 0) mask |= 1u;
    if ((X < R.bbox_x2 || X + 1 < W) && xf * (255 - yf) > 0) mask |= 2u;
    if ((Y < R.bbox_y2 || Y + 1 < H) && (255 - xf) * yf > 0) mask |= 4u;
    if (xf * yf > 0 && ((Y < R.bbox_y2 && X < R.bbox_x2) || (Y + 1 < H && X + 1 < W))) mask |= 8u;
    // taps outside the image are dropped (draw_atoms skips pos >= W*H)
    if (X >= W || Y >= H) mask &= 0u;
    if (X + 1 >= W) mask &= ~(2u | 8u);
    if (Y + 1 >= H) mask &= ~(4u | 8u);
    if (!mask) {
        recs[a] = rec;
        return;
    }
    rec.pos = uint32_t(Y * W + X);
    rec.fm = qxf | (qyf << 8) | (mask << 16);
    rec.col = col;
    rec.blob = b;
    recs[a] = rec;
    if (mask & 1u) atomicMin(&bmin[rec.pos], b);
    if (mask & 2u) atomicMin(&bmin[rec.pos + 1], b);
    if (mask & 4u) atomicMin(&bmin[rec.pos + W], b);
    if (mask & 8u) atomicMin(&bmin[rec.pos + W + 1], b);
}

// Integer bilinear weight numerators of tap k (weights are these / 255^2).
__device__ __forceinline__ void tap(const Rec &r, int W, int k, uint32_t &p, uint32_t &n) {
    const uint32_t xf = r.fm & 255u, yf = (r.fm >> 8) & 255u;
    switch (k) {
        case 0: p = r.pos; n = (255u - xf) * (255u - yf); break;
        case 1: p = r.pos + 1; n = xf * (255u - yf); break;
        case 2: p = r.pos + W; n = (255u - xf) * yf; break;
        default: p = r.pos + W + 1; n = xf * yf; break;
    }
}

struct Acc {  // dense per-pixel accumulators (structure of arrays)
    unsigned long long *cn;       // count << 40 | sum n
    unsigned long long *s[4];     // sum c * n per channel
};

__global__ __launch_bounds__(256) void k_splat(const Rec *__restrict__ recs, int64_t n_atoms, int W,
                                               const int *__restrict__ bmin, Acc acc, int *__restrict__ ovcount) {
    const int64_t a = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (a >= n_atoms) return;
    const Rec r = recs[a];
    if (r.pos == kNoPos) return;
    const uint32_t mask = r.fm >> 16;
    for (int k = 0; k < 4; ++k) {
        if (!(mask & (1u << k))) continue;
        uint32_t p, n;
        tap(r, W, k, p, n);
        if (bmin[p] == r.blob) {
            atomicAdd(&acc.cn[p], (1ull << 40) | n);
            for (int c = 0; c < 4; ++c) atomicAdd(&acc.s[c][p], (unsigned long long)(ch(r.col, c) * n));
        } else {
            atomicAdd(&ovcount[p], 1);
        }
    }
}

struct Ov {  // an overflow tap, 16 bytes
    int32_t blob;
    uint32_t n;
    uint32_t col;
    uint32_t pad;
};

__global__ __launch_bounds__(256) void k_overflow(const Rec *__restrict__ recs, int64_t n_atoms, int W,
                                                  const int *__restrict__ bmin, int *__restrict__ ovcount,
                                                  const int *__restrict__ ovstart, Ov *__restrict__ ov) {
    const int64_t a = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (a >= n_atoms) return;
    const Rec r = recs[a];
    if (r.pos == kNoPos) return;
    const uint32_t mask = r.fm >> 16;
    for (int k = 0; k < 4; ++k) {
        if (!(mask & (1u << k))) continue;
        uint32_t p, n;
        tap(r, W, k, p, n);
        if (bmin[p] != r.blob) {
            const int slot = atomicSub(&ovcount[p], 1) - 1;
            Ov o = {r.blob, n, r.col, 0u};
            ov[ovstart[p] + slot] = o;
        }
    }
}

// ------------------------------------------------------------ composite
struct Px {  // one resolved (pixel, blob) colour
    uint32_t r, g, b, a;
};

// morph.cpp:591-623 with exact integer sums: round(sum c*n / sum n), alpha
// times min(1, count / density).
__device__ Px resolve(uint64_t cnt, uint64_t sn, const uint64_t s[4], int density) {
    Px q;
    const uint64_t d2 = 2 * sn;
    q.r = uint32_t((2 * s[0] + sn) / d2);
    q.g = uint32_t((2 * s[1] + sn) / d2);
    q.b = uint32_t((2 * s[2] + sn) / d2);
    double w = density > 0 ? double(cnt) / double(density) : 0.0;
    if (w >= 1.0) {
        q.a = uint32_t((2 * s[3] + sn) / d2);
    } else {
        q.a = narrow8(round(w * (double(s[3]) / double(sn))));
    }
    return q;
}

struct Seg {  // overflow segment of one pixel
    const Ov *ov;
    int begin, end;
};

// Is pixel q (inside the image) a member of blob b's resolved pixel set?
__device__ bool member(const int *bmin, const int *ovstart, const Ov *ov, int64_t q, int b) {
    if (bmin[q] == b) return true;
    for (int i = ovstart[q]; i < ovstart[q + 1]; ++i)
        if (ov[i].blob == b) return true;
    return false;
}

// Inward feather (morph.cpp:625-674): the peel layer of pixel (x, y) in
// blob b is d - 1, d = L1 distance to the nearest pixel outside the blob
// (coordinates 0 and 65535 count as blob borders).  Returns d (> feather:
// untouched).
__device__ int feather_depth(int x, int y, int W, int H, int F, int b, const int *bmin, const int *ovstart,
                             const Ov *ov) {
    int best = F + 1;
    best = min(best, x + 1);
    best = min(best, y + 1);
    best = min(best, 65536 - x);
    best = min(best, 65536 - y);
    for (int r = 1; r < best; ++r) {
        // pixels at L1 distance exactly r
        for (int dx = -r; dx <= r; ++dx) {
            const int ady = r - (dx < 0 ? -dx : dx);
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                if (ady == 0 && sgn > 0) continue;
                const int qx = x + dx, qy = y + sgn * ady;
                bool in;
                if (qx < 0 || qy < 0 || qx >= W || qy >= H) in = false;
                else in = member(bmin, ovstart, ov, int64_t(qy) * W + qx, b);
                if (!in) return r;
            }
        }
    }
    return best;
}

struct Fold {
    double r = 0, g = 0, b = 0, a = 0, rs = 0, gs = 0, bs = 0, as = 0;
    int n = 0;
    __device__ void add(uint32_t c) {
        // morph.cpp:1357-1380
        const double sr = (c & 255u) / 255.0, sg = ((c >> 8) & 255u) / 255.0;
        const double sb = ((c >> 16) & 255u) / 255.0, sa = (c >> 24) / 255.0;
        as += sa;
        rs += sr * sa;
        gs += sg * sa;
        bs += sb * sa;
        if (n == 0) {
            r = sr; g = sg; b = sb; a = sa;
        } else {
            r = sa * sr + (1.0 - sa) * r;
            g = sa * sg + (1.0 - sa) * g;
            b = sa * sb + (1.0 - sa) * b;
            a = a + (1.0 - a) * ((c >> 24) / 255.0);
        }
        ++n;
    }
};


__global__ __launch_bounds__(256) void k_composite(amh_render R, Frame F, int *__restrict__ bmin, Acc acc,
                                                   const int *__restrict__ ovstart, const Ov *__restrict__ ov,
                                                   uint32_t *__restrict__ out, int clear_bmin,
                                                   const RenderParams *__restrict__ fp) {
    if (fp) {
        F = fp->F;
        out = fp->out;
    }
    __shared__ int lagp[512], slopep[512];
    if (R.keep_background && R.fading == kPerlin) {
        for (int i = threadIdx.x; i < 512; i += 256) {
            lagp[i] = R.lag_perm_dev[i];
            slopep[i] = R.slope_perm_dev[i];
        }
    }
    __syncthreads();
    const int W = R.width, H = R.height;
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= int64_t(W) * H) return;
    const int x = int(p % W), y = int(p / W);
    const uint32_t bg = R.keep_background ? background(R, F, x, y, lagp, slopep) : 0u;
    const int b0 = bmin[p];
    if (b0 == 0x7FFFFFFF) {
        out[p] = bg;
        return;
    }
    const int nb = R.nblobs;
    const bool texture = R.show_blobs != 1 && R.show_blobs != 2;
    Fold fold;
    auto emit = [&](int b, Px q) {
        if (texture && R.feather > 0) {
            const int d = feather_depth(x, y, W, H, R.feather, b, bmin, ovstart, ov);
            if (d <= R.feather) q.a = narrow8(round(double(q.a) * (double(d) / double(R.feather + 1))));
        }
        uint32_t c = q.r | (q.g << 8) | (q.b << 16) | (q.a << 24);
        if (!texture && R.blob_rgba_dev) c = R.blob_rgba_dev[int64_t(F.fk) * nb + b];
        if ((c >> 24) == 0) return;
        fold.add(c);
    };
    {
        const uint64_t cn = acc.cn[p];
        uint64_t s[4];
        for (int k = 0; k < 4; ++k) {
            s[k] = acc.s[k][p];
            acc.s[k][p] = 0;
        }
        acc.cn[p] = 0;
        emit(b0, resolve(cn >> 40, cn & ((1ull << 40) - 1), s, R.density));
    }
    const int o0 = ovstart[p], o1 = ovstart[p + 1];
    int last = b0;
    while (o1 > o0) {
        // next blob above `last`, with its sums (one pass per blob)
        int nbk = 0x7FFFFFFF;
        uint64_t cnt = 0, sn = 0, s[4] = {0, 0, 0, 0};
        for (int i = o0; i < o1; ++i) {
            const Ov e = ov[i];
            if (e.blob <= last || e.blob > nbk) continue;
            if (e.blob < nbk) {
                nbk = e.blob;
                cnt = sn = 0;
                s[0] = s[1] = s[2] = s[3] = 0;
            }
            ++cnt;
            sn += e.n;
            for (int k = 0; k < 4; ++k) s[k] += uint64_t(ch(e.col, k)) * e.n;
        }
        if (nbk == 0x7FFFFFFF) break;
        emit(nbk, resolve(cnt, sn, s, R.density));
        last = nbk;
    }
    if (clear_bmin) bmin[p] = 0x7FFFFFFF;
    if (fold.n == 0) {
        out[p] = bg;
        return;
    }
    double r = fold.r, g = fold.g, b = fold.b, a = fold.a;
    if (R.blend_blobs) {
        r = fold.rs / fold.as;
        g = fold.gs / fold.as;
        b = fold.bs / fold.as;
    }
    if (R.keep_background) {
        const double bgr = double(bg & 255u) / 255.0, bgg = double((bg >> 8) & 255u) / 255.0;
        const double bgb = double((bg >> 16) & 255u) / 255.0, bga = double(bg >> 24) / 255.0;
        r = a * r + (1.0 - a) * bgr;
        g = a * g + (1.0 - a) * bgg;
        b = a * b + (1.0 - a) * bgb;
        a = bga + (1.0 - bga) * a;
    }
    out[p] = rgba_pack(narrow8(round(r * 255.0)), narrow8(round(g * 255.0)), narrow8(round(b * 255.0)),
                       narrow8(round(a * 255.0)));
}

__global__ void k_fill_i32(int *p, int64_t n, int v) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------ workspace
struct Layout {
    int64_t npx, nblk;
    size_t bmin, ovcount, ovstart, sums, offs, acc, recs, ov, total;
};
Layout layout(int32_t W, int32_t H, int64_t wtot) {
    Layout L;
    L.npx = int64_t(W) * H;
    L.nblk = scan::blocks(L.npx);
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t o = 0;
    L.bmin = o; o = al(o + sizeof(int) * L.npx);
    L.ovcount = o; o = al(o + sizeof(int) * L.npx);
    L.ovstart = o; o = al(o + sizeof(int) * (L.npx + 1));
    L.sums = o; o = al(o + sizeof(int) * (L.nblk + 1));
    L.offs = o; o = al(o + sizeof(int) * (L.nblk + 1));
    L.acc = o; o = al(o + 5 * sizeof(uint64_t) * L.npx);
    L.recs = o; o = al(o + sizeof(Rec) * size_t(wtot > 0 ? wtot : 1));
    L.ov = o; o = al(o + sizeof(Ov) * 4 * size_t(wtot > 0 ? wtot : 1));
    L.total = o;
    return L;
}

Acc acc_of(char *ws, const Layout &L) {
    Acc a;
    unsigned long long *base = reinterpret_cast<unsigned long long *>(ws + L.acc);
    a.cn = base;
    for (int k = 0; k < 4; ++k) a.s[k] = base + (k + 1) * L.npx;
    return a;
}

}  // namespace

extern "C" size_t amh_render_workspace_bytes(int32_t W, int32_t H, int64_t wtot) {
    return layout(W, H, wtot).total;
}

extern "C" int amh_render_workspace_init(void *ws, size_t ws_bytes, int32_t W, int32_t H, int64_t wtot,
                                         void *stream) {
    const Layout L = layout(W, H, wtot);
    if (!ws || ws_bytes < L.total || W <= 0 || H <= 0) {
        set_error("amh_render_workspace_init: invalid argument");
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    char *base = static_cast<char *>(ws);
    if (int r = check_hip(hipMemsetAsync(base, 0, L.total, s), "amh_render_workspace_init")) return r;
    hipLaunchKernelGGL(k_fill_i32, dim3(uint32_t((L.npx + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<int *>(base + L.bmin), L.npx, 0x7FFFFFFF);
    AMH_LAUNCH_CHECK("amh_render_workspace_init");
    return AMH_OK;
}

namespace {

bool render_args_ok(const amh_render *R) {
    return R && R->width > 0 && R->height > 0 && R->nkeys > 0 && R->nchains >= 0 && R->keys_dev &&
           !(R->wtot > 0 && (!R->pts_dev || !R->chain_off_dev || !R->chain_w_dev || !R->rank_dev)) &&
           !(R->fading == kPerlin && (!R->lag_perm_dev || !R->slope_perm_dev)) && R->width <= 65535 &&
           R->height <= 65535 && R->wtot >= 0;
}

// The frame itself, on a workspace laid out by `L` whose accumulators,
// overflow counts and smallest-blob map are clean; leaves them clean.
// fp (device, may be NULL): the frame's time and output are read from it
// (graph capture; then nothing is timed, since events would be captured).
int render_frame_on(const amh_render *R, double t, char *base, const Layout &L, uint32_t *out, hipStream_t s,
                    const RenderParams *fp = nullptr);

}  // namespace

extern "C" int amh_render_frame(const amh_render *R, double t, void *ws, size_t ws_bytes, uint32_t *out,
                                void *stream) {
    if (!render_args_ok(R) || !ws || !out) {
        set_error("amh_render_frame: invalid argument");
        return AMH_E_ARG;
    }
    const Layout L = layout(R->width, R->height, R->wtot);
    if (ws_bytes < L.total) {
        set_error("amh_render_frame: workspace %zu < %zu bytes", ws_bytes, L.total);
        return AMH_E_ARG;
    }
    return render_frame_on(R, t, static_cast<char *>(ws), L, out, static_cast<hipStream_t>(stream));
}

// ------------------------------------------- allocation-free frame render
// The same frame render for callers that cannot allocate (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc): the
// workspace is a fixed array in device memory, one per GPU, held until the
// process ends.  Every call on a GPU shares it, so calls that overlap on one
// GPU race; each call clears the parts that must start clean with a fill
// kernel (no memset node, which graph capture would record).
namespace {

constexpr size_t kScratchBytes = size_t(96) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device.  The array's name
// is only a host-side stand-in, so kernels and copies get this pointer.
int scratch_ptr(char **out) {
    int dev = -1;
    if (int r = check_hip(hipGetDevice(&dev), "amh_render_frame_noalloc: hipGetDevice")) return r;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("amh_render_frame_noalloc: device %d beyond %d", dev, kMaxDevices);
        return AMH_E_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (int r = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                              "amh_render_frame_noalloc: hipGetSymbolAddress"))
            return r;
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<char *>(g_scratch_ptr[dev]);
    return AMH_OK;
}

// bmin := INT_MAX, overflow counts and accumulators := 0.
__global__ __launch_bounds__(256) void k_ws_reset(int *__restrict__ bmin, int *__restrict__ ovcount,
                                                  unsigned long long *__restrict__ acc, int64_t npx) {
    const int64_t stride = int64_t(gridDim.x) * 256;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < npx; i += stride) {
        bmin[i] = 0x7FFFFFFF;
        ovcount[i] = 0;
        for (int k = 0; k < 5; ++k) acc[int64_t(k) * npx + i] = 0ull;
    }
}

}  // namespace

extern "C" size_t amh_render_noalloc_max_bytes(void) { return kScratchBytes; }

extern "C" int amh_render_frame_noalloc(const amh_render *R, double t, uint32_t *out, void *stream) {
    if (!render_args_ok(R) || !out) {
        set_error("amh_render_frame_noalloc: invalid argument");
        return AMH_E_ARG;
    }
    const Layout L = layout(R->width, R->height, R->wtot);
    if (L.total > kScratchBytes) {
        set_error("amh_render_frame_noalloc: %dx%d with %lld atoms needs %zu bytes of workspace, more than the "
                  "%zu this entry point holds; use amh_render_frame",
                  R->width, R->height, (long long)R->wtot, L.total, kScratchBytes);
        return AMH_E_ARG;
    }
    char *base = nullptr;
    if (int r = scratch_ptr(&base)) return r;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t grid = uint32_t(std::min<int64_t>((L.npx + 255) / 256, 4096));
    hipLaunchKernelGGL(k_ws_reset, dim3(grid), dim3(256), 0, s, reinterpret_cast<int *>(base + L.bmin),
                       reinterpret_cast<int *>(base + L.ovcount),
                       reinterpret_cast<unsigned long long *>(base + L.acc), L.npx);
    AMH_LAUNCH_CHECK("amh_render_frame_noalloc");
    return render_frame_on(R, t, base, L, out, s);
}

namespace {
int render_frame_on(const amh_render *R, double t, char *base, const Layout &L, uint32_t *out, hipStream_t s,
                    const RenderParams *fp) {
    amh::ProfScope whole_(fp ? nullptr : "render_frame", s);
    int *bmin = reinterpret_cast<int *>(base + L.bmin);
    int *ovcount = reinterpret_cast<int *>(base + L.ovcount);
    int *ovstart = reinterpret_cast<int *>(base + L.ovstart);
    int *sums = reinterpret_cast<int *>(base + L.sums);
    int *offs = reinterpret_cast<int *>(base + L.offs);
    Rec *recs = reinterpret_cast<Rec *>(base + L.recs);
    Ov *ov = reinterpret_cast<Ov *>(base + L.ov);
    const Acc acc = acc_of(base, L);

    // morph::normalize_time / get_frame_key (morph.cpp:404-450, 460-464)
    Frame F;
    double integ;
    F.tn = modf(t, &integ);
    if (F.tn < 0.0) F.tn += 1.0;
    const size_t fk = size_t(F.tn * R->nkeys);
    amh_render Rc = *R;
    if (fk >= size_t(R->nkeys) || R->wtot == 0) {
        // nothing to draw: the background only
        Rc.wtot = 0;
    }
    F.fk = int(fk < size_t(R->nkeys) ? fk : 0);
    F.fnext = (F.fk + 1) % R->nkeys;
    const double dt = 1.0 / double(R->nkeys);
    F.tl = fmax(0.0, (F.tn - (double(F.fk) * dt)) / dt);
    const uint32_t gpx = uint32_t((L.npx + 255) / 256);
    const uint32_t gat = uint32_t((Rc.wtot + 255) / 256);
    if (Rc.wtot > 0) {
        if (fp) {
            hipLaunchKernelGGL(k_prepare, dim3(gat), dim3(256), 0, s, Rc, F, recs, bmin, fp);
            hipLaunchKernelGGL(k_splat, dim3(gat), dim3(256), 0, s, recs, Rc.wtot, Rc.width, bmin, acc, ovcount);
        } else {
            amh::launch_timed("k_prepare", k_prepare, dim3(gat), dim3(256), 0, s, Rc, F, recs, bmin, fp);
            amh::launch_timed("k_splat", k_splat, dim3(gat), dim3(256), 0, s, recs, Rc.wtot, Rc.width, bmin, acc,
                              ovcount);
        }
    }
    scan::exclusive(ovcount, L.npx, sums, offs, ovstart, s);
    if (Rc.wtot > 0)
        hipLaunchKernelGGL(k_overflow, dim3(gat), dim3(256), 0, s, recs, Rc.wtot, Rc.width, bmin, ovcount, ovstart,
                           ov);
    const int separate_clear = Rc.feather > 0 ? 1 : 0;
    if (fp)
        hipLaunchKernelGGL(k_composite, dim3(gpx), dim3(256), 0, s, Rc, F, bmin, acc, ovstart, ov, out,
                           separate_clear ? 0 : 1, fp);
    else
        amh::launch_timed("k_composite", k_composite, dim3(gpx), dim3(256), 0, s, Rc, F, bmin, acc, ovstart, ov,
                          out, separate_clear ? 0 : 1, fp);
    if (separate_clear)
        hipLaunchKernelGGL(k_fill_i32, dim3(gpx), dim3(256), 0, s, bmin, L.npx, 0x7FFFFFFF);
    AMH_LAUNCH_CHECK("amh_render_frame");
    return AMH_OK;
}
}  // namespace

// --------------------------------------------------- frames from a graph
// One frame's launches captured once and replayed for every later frame of
// the same chains and options; only the frame's time and output, in a small
// device block, change between replays.
struct amh_render_graph_s {
    amh_render R;
    char *ws = nullptr;
    Layout L;
    RenderParams *fp = nullptr;  // device
    hipGraphExec_t exec = nullptr;
};

extern "C" int amh_render_graph_create(void **graph, const amh_render *R, void *ws, size_t ws_bytes, void *stream) {
    if (!graph || !render_args_ok(R) || !ws || !stream) {
        set_error("amh_render_graph_create: invalid argument (a created stream is required)");
        return AMH_E_ARG;
    }
    *graph = nullptr;
    const Layout L = layout(R->width, R->height, R->wtot);
    if (ws_bytes < L.total) {
        set_error("amh_render_graph_create: workspace %zu < %zu bytes", ws_bytes, L.total);
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    amh_render_graph_s *g = new amh_render_graph_s;
    g->R = *R;
    g->ws = static_cast<char *>(ws);
    g->L = L;
    if (int r = check_hip(hipMalloc(&g->fp, sizeof(RenderParams)), "amh_render_graph_create")) {
        delete g;
        return r;
    }
    int r = check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (!r) {
        // time 0 only fixes which launches the graph holds (a frame with
        // atoms); every replay reads its own frame from g->fp
        const int rr = render_frame_on(&g->R, 0.0, g->ws, g->L, nullptr, s, g->fp);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &gr);
        r = rr ? rr : check_hip(e, "hipStreamEndCapture");
        if (!r) r = check_hip(hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0), "hipGraphInstantiate");
        if (gr) (void)hipGraphDestroy(gr);
    }
    if (r) {
        (void)hipFree(g->fp);
        delete g;
        return r;
    }
    *graph = g;
    return AMH_OK;
}

extern "C" int amh_render_graph_frame(void *graph, double t, uint32_t *out, void *stream) {
    amh_render_graph_s *g = static_cast<amh_render_graph_s *>(graph);
    if (!g || !out) {
        set_error("amh_render_graph_frame: invalid argument");
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    RenderParams v;
    v.F = frame_of(t, g->R.nkeys);
    v.out = out;
    // a frame with nothing to draw (no atoms, or a time past the keys)
    // launches fewer kernels than the graph holds: it runs eagerly
    if (g->R.wtot == 0 || v.F.fk >= g->R.nkeys) return render_frame_on(&g->R, t, g->ws, g->L, out, s);
    hipLaunchKernelGGL(k_set_render_params, dim3(1), dim3(1), 0, s, g->fp, v);
    if (int r = check_hip(hipGetLastError(), "amh_render_graph_frame")) return r;
    return check_hip(hipGraphLaunch(g->exec, s), "amh_render_graph_frame");
}

extern "C" int amh_render_graph_destroy(void *graph) {
    amh_render_graph_s *g = static_cast<amh_render_graph_s *>(graph);
    if (!g) return AMH_OK;
    (void)hipDeviceSynchronize();  // replays may still be queued on the caller's stream
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    (void)hipFree(g->fp);
    delete g;
    return AMH_OK;
}
