// Shading helpers shared by the atom and fluid renderers: Perlin fBm, the
// colour fades, key-frame lookups and the cross-dissolved background, in
// FP64 with the reference's operation order.
#pragma once
#include "colour.h"
#include "common.h"
#include "../../include/atomorph_hip.h"

namespace amh {
namespace shade {

constexpr int kNone = 2, kLinear = 3, kSpline = 4, kCosine = 5, kPerlin = 6;

struct Frame {  // per-frame scalars, computed on the host
    double tn;   // normalised global time
    double tl;   // time within the key segment
    int fk;      // key frame (row y)
    int fnext;   // next key frame (row y + 1)
};

// ------------------------------------------------------------- Perlin
// perlin.cpp:19-69 on the z = 0 plane.  With z = 0 the reference's outer
// lerp has weight fade(0) = 0 and returns its first argument unchanged, so
// the z = 1 half never contributes and is not evaluated.
__device__ __forceinline__ double pfade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
__device__ __forceinline__ double plerp(double t, double a, double b) { return a + t * (b - a); }
__device__ __forceinline__ double pgrad(int hash, double x, double y) {
    const int h = hash & 15;
    const double u = h < 8 ? x : y, v = h < 4 ? y : (h == 12 || h == 14) ? x : 0.0;
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}
__device__ double pnoise(const int *p, double x, double y) {
    const int X = static_cast<int>(floor(x)) & 255;
    const int Y = static_cast<int>(floor(y)) & 255;
    x -= floor(x);
    y -= floor(y);
    const double u = pfade(x), v = pfade(y);
    const int A = p[X] + Y, AA = p[A], AB = p[A + 1];
    const int B = p[X + 1] + Y, BA = p[B], BB = p[B + 1];
    return plerp(v, plerp(u, pgrad(p[AA], x, y), pgrad(p[BA], x - 1, y)),
                 plerp(u, pgrad(p[AB], x, y - 1), pgrad(p[BB], x - 1, y - 1)));
}
// Out of line (scalar arguments): inlined into a kernel, its octaves' table
// lookups are scheduled together with the caller's values and the kernel
// takes about forty more registers.
__device__ __noinline__ double poctave(const int *p, double x, double y) {
    double result = 0.0, amp = 1.0;
    for (int i = 0; i < 8; ++i) {
        result += pnoise(p, x, y) * amp;
        x *= 2.0;
        y *= 2.0;
        amp *= 0.5;
    }
    return result;
}

// ------------------------------------------------------------- colours
__device__ __forceinline__ uint32_t ch(uint32_t c, int k) { return (c >> (8 * k)) & 255u; }

// morph.cpp:1481-1488
__device__ uint32_t mix(uint32_t c1, uint32_t c2, double w1) {
    uint32_t o = 0;
    for (int k = 0; k < 4; ++k)
        o |= narrow8(round(w1 * double(ch(c1, k)) + (1.0 - w1) * double(ch(c2, k)))) << (8 * k);
    return o;
}
// morph.cpp:1467-1479 (cosine ramp when fading is COSINE or PERLIN)
__device__ uint32_t fade_mix(uint32_t c1, uint32_t c2, double lag, double slope, double str, int fading) {
    if (fading == kCosine || fading == kPerlin) {
        const double pi = 3.14159265358;
        const double s = (slope + 0.1) / 1.1;
        const double l = (1.0 - s) * lag;
        if (str <= l) str = 0.0;
        else if (str >= (l + s)) str = 1.0;
        else str = ((-cos((str - l) * (pi / s)) + 1.0) / 2.0);
    }
    return mix(c1, c2, str);
}
// The colour fade of get_pixels / get_background for perlin coordinates
// (px, py) (only used for PERLIN).
__device__ uint32_t fade_colour(uint32_t c1, uint32_t c2, double px, double py, double str, int fading,
                                const int *lagp, const int *slopep) {
    if (fading == kPerlin) {
        const double lag = poctave(lagp, px, py) * 0.5 + 0.5;
        const double slope = poctave(slopep, px, py) * 0.5 + 0.5;
        return fade_mix(c1, c2, lag, slope, str, fading);
    }
    if (fading == kCosine) return fade_mix(c1, c2, 0.5, 0.5, str, fading);
    return mix(c1, c2, str);
}

__device__ __forceinline__ uint32_t key_px(const uint32_t *keys, int f, int W, int H, uint32_t x, uint32_t y) {
    // branch-free: a pixel outside the frame reads the frame's first pixel
    // and returns 0, so the load is never waited for inside a branch
    const bool in = x < uint32_t(W) && y < uint32_t(H);
    const uint32_t c = keys[int64_t(f) * W * H + (in ? int64_t(y) * W + x : 0)];
    return (in && (c >> 24)) ? c : 0u;
}


__device__ uint32_t background(const amh_render &R, const Frame &F, int x, int y, const int *lagp,
                               const int *slopep) {
    const uint32_t c1 = key_px(R.keys_dev, F.fk, R.width, R.height, x, y);
    const uint32_t c2 = key_px(R.keys_dev, F.fnext, R.width, R.height, x, y);
    double px = 0.0, py = 0.0;
    if (R.fading == kPerlin) {
        const double bw = R.bbox_x2 - R.bbox_x1 + 1.0, bh = R.bbox_y2 - R.bbox_y1 + 1.0;
        px = ((x - R.bbox_x1) / double(bw)) * 8.0;
        py = ((y - R.bbox_y1) / double(bh)) * 8.0;
    }
    return fade_colour(c1, c2, px, py, 1.0 - F.tl, R.fading, lagp, slopep);
}

// Trajectory point of atom a at the frame's time (morph.cpp:525-531 and
// 1501-1515): NONE keeps the source point, LINEAR interpolates the fixed
// point coordinates (truncating), SPLINE evaluates the closed Catmull-Rom
// through the atom's key positions (spline.cpp:29-57) and rounds the
// fraction to /255.
// The key points a SPLINE trajectory reads (the four around the frame's
// segment); other motions read none.  Loaded apart from traj_point so that a
// caller can issue them with its other loads.
__device__ __forceinline__ void traj_keys(const amh_render &R, const Frame &F, int64_t a, uint64_t q[4]) {
    if (R.motion != kSpline) return;
    const int h = R.nkeys;
    const double dt = 1.0 / h;
    const int pk = (int)(F.tn / dt);
    auto wrap = [h](int i) { return i < 0 ? h - 1 : (i >= h ? i - h : i); };
    for (int k = 0; k < 4; ++k) q[k] = R.pts_dev[int64_t(wrap(pk - 1 + k)) * R.wtot + a];
}

// traj_keys from the atom's own record ka[nkeys] (its key points
// atom-major).
__device__ __forceinline__ void traj_keys_of(const amh_render &R, const Frame &F, const uint64_t *ka, uint64_t q[4]) {
    if (R.motion != kSpline) return;
    const int h = R.nkeys;
    const double dt = 1.0 / h;
    const int pk = (int)(F.tn / dt);
    auto wrap = [h](int i) { return i < 0 ? h - 1 : (i >= h ? i - h : i); };
    for (int k = 0; k < 4; ++k) q[k] = ka[wrap(pk - 1 + k)];
}

__device__ inline void traj_point(const amh_render &R, const Frame &F, int64_t a, uint64_t p1, uint64_t p2,
                                  uint32_t &qx, uint32_t &qy, uint32_t &qxf, uint32_t &qyf,
                                  const uint64_t *keys = nullptr) {
    const int h = R.nkeys;
    qx = pt_x(p1); qy = pt_y(p1); qxf = pt_xf(p1); qyf = pt_yf(p1);
    if (R.motion == kLinear) {
        const double w1 = 1.0 - F.tl;
        const double x1 = 256.0 * pt_x(p1) + pt_xf(p1), y1 = 256.0 * pt_y(p1) + pt_yf(p1);
        const double x2 = 256.0 * pt_x(p2) + pt_xf(p2), y2 = 256.0 * pt_y(p2) + pt_yf(p2);
        const double xx = (w1 * x1 + (1.0 - w1) * x2), yy = (w1 * y1 + (1.0 - w1) * y2);
        qx = narrow16(xx / 256.0);
        qxf = narrow8(xx - double(int(qx) * 256));
        qy = narrow16(yy / 256.0);
        qyf = narrow8(yy - double(int(qy) * 256));
    } else if (R.motion == kSpline) {
        const double dt = 1.0 / h;
        const int pk = (int)(F.tn / dt);
        const double lt = (F.tn - dt * (double)pk) / dt;
        const double t2 = lt * lt, t3 = t2 * lt;
        const double b1 = 0.5 * (-t3 + 2.0 * t2 - lt);
        const double b2 = 0.5 * (3.0 * t3 - 5.0 * t2 + 2.0);
        const double b3 = 0.5 * (-3.0 * t3 + 4.0 * t2 + lt);
        const double b4 = 0.5 * (t3 - t2);
        uint64_t q[4];
        if (keys) {
            for (int k = 0; k < 4; ++k) q[k] = keys[k];
        } else {
            traj_keys(R, F, a, q);
        }
        const uint64_t q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
        auto cx = [](uint64_t q) { return pt_x(q) + pt_xf(q) / double(256); };
        auto cy = [](uint64_t q) { return pt_y(q) + pt_yf(q) / double(256); };
        const double sx = cx(q0) * b1 + cx(q1) * b2 + cx(q2) * b3 + cx(q3) * b4;
        const double sy = cy(q0) * b1 + cy(q1) * b2 + cy(q2) * b3 + cy(q3) * b4;
        double ig;
        double fr = modf(sx, &ig);
        qx = narrow16(ig);
        qxf = narrow8(round(fr * 255));
        fr = modf(sy, &ig);
        qy = narrow16(ig);
        qyf = narrow8(round(fr * 255));
    }
}

// The PERLIN fade's lag and slope at Perlin coordinates (px, py).
__device__ __forceinline__ double2 lag_slope_at(double px, double py, const int *lagp, const int *slopep) {
    return double2{poctave(lagp, px, py) * 0.5 + 0.5, poctave(slopep, px, py) * 0.5 + 0.5};
}

// The PERLIN fade's lag and slope of an atom with source point p1
// (morph.cpp:537-550): the Perlin coordinates follow p1 over the bounding
// box.  They depend on p1, the bounding box and the two tables only, not on
// the time.
__device__ __forceinline__ void atom_lag_slope(const amh_render &R, uint64_t p1, const int *lagp, const int *slopep,
                                               double &lag, double &slope) {
    const double bw = R.bbox_x2 - R.bbox_x1 + 1.0, bh = R.bbox_y2 - R.bbox_y1 + 1.0;
    const double perx = (((int(pt_x(p1)) - R.bbox_x1) * 256 + int(pt_xf(p1))) / double(bw * 256)) * 8.0;
    const double pery = (((int(pt_y(p1)) - R.bbox_y1) * 256 + int(pt_yf(p1))) / double(bh * 256)) * 8.0;
    const double2 v = lag_slope_at(perx, pery, lagp, slopep);
    lag = v.x;
    slope = v.y;
}

// Colour of an atom at the frame's time (morph.cpp:537-550); for PERLIN
// with its lag and slope (atom_lag_slope) given.
__device__ inline uint32_t atom_colour_ls(const amh_render &R, const Frame &F, uint32_t c1, uint32_t c2, double lag,
                                          double slope) {
    const double str = 1.0 - F.tl;
    if (R.fading == kPerlin) return fade_mix(c1, c2, lag, slope, str, R.fading);
    if (R.fading == kCosine) return fade_mix(c1, c2, 0.5, 0.5, str, R.fading);
    return mix(c1, c2, str);
}

__device__ inline uint32_t atom_colour(const amh_render &R, const Frame &F, uint64_t p1, uint32_t c1, uint32_t c2,
                                       const int *lagp, const int *slopep) {
    double lag = 0.5, slope = 0.5;
    if (R.fading == kPerlin) atom_lag_slope(R, p1, lagp, slopep, lag, slope);
    return atom_colour_ls(R, F, c1, c2, lag, slope);
}

__device__ __forceinline__ int find_chain(const int64_t *off, int nchains, int64_t a) {
    int lo = 0, hi = nchains - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= a) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Chain of atom a (the last c with off[c] <= a; off ascending from 0) for a
// workgroup of consecutive atoms starting at a0 = blockIdx.x * blockDim.x.
// The chain of a0 is found by a blockDim-way search: each round every
// thread probes one offset and __syncthreads_count of the probes at or
// below a0 narrows the range blockDim-fold -- one round, one load per thread
// for up to blockDim chains, instead of a chain of log2(chains) dependent
// loads by one thread.  Then each thread walks forward over the few chains
// that start inside the workgroup.  Every thread of the workgroup calls it.
__device__ __forceinline__ int block_chain(const int64_t *off, int nchains, int64_t a) {
    const int64_t a0 = int64_t(blockIdx.x) * blockDim.x;
    const int nt = int(blockDim.x), t = int(threadIdx.x);
    int lo = 0, n = nchains;  // the chain of a0 is in [lo, lo + n); off[lo] <= a0
    while (n > 1) {
        const int stride = (n + nt - 1) / nt;
        const bool le = t * stride < n && off[lo + t * stride] <= a0;
        const int below = __syncthreads_count(le);  // probes at or below a0: a prefix, >= 1
        lo += (below - 1) * stride;
        n = min(stride, n - (below - 1) * stride);
    }
    int c = lo;
    while (c + 1 < nchains && off[c + 1] <= a) ++c;
    return c;
}

// Host side: frame scalars of global time t (morph::normalize_time,
// get_frame_key, morph.cpp:404-464); fk = nkeys when t maps past the keys.
inline Frame frame_of(double t, int nkeys) {
    Frame F;
    double integ;
    F.tn = modf(t, &integ);
    if (F.tn < 0.0) F.tn += 1.0;
    const size_t fk = size_t(F.tn * nkeys);
    F.fk = int(fk < size_t(nkeys) ? fk : size_t(nkeys));
    F.fnext = (F.fk + 1) % (nkeys > 0 ? nkeys : 1);
    const double dt = 1.0 / double(nkeys);
    F.tl = fmax(0.0, (F.tn - (double(F.fk < nkeys ? F.fk : 0) * dt)) / dt);
    return F;
}

}  // namespace shade
}  // namespace amh
