// Fluid path of the renderer: morph::step_fluid + FluidModel::step +
// morph::draw_fluid (morph.cpp:807-1300, fluidmodel.cpp:165-580).
//
// Particle i follows atom i of the chains.  In key segment y -> y+1 a chain
// keeps round((1-t)|A| + t|B|) particles active (|A|, |B| = its blob sizes,
// morph.cpp:836-839); which atoms carry them is fixed per segment by a rank
// (atoms whose both ends hold fluid first, then source-only, then
// destination-only).  Each active particle is attracted to its atom's
// trajectory point and takes its colour (update_particle, morph.cpp:680-805).
//
// FluidModel::step as six grid-synchronised phases: P2G scatter (mass,
// density, density gradient, strength-weighted colour), pressure and wall
// forces, node acceleration, G2P with the trajectory attractor and momentum
// scatter, node velocity, advection with colour diffusion, freedom clamp and
// walls.  Particle state is FP64; the grid accumulates in FP32 atomics.
#include "colour.h"
#include "common.h"
#include "../../include/atomorph_hip.h"
#include "shade.h"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace amh;
using namespace amh::shade;

namespace {

constexpr int kT = 256;

struct Particles {  // structure of arrays, device
    double *x, *y, *u, *v, *gx, *gy, *fr;
    double *R, *G, *B, *A, *r, *g, *b, *a;
    double *strength;
    int32_t *active, *mature, *owner, *key;  // owner: first at its source
};

struct Grid {  // FP32 node fields, [gsx * gsy], node (i, j) at i * gsy + j
    float *m, *d, *gx, *gy, *u, *v, *ax, *ay, *r, *g, *b, *a, *w;
};

__device__ __forceinline__ void weights(double x, double y, int &cx, int &cy, double px[3], double py[3],
                                        double gx[3], double gy[3]) {
    // quadratic B-spline weights and gradients (fluidmodel.cpp:191-215)
    cx = static_cast<int>(x - 0.5);
    cy = static_cast<int>(y - 0.5);
    double t = cx - x;
    px[0] = (0.5 * t * t + 1.5 * t + 1.125); gx[0] = (t + 1.5);
    t += 1.0;
    px[1] = (-t * t + 0.75); gx[1] = (-2.0 * t);
    t += 1.0;
    px[2] = (0.5 * t * t - 1.5 * t + 1.125); gx[2] = (t - 1.5);
    t = cy - y;
    py[0] = (0.5 * t * t + 1.5 * t + 1.125); gy[0] = (t + 1.5);
    t += 1.0;
    py[1] = (-t * t + 0.75); gy[1] = (-2.0 * t);
    t += 1.0;
    py[2] = (0.5 * t * t - 1.5 * t + 1.125); gy[2] = (t - 1.5);
}

__device__ __forceinline__ bool node_ok(int i, int j, int gsx, int gsy) { return i >= 0 && j >= 0 && i < gsx && j < gsy; }

__global__ void k_p2g(Particles P, int n, Grid Gd, int gsx, int gsy) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n || !P.active[i]) return;
    int cx, cy;
    double px[3], py[3], gx[3], gy[3];
    weights(P.x[i], P.y[i], cx, cy, px, py, gx, gy);
    const bool col = P.mature[i] && P.strength[i] > 0.0;
    const double s = P.strength[i];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            if (!node_ok(cx + a, cy + b, gsx, gsy)) continue;
            const int64_t k = int64_t(cx + a) * gsy + cy + b;
            const double phi = px[a] * py[b];
            atomicAdd(&Gd.m[k], float(phi));
            atomicAdd(&Gd.d[k], float(phi));
            atomicAdd(&Gd.gx[k], float(gx[a] * py[b]));
            atomicAdd(&Gd.gy[k], float(px[a] * gy[b]));
            if (col) {
                // strength-weighted running mean == sum(s c) / sum(s)
                atomicAdd(&Gd.r[k], float(s * P.r[i]));
                atomicAdd(&Gd.g[k], float(s * P.g[i]));
                atomicAdd(&Gd.b[k], float(s * P.b[i]));
                atomicAdd(&Gd.a[k], float(s * P.a[i]));
                atomicAdd(&Gd.w[k], float(s));
            }
        }
}

__device__ __forceinline__ float nd(const float *f, int i, int j, int gsx, int gsy) {
    return node_ok(i, j, gsx, gsy) ? f[int64_t(i) * gsy + j] : 0.0f;
}

// pressure from the bicubic density interpolation, wall forces
// (fluidmodel.cpp:251-357)
__global__ void k_forces(Particles P, int n, Grid Gd, int gsx, int gsy) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n || !P.active[i]) return;
    const double x = P.x[i], y = P.y[i];
    const int cx = static_cast<int>(x), cy = static_cast<int>(y);
    const double d01 = nd(Gd.d, cx, cy, gsx, gsy), d02 = nd(Gd.d, cx, cy + 1, gsx, gsy);
    const double d11 = nd(Gd.d, cx + 1, cy, gsx, gsy), d12 = nd(Gd.d, cx + 1, cy + 1, gsx, gsy);
    const double gx01 = nd(Gd.gx, cx, cy, gsx, gsy), gx02 = nd(Gd.gx, cx, cy + 1, gsx, gsy);
    const double gx11 = nd(Gd.gx, cx + 1, cy, gsx, gsy), gx12 = nd(Gd.gx, cx + 1, cy + 1, gsx, gsy);
    const double gy01 = nd(Gd.gy, cx, cy, gsx, gsy), gy02 = nd(Gd.gy, cx, cy + 1, gsx, gsy);
    const double gy11 = nd(Gd.gy, cx + 1, cy, gsx, gsy), gy12 = nd(Gd.gy, cx + 1, cy + 1, gsx, gsy);
    const double pdx = d11 - d01, pdy = d02 - d01;
    const double C20 = 3.0 * pdx - gx11 - 2.0 * gx01;
    const double C02 = 3.0 * pdy - gy02 - 2.0 * gy01;
    const double C30 = -2.0 * pdx + gx11 + gx01;
    const double C03 = -2.0 * pdy + gy02 + gy01;
    const double csum1 = d01 + gy01 + C02 + C03;
    const double csum2 = d01 + gx01 + C20 + C30;
    const double C21 = 3.0 * d12 - 2.0 * gx02 - gx12 - 3.0 * csum1 - C20;
    const double C31 = -2.0 * d12 + gx02 + gx12 + 2.0 * csum1 - C30;
    const double C12 = 3.0 * d12 - 2.0 * gy11 - gy12 - 3.0 * csum2 - C02;
    const double C13 = -2.0 * d12 + gy11 + gy12 + 2.0 * csum2 - C03;
    const double C11 = gx02 - C13 - C12 - gx01;
    const double u = x - cx, u2 = u * u, u3 = u * u2;
    const double v = y - cy, v2 = v * v, v3 = v * v2;
    const double density = d01 + gx01 * u + gy01 * v + C20 * u2 + C02 * v2 + C30 * u3 + C03 * v3 + C21 * u2 * v +
                           C31 * u3 * v + C12 * u * v2 + C13 * u * v3 + C11 * u * v;
    double pressure = density - 1.0;
    if (pressure > 2.0) pressure = 2.0;
    double fx = 0.0, fy = 0.0;
    if (x < 4.0) fx += (4.0 - x);
    else if (x > gsx - 5) fx += (gsx - 5 - x);
    if (y < 4.0) fy += (4.0 - y);
    else if (y > gsy - 5) fy += (gsy - 5 - y);
    int pcx, pcy;
    double px[3], py[3], gx[3], gy[3];
    weights(x, y, pcx, pcy, px, py, gx, gy);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            if (!node_ok(pcx + a, pcy + b, gsx, gsy)) continue;
            const int64_t k = int64_t(pcx + a) * gsy + pcy + b;
            const double phi = px[a] * py[b];
            atomicAdd(&Gd.ax[k], float(-((gx[a] * py[b]) * pressure) + fx * phi));
            atomicAdd(&Gd.ay[k], float(-((px[a] * gy[b]) * pressure) + fy * phi));
        }
}

__global__ void k_node_div(const float *m, float *f1, float *f2, int64_t nn) {
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k >= nn) return;
    const float mm = m[k];
    if (mm > 0.0f) {
        f1[k] /= mm;
        f2[k] /= mm;
    }
}

// Move (x1, y1) towards (x2, y2) by `a` (capped at the distance), the
// construction used by the attractor and the freedom clamp.
__device__ __forceinline__ void toward(double x1, double y1, double x2, double y2, double a, double &dx_out,
                                       double &dy_out) {
    dx_out = dy_out = 0.0;
    const double A = fabs(y1 - y2), B = fabs(x1 - x2), Cd = sqrt(A * A + B * B);
    if (a >= Cd) a = Cd;
    if (B <= 0.0) {
        if (y2 <= y1) dy_out -= a;
        else dy_out += a;
    } else if (Cd > 0.0) {
        const double dx = (a * B) / Cd, dy = (A * dx) / B;
        if (x1 <= x2) dx_out += dx;
        else dx_out -= dx;
        if (y1 <= y2) dy_out += dy;
        else dy_out -= dy;
    }
}

__global__ void k_g2p(Particles P, int n, Grid Gd, int gsx, int gsy) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n || !P.active[i]) return;
    int cx, cy;
    double px[3], py[3], gx[3], gy[3];
    const double x = P.x[i], y = P.y[i];
    weights(x, y, cx, cy, px, py, gx, gy);
    double cax, cay;  // attractor, 0.03 per step (fluidmodel.cpp:383-410)
    toward(x, y, P.gx[i], P.gy[i], 0.03, cax, cay);
    double u = P.u[i], v = P.v[i];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const double phi = px[a] * py[b];
            const double ax = nd(Gd.ax, cx + a, cy + b, gsx, gsy), ay = nd(Gd.ay, cx + a, cy + b, gsx, gsy);
            u += phi * (ax + cax);
            v += phi * (ay + cay);
        }
    P.u[i] = u;
    P.v[i] = v;
    double mu = u, mv = v;
    if (!P.mature[i]) {
        mu *= 0.0;
        mv *= 0.0;
    }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            if (!node_ok(cx + a, cy + b, gsx, gsy)) continue;
            const int64_t k = int64_t(cx + a) * gsy + cy + b;
            const double phi = px[a] * py[b];
            atomicAdd(&Gd.u[k], float(phi * mu));
            atomicAdd(&Gd.v[k], float(phi * mv));
        }
}

__device__ __forceinline__ double jitter(uint64_t seed, int i, int k) {
    // replaces rand() / RAND_MAX * 0.01 of the wall bounce (fluidmodel.cpp:554-564)
    return double(mix64(seed ^ (uint64_t(i) << 3) ^ uint64_t(k)) >> 11) * (1.0 / 9007199254740992.0) * 0.01;
}

__global__ void k_advect(Particles P, int n, Grid Gd, int gsx, int gsy, int steps_left, double freedom_radius,
                         uint64_t seed) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n || !P.active[i]) return;
    int cx, cy;
    double px[3], py[3], gx[3], gy[3];
    double x = P.x[i], y = P.y[i];
    weights(x, y, cx, cy, px, py, gx, gy);
    double gu = 0, gv = 0, nR = 0, nG = 0, nB = 0, nA = 0, wsum = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            if (!node_ok(cx + a, cy + b, gsx, gsy)) continue;
            const int64_t k = int64_t(cx + a) * gsy + cy + b;
            const double phi = px[a] * py[b];
            gu += phi * Gd.u[k];
            gv += phi * Gd.v[k];
            const double w = Gd.w[k];
            if (w > 0.0) {
                // node colour = sum(s c)/sum(s); weighted by sum(s)
                wsum += w;
                nR += Gd.r[k];
                nG += Gd.g[k];
                nB += Gd.b[k];
                nA += Gd.a[k];
            }
        }
    if (wsum > 0.0) {  // colour diffusion (fluidmodel.cpp:470-495)
        nR /= wsum; nG /= wsum; nB /= wsum; nA /= wsum;
        if (!P.mature[i]) {
            P.r[i] = nR; P.g[i] = nG; P.b[i] = nB; P.a[i] = nA;
        } else {
            const double wr = fabs(nR - P.R[i]), wg = fabs(nG - P.G[i]), wb = fabs(nB - P.B[i]),
                         wa = fabs(nA - P.A[i]);
            P.r[i] = (1.0 - wr) * P.r[i] + wr * nR;
            P.g[i] = (1.0 - wg) * P.g[i] + wg * nG;
            P.b[i] = (1.0 - wb) * P.b[i] + wb * nB;
            P.a[i] = (1.0 - wa) * P.a[i] + wa * nA;
        }
    }
    x += gu;
    y += gv;
    {   // freedom radius (fluidmodel.cpp:501-539)
        const double A = fabs(y - P.gy[i]), B = fabs(x - P.gx[i]), Cd = sqrt(A * A + B * B);
        const double r = freedom_radius * P.fr[i];
        if (Cd > r) {
            double mx, my;
            toward(x, y, P.gx[i], P.gy[i], Cd - r, mx, my);
            const double w = 1.0 / (steps_left + 1);
            x += mx * w;
            y += my * w;
        }
    }
    double u = P.u[i], v = P.v[i];
    u += gu - u;
    v += gv - v;
    if (x < 1.0) { x = 1.0 + jitter(seed, i, 0); u = 0.0; }
    else if (x > gsx - 2) { x = gsx - 2 - jitter(seed, i, 1); u = 0.0; }
    if (y < 1.0) { y = 1.0 + jitter(seed, i, 2); v = 0.0; }
    else if (y > gsy - 2) { y = gsy - 2 - jitter(seed, i, 3); v = 0.0; }
    P.x[i] = x; P.y[i] = y; P.u[i] = u; P.v[i] = v;
}

__global__ void k_fill_f32(float *p, int64_t n) {
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k < n) p[k] = 0.0f;
}

}  // namespace

// ------------------------------------------------------------- the model
struct amh_fluid_s {
    int32_t gsx = 0, gsy = 0, count = 0;
    void *buf = nullptr;       // one allocation for particles + grid + render
    Particles P;
    Grid G;
    // render buffers
    double *acc = nullptr;     // [npx][5] sum w, sum w*c
    int32_t *mask = nullptr;   // [npx] blob membership / feather state
    uint32_t *img = nullptr;   // [npx] colours during feather
    int32_t W = 0, H = 0;
    uint64_t frame = 0;        // frames rendered (jitter seed)
};

namespace {
size_t fluid_bytes(int gsx, int gsy, int count, int W, int H, int nkeys) {
    const size_t n = size_t(count), nn = size_t(gsx) * gsy, npx = size_t(W) * H;
    (void)nkeys;
    return n * 8 * 16 + n * 4 * 4 + nn * 4 * 13 + npx * 8 * 5 + npx * 4 * 2 + 8192;
}
}  // namespace

extern "C" int amh_fluid_create(void **fluid, int32_t gsx, int32_t gsy, int32_t count, int32_t width, int32_t height,
                                int32_t nkeys, void *stream) {
    if (!fluid || gsx < 8 || gsy < 8 || count < 0 || width <= 0 || height <= 0 || nkeys <= 0) {
        set_error("amh_fluid_create: invalid argument");
        return AMH_E_ARG;
    }
    amh_fluid_s *f = new amh_fluid_s;
    f->gsx = gsx; f->gsy = gsy; f->count = count; f->W = width; f->H = height;
    const size_t bytes = fluid_bytes(gsx, gsy, count, width, height, nkeys);
    if (int r = check_hip(hipMalloc(&f->buf, bytes), "amh_fluid_create")) {
        delete f;
        return r;
    }
    char *p = static_cast<char *>(f->buf);
    auto take = [&](size_t b) { char *q = p; p += (b + 255) & ~size_t(255); return q; };
    const size_t n = size_t(count), nn = size_t(gsx) * gsy, npx = size_t(width) * height;
    double **d[] = {&f->P.x, &f->P.y, &f->P.u, &f->P.v, &f->P.gx, &f->P.gy, &f->P.fr, &f->P.R, &f->P.G, &f->P.B,
                    &f->P.A, &f->P.r, &f->P.g, &f->P.b, &f->P.a, &f->P.strength};
    for (double **q : d) *q = reinterpret_cast<double *>(take(n * 8));
    f->P.active = reinterpret_cast<int32_t *>(take(n * 4));
    f->P.mature = reinterpret_cast<int32_t *>(take(n * 4));
    f->P.owner = reinterpret_cast<int32_t *>(take(n * 4));
    f->P.key = reinterpret_cast<int32_t *>(take(n * 4));
    float **g[] = {&f->G.m, &f->G.d, &f->G.gx, &f->G.gy, &f->G.u, &f->G.v, &f->G.ax, &f->G.ay,
                   &f->G.r, &f->G.g, &f->G.b, &f->G.a, &f->G.w};
    float *grid = reinterpret_cast<float *>(take(nn * 4 * 13));  // contiguous: cleared in one pass
    for (int k = 0; k < 13; ++k) *g[k] = grid + size_t(k) * nn;
    f->acc = reinterpret_cast<double *>(take(npx * 8 * 5));
    f->mask = reinterpret_cast<int32_t *>(take(npx * 4));
    f->img = reinterpret_cast<uint32_t *>(take(npx * 4));
    if (int r = check_hip(hipMemsetAsync(f->buf, 0, bytes, static_cast<hipStream_t>(stream)), "amh_fluid_create")) {
        (void)hipFree(f->buf);
        delete f;
        return r;
    }
    *fluid = f;
    return AMH_OK;
}

extern "C" int amh_fluid_destroy(void *fluid) {
    amh_fluid_s *f = static_cast<amh_fluid_s *>(fluid);
    if (!f) return AMH_OK;
    if (f->buf) (void)hipFree(f->buf);
    delete f;
    return AMH_OK;
}

// Particle state in the oracle's 20-double layout (x y u v gravity_x
// gravity_y freedom_r R G B A r g b a strength active mature - -).
extern "C" int amh_fluid_set_particles(void *fluid, const double *host, void *stream) {
    amh_fluid_s *f = static_cast<amh_fluid_s *>(fluid);
    if (!f || !host) { set_error("amh_fluid_set_particles: NULL"); return AMH_E_ARG; }
    const size_t n = size_t(f->count);
    std::vector<double> col(n);
    std::vector<int32_t> ic(n);
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *dst[] = {f->P.x, f->P.y, f->P.u, f->P.v, f->P.gx, f->P.gy, f->P.fr, f->P.R, f->P.G, f->P.B,
                     f->P.A, f->P.r, f->P.g, f->P.b, f->P.a, f->P.strength};
    for (int k = 0; k < 16; ++k) {
        for (size_t i = 0; i < n; ++i) col[i] = host[20 * i + k];
        if (int r = check_hip(hipMemcpyAsync(dst[k], col.data(), n * 8, hipMemcpyHostToDevice, s), "set")) return r;
        if (int r = check_hip(hipStreamSynchronize(s), "set")) return r;
    }
    int32_t *idst[] = {f->P.active, f->P.mature};
    for (int k = 0; k < 2; ++k) {
        for (size_t i = 0; i < n; ++i) ic[i] = host[20 * i + 16 + k] != 0.0 ? 1 : 0;
        if (int r = check_hip(hipMemcpyAsync(idst[k], ic.data(), n * 4, hipMemcpyHostToDevice, s), "set")) return r;
        if (int r = check_hip(hipStreamSynchronize(s), "set")) return r;
    }
    return AMH_OK;
}

extern "C" int amh_fluid_get_particles(void *fluid, double *host, void *stream) {
    amh_fluid_s *f = static_cast<amh_fluid_s *>(fluid);
    if (!f || !host) { set_error("amh_fluid_get_particles: NULL"); return AMH_E_ARG; }
    const size_t n = size_t(f->count);
    std::vector<double> col(n);
    std::vector<int32_t> ic(n);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const double *src[] = {f->P.x, f->P.y, f->P.u, f->P.v, f->P.gx, f->P.gy, f->P.fr, f->P.R, f->P.G, f->P.B,
                           f->P.A, f->P.r, f->P.g, f->P.b, f->P.a, f->P.strength};
    for (int k = 0; k < 16; ++k) {
        if (int r = check_hip(hipMemcpyAsync(col.data(), src[k], n * 8, hipMemcpyDeviceToHost, s), "get")) return r;
        if (int r = check_hip(hipStreamSynchronize(s), "get")) return r;
        for (size_t i = 0; i < n; ++i) host[20 * i + k] = col[i];
    }
    const int32_t *isrc[] = {f->P.active, f->P.mature};
    for (int k = 0; k < 2; ++k) {
        if (int r = check_hip(hipMemcpyAsync(ic.data(), isrc[k], n * 4, hipMemcpyDeviceToHost, s), "get")) return r;
        if (int r = check_hip(hipStreamSynchronize(s), "get")) return r;
        for (size_t i = 0; i < n; ++i) host[20 * i + 16 + k] = ic[i];
    }
    return AMH_OK;
}

// One FluidModel::step (fluidmodel.cpp:165-580).
extern "C" int amh_fluid_step(void *fluid, int32_t steps_left, double freedom_radius, uint64_t seed, void *stream) {
    amh_fluid_s *f = static_cast<amh_fluid_s *>(fluid);
    if (!f) { set_error("amh_fluid_step: NULL"); return AMH_E_ARG; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t nn = int64_t(f->gsx) * f->gsy;
    const uint32_t gn = uint32_t((nn + kT - 1) / kT), gp = uint32_t((f->count + kT - 1) / kT);
    // the grid's 13 fields are contiguous: clear them in one pass
    hipLaunchKernelGGL(k_fill_f32, dim3(uint32_t((nn * 13 + kT - 1) / kT)), dim3(kT), 0, s, f->G.m, nn * 13);
    if (f->count > 0) {
        {
            amh::ProfScope ps_("k_p2g", s);
            hipLaunchKernelGGL(k_p2g, dim3(gp), dim3(kT), 0, s, f->P, f->count, f->G, f->gsx, f->gsy);
        }
        {
            amh::ProfScope ps_("k_forces", s);
            hipLaunchKernelGGL(k_forces, dim3(gp), dim3(kT), 0, s, f->P, f->count, f->G, f->gsx, f->gsy);
        }
        hipLaunchKernelGGL(k_node_div, dim3(gn), dim3(kT), 0, s, f->G.m, f->G.ax, f->G.ay, nn);
        {
            amh::ProfScope ps_("k_g2p", s);
            hipLaunchKernelGGL(k_g2p, dim3(gp), dim3(kT), 0, s, f->P, f->count, f->G, f->gsx, f->gsy);
        }
        hipLaunchKernelGGL(k_node_div, dim3(gn), dim3(kT), 0, s, f->G.m, f->G.u, f->G.v, nn);
        {
            amh::ProfScope ps_("k_advect", s);
            hipLaunchKernelGGL(k_advect, dim3(gp), dim3(kT), 0, s, f->P, f->count, f->G, f->gsx, f->gsy, int(steps_left),
                               freedom_radius, seed);
        }
    }
    AMH_LAUNCH_CHECK("amh_fluid_step");
    return AMH_OK;
}

// ----------------------------------------------------------- frame path
namespace {

// Particles as morph::synchronize creates them (morph.cpp:243-262).
__global__ void k_fluid_reset(Particles P, int n, double x0, double y0) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    P.x[i] = x0; P.y[i] = y0; P.u[i] = 0.0; P.v[i] = 0.0;
    P.gx[i] = x0; P.gy[i] = y0; P.fr[i] = 1.0;
    P.R[i] = 0.0; P.G[i] = 1.0; P.B[i] = 0.0; P.A[i] = 1.0;
    P.r[i] = 0.0; P.g[i] = 1.0; P.b[i] = 0.0; P.a[i] = 1.0;
    P.strength[i] = 1.0;
    P.active[i] = 0; P.mature[i] = 0; P.owner[i] = 0; P.key[i] = -1;
}

// step_fluid's bookkeeping and update_particle (morph.cpp:680-1050).
__global__ __launch_bounds__(kT) void k_fluid_update(amh_render R, Frame F, Particles P,
                                                    const int32_t *__restrict__ rank,
                                                    const int32_t *__restrict__ surf,
                                                    const double *__restrict__ bcol, double sp) {
    __shared__ int lagp[512], slopep[512];
    if (R.fading == kPerlin) {
        for (int i = threadIdx.x; i < 512; i += kT) {
            lagp[i] = R.lag_perm_dev[i];
            slopep[i] = R.slope_perm_dev[i];
        }
    }
    __syncthreads();
    const int64_t a = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (a >= R.wtot) return;
    const int c = find_chain(R.chain_off_dev, R.nchains, a);
    const int nA = surf[int64_t(F.fk) * R.nchains + c], nB = surf[int64_t(F.fnext) * R.nchains + c];
    const int limit = int(round((1.0 - F.tl) * double(nA) + F.tl * double(nB)));
    if (rank[int64_t(F.fk) * R.wtot + a] >= limit) {
        P.active[a] = 0;
        return;
    }
    if (!(P.active[a] && P.key[a] == F.fk)) {  // a new particle (Particle::clear, fluidmodel.cpp:80-95)
        P.u[a] = 0.0; P.v[a] = 0.0;
        P.strength[a] = 1.0;
        P.mature[a] = 0;
        P.owner[a] = 1;
        P.key[a] = F.fk;
    }
    const uint64_t p1 = R.pts_dev[int64_t(F.fk) * R.wtot + a], p2 = R.pts_dev[int64_t(F.fnext) * R.wtot + a];
    const bool h1 = pt_flags(p1) & HAS_PIXEL, h2 = pt_flags(p2) & HAS_PIXEL;
    const int W = R.width, H = R.height;
    uint32_t c1, c2;
    if (h1 && !h2) {
        c1 = key_px(R.keys_dev, F.fk, W, H, pt_x(p1), pt_y(p1));
        c2 = nB > 0 ? (c1 & 0x00FFFFFFu) : c1;
    } else if (h2 && !h1) {
        c2 = key_px(R.keys_dev, F.fnext, W, H, pt_x(p2), pt_y(p2));
        c1 = nA > 0 ? (c2 & 0x00FFFFFFu) : c2;
    } else {
        c1 = key_px(R.keys_dev, F.fk, W, H, pt_x(p1), pt_y(p1));
        c2 = key_px(R.keys_dev, F.fnext, W, H, pt_x(p2), pt_y(p2));
    }
    const uint32_t col = atom_colour(R, F, p1, c1, c2, lagp, slopep);
    uint32_t qx, qy, qxf, qyf;
    traj_point(R, F, a, p1, p2, qx, qy, qxf, qyf);
    // point2xy works in float (atomorph.h:316-319)
    const float ptx = float(qx) + float(qxf) / float(256), pty = float(qy) + float(qyf) / float(256);
    const double gxv = fmin(fmax((double(ptx) * sp) + 10.0, 1.0), double(W * sp + 10.0));
    const double gyv = fmin(fmax((double(pty) * sp) + 10.0, 1.0), double(H * sp + 10.0));
    P.gx[a] = gxv;
    P.gy[a] = gyv;
    double Rr = ch(col, 0) / 255.0, Gg = ch(col, 1) / 255.0, Bb = ch(col, 2) / 255.0, Aa = ch(col, 3) / 255.0;
    if (R.show_blobs == 2 && bcol) {  // DISTINCT
        const double *q = bcol + (int64_t(F.fk) * R.nchains + c) * 4;
        Rr = q[0]; Gg = q[1]; Bb = q[2]; Aa = 1.0;
    } else if (R.show_blobs == 1 && bcol) {  // AVERAGE: the blob's mean colour
        const double *q = bcol + (int64_t(F.fk) * R.nchains + c) * 4;
        Rr = q[0]; Gg = q[1]; Bb = q[2]; Aa = q[3];
    }
    P.R[a] = Rr; P.G[a] = Gg; P.B[a] = Bb; P.A[a] = Aa;
    const double t = F.tl;
    double r = t * Rr + (1.0 - t) * P.r[a], g = t * Gg + (1.0 - t) * P.g[a];
    double b = t * Bb + (1.0 - t) * P.b[a], al = t * Aa + (1.0 - t) * P.a[a];
    if (!P.mature[a]) {
        if (P.owner[a]) {
            P.x[a] = gxv;
            P.y[a] = gyv;
            r = Rr; g = Gg; b = Bb; al = Aa;
        }
        P.strength[a] = nA == 0 ? 0.1 : 1.0;
        P.fr[a] = 1.0;
    }
    P.r[a] = r; P.g[a] = g; P.b[a] = b; P.a[a] = al;
    P.active[a] = 1;
}

// draw_fluid, splat (morph.cpp:1080-1167)
__global__ void k_fluid_splat(Particles P, int n, double t, int W, int H, double sp, int bx1, int by1, int bx2,
                              int by2, double *acc) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n || !P.active[i]) return;
    const double s = P.strength[i];
    const double w = (1.0 / (100.0 * pow((t - (1.0 * (1.0 - s))), 2.0) + 1.0)) * pow(s, 2.0);
    const double r = (w * P.R[i]) + (1.0 - w) * P.r[i], g = (w * P.G[i]) + (1.0 - w) * P.g[i];
    const double b = (w * P.B[i]) + (1.0 - w) * P.b[i], al = (w * P.A[i]) + (1.0 - w) * P.a[i];
    P.r[i] = r; P.g[i] = g; P.b[i] = b; P.a[i] = al;
    double px, py;
    const double x = P.x[i], y = P.y[i];
    if (x <= 10.0) px = 0.0;
    else if (x >= W * sp + 10.0) px = (W - 1);
    else px = (x - 10.0) / sp;
    if (y <= 10.0) py = 0.0;
    else if (y >= H * sp + 10.0) py = (H - 1);
    else py = (y - 10.0) / sp;
    const int X = min(int(px), W - 1), Y = min(int(py), H - 1);
    P.mature[i] = 1;
    double ix, iy;
    const double xf = modf(px, &ix), yf = modf(py, &iy);
    if (ix >= W || iy >= H) {
        if (ix > bx2 || ix < bx1 || iy > by2 || iy < by1) return;
    }
    const double w11 = ((1.0 - xf) * (1.0 - yf)), w21 = (xf * (1.0 - yf));
    const double w12 = ((1.0 - xf) * yf), w22 = (xf * yf);
    const uint32_t col = rgba_pack(narrow8(round(r * 255.0)), narrow8(round(g * 255.0)), narrow8(round(b * 255.0)),
                                   narrow8(round(al * 255.0)));
    auto add = [&](int qx, int qy, double wt) {
        if (qx < 0 || qy < 0 || qx >= W || qy >= H) return;
        double *q = acc + (int64_t(qy) * W + qx) * 5;
        atomicAdd(&q[0], wt);
        for (int k = 0; k < 4; ++k) atomicAdd(&q[1 + k], double(ch(col, k)) * wt);
    };
    if (w11 > 0.0) add(X, Y, w11);
    if ((X < bx2 || X + 1 < W) && w21 > 0.0) add(X + 1, Y, w21);
    if ((Y < by2 || Y + 1 < H) && w12 > 0.0) add(X, Y + 1, w12);
    if (w22 > 0.0 && ((Y < by2 && X < bx2) || (Y + 1 < H && X + 1 < W))) add(X + 1, Y + 1, w22);
}

// resolve (morph.cpp:1169-1209); clears the accumulators
__global__ void k_fluid_resolve(double *acc, int64_t npx, double t, int keep_bg, uint32_t *img, int32_t *mask) {
    const int64_t p = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (p >= npx) return;
    double *q = acc + p * 5;
    const double ws = q[0];
    if (ws > 0.0) {
        const double r = round(q[1] / ws), g = round(q[2] / ws), b = round(q[3] / ws);
        double a = round(q[4] / ws);
        if (keep_bg) a *= 1.0 - pow(t, 24.0);
        img[p] = rgba_pack(narrow8(r), narrow8(g), narrow8(b), narrow8(a));
        mask[p] = 1;
        for (int k = 0; k < 5; ++k) q[k] = 0.0;
    } else {
        mask[p] = 0;
    }
}

// One outward feather layer (morph.cpp:1211-1262): empty pixels next to the
// blob take the mean colour of their filled 4-neighbours, alpha scaled.
__global__ void k_fluid_feather(int W, int H, double alpha, uint32_t *img, int32_t *mask) {
    const int64_t p = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (p >= int64_t(W) * H || mask[p] == 1) return;
    const int x = int(p % W), y = int(p / W);
    // neighbours in increasing position order (the std::set order)
    int64_t nb[4];
    int n = 0;
    if (y > 0 && mask[p - W] == 1) nb[n++] = p - W;
    if (x > 0 && mask[p - 1] == 1) nb[n++] = p - 1;
    if (x + 1 < W && mask[p + 1] == 1) nb[n++] = p + 1;
    if (y + 1 < H && mask[p + W] == 1) nb[n++] = p + W;
    if (n == 0) return;
    double r = 0, g = 0, b = 0, a = 0;
    const double w = n;
    for (int k = 0; k < n; ++k) {
        const uint32_t c = img[nb[k]];
        r += (ch(c, 0) / 255.0) / w;
        g += (ch(c, 1) / 255.0) / w;
        b += (ch(c, 2) / 255.0) / w;
        a += (ch(c, 3) / 255.0) / w;
    }
    a *= alpha;
    img[p] = rgba_pack(narrow8(round(r * 255.0)), narrow8(round(g * 255.0)), narrow8(round(b * 255.0)),
                       narrow8(round(a * 255.0)));
    mask[p] = 2;
}
__global__ void k_fluid_commit(int32_t *mask, int64_t npx) {
    const int64_t p = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (p < npx && mask[p] == 2) mask[p] = 1;
}

// background and the final blend (morph.cpp:1264-1298, 1405-1417)
__global__ __launch_bounds__(kT) void k_fluid_compose(amh_render R, Frame F, const uint32_t *img, const int32_t *mask,
                                                      uint32_t *out) {
    __shared__ int lagp[512], slopep[512];
    if (R.keep_background && R.fading == kPerlin) {
        for (int i = threadIdx.x; i < 512; i += kT) {
            lagp[i] = R.lag_perm_dev[i];
            slopep[i] = R.slope_perm_dev[i];
        }
    }
    __syncthreads();
    const int64_t p = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (p >= int64_t(R.width) * R.height) return;
    const int x = int(p % R.width), y = int(p / R.width);
    const uint32_t bg = R.keep_background ? background(R, F, x, y, lagp, slopep) : 0u;
    if (mask[p] != 1) {
        out[p] = bg;
        return;
    }
    uint32_t c = img[p];
    if (R.keep_background) {
        const double bgr = double(ch(bg, 0)) / 255.0, bgg = double(ch(bg, 1)) / 255.0;
        const double bgb = double(ch(bg, 2)) / 255.0, bga = double(ch(bg, 3)) / 255.0;
        double r = ch(c, 0) / 255.0, g = ch(c, 1) / 255.0, b = ch(c, 2) / 255.0, a = ch(c, 3) / 255.0;
        r = a * r + (1.0 - a) * bgr;
        g = a * g + (1.0 - a) * bgg;
        b = a * b + (1.0 - a) * bgb;
        a = bga + (1.0 - bga) * a;
        c = rgba_pack(narrow8(round(r * 255.0)), narrow8(round(g * 255.0)), narrow8(round(b * 255.0)),
                      narrow8(round(a * 255.0)));
    }
    out[p] = c;
}

}  // namespace

extern "C" int amh_fluid_reset(void *fluid, double sparseness, void *stream) {
    amh_fluid_s *f = static_cast<amh_fluid_s *>(fluid);
    if (!f || !(sparseness > 0.0)) { set_error("amh_fluid_reset: invalid argument"); return AMH_E_ARG; }
    if (f->count > 0)
        hipLaunchKernelGGL(k_fluid_reset, dim3((f->count + kT - 1) / kT), dim3(kT), 0,
                           static_cast<hipStream_t>(stream), f->P, f->count, f->W * sparseness / 2.0 - 10,
                           f->H * sparseness / 2.0 - 10);
    f->frame = 0;
    AMH_LAUNCH_CHECK("amh_fluid_reset");
    return AMH_OK;
}

extern "C" int amh_fluid_render(void *fluid, const amh_fluid_frame *fa, double t, uint32_t *out, void *stream) {
    amh_fluid_s *f = static_cast<amh_fluid_s *>(fluid);
    if (!f || !fa || !fa->r || !out || !fa->rank_dev || !fa->surface_dev || !(fa->sparseness > 0.0) ||
        fa->fluidsteps < 0) {
        set_error("amh_fluid_render: invalid argument");
        return AMH_E_ARG;
    }
    const amh_render &R = *fa->r;
    if (R.width != f->W || R.height != f->H || R.wtot != f->count || !R.keys_dev || !R.pts_dev ||
        (R.fading == kPerlin && (!R.lag_perm_dev || !R.slope_perm_dev))) {
        set_error("amh_fluid_render: render arguments do not match the model (%dx%d, %lld atoms)", f->W, f->H,
                  (long long)f->count);
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Frame F = frame_of(t, R.nkeys);
    const int64_t npx = int64_t(f->W) * f->H;
    const uint32_t gpx = uint32_t((npx + kT - 1) / kT);
    if (F.fk < R.nkeys && f->count > 0) {
        const uint32_t gp = uint32_t((f->count + kT - 1) / kT);
        {
            amh::ProfScope ps_("k_fluid_update", s);
            hipLaunchKernelGGL(k_fluid_update, dim3(gp), dim3(kT), 0, s, R, F, f->P, fa->rank_dev, fa->surface_dev,
                               fa->blob_col_dev, fa->sparseness);
        }
        // morph.cpp:1050-1054
        const double freedom = sqrt(double(f->W * f->W + f->H * f->H)) * fa->sparseness / 4.0;
        const double radius = freedom * (1.0 - F.tl);
        for (int k = 0; k < fa->fluidsteps; ++k) {
            const int r = amh_fluid_step(f, fa->fluidsteps - (k + 1), radius, mix64(f->frame * 131 + k), stream);
            if (r) return r;
        }
        {
            amh::ProfScope ps_("k_fluid_splat", s);
            hipLaunchKernelGGL(k_fluid_splat, dim3(gp), dim3(kT), 0, s, f->P, f->count, F.tl, f->W, f->H, fa->sparseness,
                               R.bbox_x1, R.bbox_y1, R.bbox_x2, R.bbox_y2, f->acc);
        }
    }
    hipLaunchKernelGGL(k_fluid_resolve, dim3(gpx), dim3(kT), 0, s, f->acc, npx, F.tl, R.keep_background, f->img,
                       f->mask);
    for (int layer = 0; layer < R.feather; ++layer) {
        const double alpha = 1.0 - (double(layer + 1) / double(R.feather + 1));
        hipLaunchKernelGGL(k_fluid_feather, dim3(gpx), dim3(kT), 0, s, f->W, f->H, alpha, f->img, f->mask);
        hipLaunchKernelGGL(k_fluid_commit, dim3(gpx), dim3(kT), 0, s, f->mask, npx);
    }
    hipLaunchKernelGGL(k_fluid_compose, dim3(gpx), dim3(kT), 0, s, R, F, f->img, f->mask, out);
    ++f->frame;
    AMH_LAUNCH_CHECK("amh_fluid_render");
    return AMH_OK;
}
