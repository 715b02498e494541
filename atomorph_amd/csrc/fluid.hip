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
// walls.  Particle state is FP64; the grid is FP32, summed per tile in LDS.
#include "colour.h"
#include "common.h"
#include "../../include/atomorph_hip.h"
#include "scan.h"
#include "shade.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

using namespace amh;
using namespace amh::shade;

namespace {

constexpr int kT = 256;

struct Particles {  // structure of arrays, device
    double *x, *y, *u, *v, *gx, *gy, *fr;
    double *R, *G, *B, *A, *r, *g, *b, *a;
    double *strength;
    double *u2, *v2;  // the velocity between G2P and advection
    int32_t *active, *mature, *owner, *key;  // owner: first at its source
    int32_t *atom;  // in the cell-ordered copy: the particle (atom) in each slot
};
constexpr int kParticleDoubles = 18;
constexpr int kParticleInts = 5;

// FP32 node fields, [gsx * gsy], node (i, j) at i * gsy + j.  FluidModel's
// node density accumulates the same weights as its mass, so one field (m)
// serves both.
// Node fields, FP32.  Mass, momentum and colour weight are one 16-byte
// record per node and the colour sums another, so advection reads a node's
// values with two loads; the density gradient and accelerations are read
// in tiles through LDS and stay one array per field.
struct Grid {
    float *muvw;  // [nodes][4]: m, u, v, w
    float *col;   // [nodes][4]: r, g, b, a
    float *gx, *gy, *ax, *ay;
};
constexpr int kGridFields = 12;  // floats per node

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

// ---------------------------------------------------- stencil scatter
// Particles scatter into the 3x3 nodes around them.  Neighbouring atoms of a
// chain stay close together, so the nodes one block's particles touch form a
// small box: the block sums its contributions into an LDS copy of that box
// (LDS atomics) and then adds each touched node to the grid once per field.
// That turns ~9 contended global atomics per particle and field into one per
// node.  A block whose box exceeds the tile scatters straight to the grid.
struct Box {
    int x0, y0, w, h;
    bool fits;
};

// Bounding box of the E x E stencils at (cx, cy) of the block's valid
// particles; fits when it has at most NODES entries.
template <int NODES, int E = 3>
__device__ Box block_box(bool valid, int cx, int cy, int *red) {
    int mnx = valid ? cx : 0x7FFFFFFF, mny = valid ? cy : 0x7FFFFFFF;
    int mxx = valid ? cx : -0x7FFFFFFF, mxy = valid ? cy : -0x7FFFFFFF;
    for (int off = 32; off > 0; off >>= 1) {
        mnx = min(mnx, __shfl_xor(mnx, off, 64));
        mny = min(mny, __shfl_xor(mny, off, 64));
        mxx = max(mxx, __shfl_xor(mxx, off, 64));
        mxy = max(mxy, __shfl_xor(mxy, off, 64));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave * 4 + 0] = mnx;
        red[wave * 4 + 1] = mny;
        red[wave * 4 + 2] = mxx;
        red[wave * 4 + 3] = mxy;
    }
    __syncthreads();
    for (int w = 0; w < kT / 64; ++w) {
        mnx = min(mnx, red[w * 4 + 0]);
        mny = min(mny, red[w * 4 + 1]);
        mxx = max(mxx, red[w * 4 + 2]);
        mxy = max(mxy, red[w * 4 + 3]);
    }
    Box b;
    b.x0 = mnx;
    b.y0 = mny;
    const int64_t w = int64_t(mxx) - mnx + E, h = int64_t(mxy) - mny + E;
    b.fits = mnx <= mxx && w * h <= NODES;
    b.w = b.fits ? int(w) : 0;
    b.h = b.fits ? int(h) : 0;
    return b;
}

// Runs of consecutive lanes whose particles share a stencil origin form
// segments (the cell order makes such runs common).  seg_sum leaves each
// segment's total in its last lane, so one lane per segment scatters.
struct Seg {
    int lane, start;  // start: first lane of this lane's segment
    bool tail;        // last lane of its segment
};

// The neighbours' origins come by DPP wave shifts and the segment starts
// from a ballot of the heads, so no LDS round trip is on this path.
__device__ __forceinline__ Seg segments(bool valid, int cx, int cy) {
    Seg s;
    s.lane = threadIdx.x & 63;
    // origin key; an invalid lane matches no origin (nodes are < 65535)
    const int key = valid ? int((uint32_t(cx) << 16) | (uint32_t(cy) & 0xFFFFu)) : -1;
    const int prev = __builtin_amdgcn_update_dpp(-1, key, 0x138, 0xF, 0xF, false);  // wave_shr:1, lane - 1
    const int next = __builtin_amdgcn_update_dpp(-1, key, 0x130, 0xF, 0xF, false);  // wave_shl:1, lane + 1
    const bool head = !valid || s.lane == 0 || prev != key;
    s.tail = !valid || s.lane == 63 || next != key;
    const uint64_t heads = __ballot(head) & ((2ull << s.lane) - 1ull);  // heads at or below this lane
    s.start = 63 - __clzll(static_cast<long long>(heads));               // lane 0 is always a head
    return s;
}

// DPP move of v: row_shr:k (0x110 + k), row_bcast:15 (0x142) or
// row_bcast:31 (0x143); lanes without a source (or outside row_mask) get 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}

// Inclusive segmented sum over lanes [start, lane]; every lane of the wave
// must take part.
__device__ __forceinline__ float seg_sum(float v, const Seg &s) {
    float t;
    t = dpp<0x111, 0xF>(v);
    if (s.lane - 1 >= s.start) v += t;
    t = dpp<0x112, 0xF>(v);
    if (s.lane - 2 >= s.start) v += t;
    t = dpp<0x114, 0xF>(v);
    if (s.lane - 4 >= s.start) v += t;
    t = dpp<0x118, 0xF>(v);
    if (s.lane - 8 >= s.start) v += t;
    t = dpp<0x142, 0xA>(v);  // lane 15 -> row 1, lane 47 -> row 3
    if ((s.lane & 16) && s.start <= (s.lane & ~15) - 1) v += t;
    t = dpp<0x143, 0xC>(v);  // lane 31 -> rows 2 and 3
    if (s.lane >= 32 && s.start <= 31) v += t;
    return v;
}

// ------------------------------------------------------------ P2G
// Mass (= density), density gradient and strength-weighted colour
// (fluidmodel.cpp:191-249).  With e = x - cx - 1 in [-0.5, 0.5) the
// quadratic B-spline weights and their gradients are polynomials in e:
//   px[a] = sum_p kW[a][p] e^p,   gx[a] = sum_p kD[a][p] e^p,
// so the node sums of px[a] py[b], gx[a] py[b] and px[a] gy[b] over the
// particles of one cell follow from the cell's nine moments
// M[p][q] = sum e^p h^q (h likewise in y).  A particle therefore adds 9
// moments and 5 colour sums to its one cell instead of 8 values to each of 9
// nodes; the nodes are assembled from the 3x3 cells around them.
__constant__ float kW[3][3] = {{0.125f, -0.5f, 0.5f}, {0.75f, 0.0f, -1.0f}, {0.125f, 0.5f, 0.5f}};
__constant__ float kD[3][3] = {{0.5f, -1.0f, 0.0f}, {0.0f, 2.0f, 0.0f}, {-0.5f, -1.0f, 0.0f}};
constexpr int kCellSums = 14;  // 9 moments (p-major), then s*r, s*g, s*b, s*a, s

// The 8 node values (m, gx, gy, r, g, b, a, w) that the sums c[] of cell
// (node - (a, b)) contribute to the node.
__device__ __forceinline__ void cell_to_node(const float *c, int a, int b, float *v) {
    float xw[3], xd[3];  // sum_p kW[a][p] M[p][q], sum_p kD[a][p] M[p][q]
    for (int q = 0; q < 3; ++q) {
        xw[q] = kW[a][0] * c[q] + kW[a][1] * c[3 + q] + kW[a][2] * c[6 + q];
        xd[q] = kD[a][0] * c[q] + kD[a][1] * c[3 + q] + kD[a][2] * c[6 + q];
    }
    v[0] = kW[b][0] * xw[0] + kW[b][1] * xw[1] + kW[b][2] * xw[2];
    v[1] = kW[b][0] * xd[0] + kW[b][1] * xd[1] + kW[b][2] * xd[2];
    v[2] = kD[b][0] * xw[0] + kD[b][1] * xw[1] + kD[b][2] * xw[2];
    for (int k = 0; k < 5; ++k) v[3 + k] = c[9 + k];
}

__device__ __forceinline__ float nd(const float *f, int i, int j, int gsx, int gsy) {
    return node_ok(i, j, gsx, gsy) ? f[int64_t(i) * gsy + j] : 0.0f;
}

// A node's acceleration and velocity as FluidModel reads them after dividing
// them by the node mass where it is positive (fluidmodel.cpp:359-364,
// 437-443); the readers divide, so no pass over the grid is needed.
__device__ __forceinline__ float per_mass(float v, float m) { return m > 0.0f ? v / m : v; }
// f[k] / mass of node (i, j); muvw holds the mass of node k at 4k.
__device__ __forceinline__ float nd_m(const float *f, const float *muvw, int i, int j, int gsx, int gsy) {
    if (!node_ok(i, j, gsx, gsy)) return 0.0f;
    const int64_t k = int64_t(i) * gsy + j;
    return per_mass(f[k], muvw[4 * k]);
}

// pressure from the bicubic density interpolation, wall forces
// (fluidmodel.cpp:251-357)
// node(f, i, j): field f (0 mass, 1 density gradient x, 2 gradient y) of
// node (i, j), 0 outside the grid.
template <class Node>
__device__ __forceinline__ void particle_forces(double x, double y, const Node &node, int gsx, int gsy,
                                                double &pressure_out, double &fx_out, double &fy_out) {
    const int cx = static_cast<int>(x), cy = static_cast<int>(y);
    const double d01 = node(0, cx, cy), d02 = node(0, cx, cy + 1);
    const double d11 = node(0, cx + 1, cy), d12 = node(0, cx + 1, cy + 1);
    const double gx01 = node(1, cx, cy), gx02 = node(1, cx, cy + 1);
    const double gx11 = node(1, cx + 1, cy), gx12 = node(1, cx + 1, cy + 1);
    const double gy01 = node(2, cx, cy), gy02 = node(2, cx, cy + 1);
    const double gy11 = node(2, cx + 1, cy), gy12 = node(2, cx + 1, cy + 1);
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
    pressure_out = pressure;
    fx_out = fx;
    fy_out = fy;
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

// ------------------------------------------------------------------ tiles
// The grid's nodes are owned by kTile x kTile tiles.  Each step every active
// particle is binned into each tile that holds a node of its 3x3 stencil (1,
// 2 or 4 tiles; a counting sort of slot numbers); the entry for the tile of
// its own cell is marked as its home.  The block of a tile reads its bin,
// sums what those particles add to its nodes in LDS and stores every node
// once.  No phase of the step uses a global atomic (device-wide atomics are
// performed past the per-XCD L2) and no phase needs the grid cleared first.
constexpr int kTile = 8;
constexpr int kReach = kTile + 2;       // cells per axis that reach a tile's nodes
constexpr int kHome = int(0x80000000);  // bin entry flag: the particle's home tile

struct Bins {
    int32_t ntx, nty;   // tiles per axis
    int32_t *count[2];  // [tiles] per step parity; the other parity is zero
    int32_t *start;     // [tiles + 1]
    int32_t *rank;      // [4][n]: place of particle i's j-th entry in its bin
    int32_t *list;      // [<= 4n] slot | kHome, grouped by bin
    // Work items of kT entries at most, each one workgroup (chunk NULL:
    // whole bins, one workgroup per tile, from start[]).  A dense bin is
    // split over several items, which add their node sums with atomics to
    // nodes cleared when the bins were placed; a tile with one item stores
    // its nodes.  Only tiles with entries get items, plus one (empty) item
    // for each tile that had entries in the step before and has none now,
    // which stores zeros: a tile empty in both steps keeps the zeros stored
    // then.  chunk[b] = (tile, entries, only item of its tile, 0), b <
    // *nchunks, in tile order; item b's entries are list[kT * b ...], so a
    // tile kernel loads its entry together with the descriptor.
    int4 *chunk;
    int32_t *nchunks;     // this step's item count
    int32_t max_chunks;   // capacity of chunk (an upper bound of *nchunks)
    int32_t *items;       // [tiles] items per tile (scratch of grids beyond kLdsTiles)
    int32_t par;          // this step's count parity (the other is cleared by k_p2g)
};
constexpr int kLdsTiles = 4096;  // up to this many tiles the placing blocks scan the counts in LDS

// The tiles holding the stencil nodes of a particle at (x, y): tx0..tx1 x
// ty0..ty1 (nodes clamped to the grid); (tx0, ty0) is its home.
__device__ __forceinline__ void stencil_tiles(double x, double y, int gsx, int gsy, int &tx0, int &tx1, int &ty0,
                                              int &ty1) {
    const int cx = static_cast<int>(x - 0.5), cy = static_cast<int>(y - 0.5);
    tx0 = min(max(cx, 0), gsx - 1) / kTile;
    tx1 = min(max(cx + 2, 0), gsx - 1) / kTile;
    ty0 = min(max(cy, 0), gsy - 1) / kTile;
    ty1 = min(max(cy + 2, 0), gsy - 1) / kTile;
}

// Entry j (0..3) of a particle: tile (tx0 + (j >> 1), ty0 + (j & 1)) when it
// exists, else -1.
__device__ __forceinline__ int entry_tile(int j, int tx0, int tx1, int ty0, int ty1, int nty) {
    const int bx = tx0 + (j >> 1), by = ty0 + (j & 1);
    return (bx <= tx1 && by <= ty1) ? bx * nty + by : -1;
}

// What a particle is binned by: tiles(i, tx0, tx1, ty0, ty1) gives the
// block tx0..tx1 x ty0..ty1 of tiles it enters ((tx0, ty0) its home) and
// returns false for a particle that is not binned.
struct StencilTiles {  // the step: tiles holding the 3x3 stencil nodes
    Particles P;
    int gsx, gsy;
    Grid G;
    __device__ __forceinline__ bool operator()(int i, int &tx0, int &tx1, int &ty0, int &ty1) const {
        // the three loads together: the position is read whatever the flag
        const int act = P.active[i];
        stencil_tiles(P.x[i], P.y[i], gsx, gsy, tx0, tx1, ty0, ty1);
        return act != 0;
    }
    // Zeroes every field of tile t's nodes (the workgroup together).
    __device__ __forceinline__ void clear_tile(int t, int nty) const {
        const int x0 = (t / nty) * kTile, y0 = (t % nty) * kTile;
        for (int e = threadIdx.x; e < kTile * kTile * kGridFields; e += kT) {
            const int node = e % (kTile * kTile), field = e / (kTile * kTile);
            const int ni = x0 + node / kTile, nj = y0 + node % kTile;
            if (ni >= gsx || nj >= gsy) continue;
            const int64_t k = int64_t(ni) * gsy + nj;
            if (field < 4) G.muvw[4 * k + field] = 0.0f;
            else if (field < 8) G.col[4 * k + field - 4] = 0.0f;
            else {
                float *f[4] = {G.gx, G.gy, G.ax, G.ay};
                f[field - 8][k] = 0.0f;
            }
        }
    }
};

constexpr int kPxTile = 32;  // pixels per side of a splat tile

constexpr uint32_t kNoTap = 0xFFFFFFFFu;  // splat: particle draws nothing (X, Y <= 65534)

// The pixel tiles of a splat's taps at v = X | Y << 16 (v != kNoTap).
__device__ __forceinline__ void splat_tiles(uint32_t v, int W, int H, int &tx0, int &tx1, int &ty0, int &ty1) {
    const int X = int(v & 0xFFFFu), Y = int(v >> 16);
    tx0 = X / kPxTile;
    tx1 = min(X + 1, W - 1) / kPxTile;
    ty0 = Y / kPxTile;
    ty1 = min(Y + 1, H - 1) / kPxTile;
}

struct SplatTiles {  // the splat: pixel tiles holding the 2x2 taps at (X, Y)
    const uint32_t *xy;  // X | Y << 16, kNoTap = nothing to draw
    int W, H;
    __device__ __forceinline__ void clear_tile(int, int) const {}  // not chunked
    __device__ __forceinline__ bool operator()(int i, int &tx0, int &tx1, int &ty0, int &ty1) const {
        const uint32_t v = xy[i];
        if (v == kNoTap) return false;
        splat_tiles(v, W, H, tx0, tx1, ty0, ty1);
        return true;
    }
};

// Counts slot i's entries (tiles tx0..tx1 x ty0..ty1 when act) into
// B.count[par] and stores each entry's place in its bin in B.rank.  Every
// lane of the wave calls it (it groups the wave's lanes with ballots).
__device__ __forceinline__ void count_entries(bool act, int tx0, int tx1, int ty0, int ty1, int i, int n,
                                              const Bins &B, int par) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    // Group the wave's lanes by bin, per entry: lane lead[j] leads the lanes
    // holding bin t[j] and adds their number to the bin; all atomics of the
    // wave are then in flight together.
    int t[4], lead[4], pre[4], total[4];
    for (int j = 0; j < 4; ++j) {
        t[j] = act ? entry_tile(j, tx0, tx1, ty0, ty1, B.nty) : -1;
        lead[j] = -1;
        pre[j] = 0;
        total[j] = 0;
        uint64_t todo = __ballot(t[j] >= 0);
        while (todo) {
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const int tl = __shfl(t[j], leader, 64);
            const uint64_t same = __ballot(t[j] == tl);
            if (t[j] == tl) {
                lead[j] = leader;
                pre[j] = int(__popcll(same & below));
            }
            if (lane == leader) total[j] = int(__popcll(same));
            todo &= ~same;
        }
    }
    int base[4];
    for (int j = 0; j < 4; ++j) base[j] = total[j] > 0 ? atomicAdd(&B.count[par][t[j]], total[j]) : 0;
    for (int j = 0; j < 4; ++j) {
        const int b = __shfl(base[j], lead[j] < 0 ? lane : lead[j], 64);
        if (i < n) B.rank[int64_t(j) * n + i] = lead[j] < 0 ? 0 : b + pre[j];
    }
}

template <class Tiles>
__global__ __launch_bounds__(kT) void k_bin_count(Tiles tiles_of, int n, Bins B, int par) {
    const int i = blockIdx.x * kT + threadIdx.x;
    int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
    const bool act = i < n && tiles_of(i, tx0, tx1, ty0, ty1);
    count_entries(act, tx0, tx1, ty0, ty1, i, n, B, par);
}

// Exclusive scan of value(k) over tiles k < tiles into out[0..tiles] (out[tiles]
// = total) by the whole block in LDS: thread t takes a run of tiles.  value
// is called once per tile, by the thread of its run, in tile order.
template <class Value>
__device__ __forceinline__ void block_scan_tiles(int tiles, const Value &value, int *out, int *wsum) {
    const int per = (tiles + kT - 1) / kT;
    const int lo = min(int(threadIdx.x) * per, tiles), hi = min(lo + per, tiles);
    int sum = 0;
    for (int k = lo; k < hi; ++k) {
        const int v = value(k);
        out[k] = sum;
        sum += v;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int off = incl - sum;
    for (int w = 0; w < wave; ++w) off += wsum[w];
    for (int k = lo; k < hi; ++k) out[k] += off;
    if (threadIdx.x == kT - 1) out[tiles] = off + sum;
    __syncthreads();
}

// Work items of a tile with c entries that had `before` entries in the step
// before: ceil(c / kT) chunks, or one empty item (which zeroes the tile's
// nodes) when the tile has just emptied, or none.
__device__ __forceinline__ int work_items(int c, int before) {
    return c > 0 ? (c + kT - 1) / kT : (before > 0 ? 1 : 0);
}

// The nc work items of tile t (c entries) at B.chunk[c0 ...].
__device__ __forceinline__ void put_items(const Bins &B, int c0, int nc, int t, int c) {
    for (int j = 0; j < nc; ++j) B.chunk[c0 + j] = make_int4(t, max(0, min(kT, c - j * kT)), nc == 1, 0);
}

// Nothing more for particle i (k_bin_place's default tail).
struct NoTail {
    __device__ void operator()(int) const {}
};

// Particles into whole bins (the splat's pixel tiles; the tiles recomputed
// from the unchanged positions) and the other parity's counts cleared for
// the next use.  With kLds every block scans the counts into LDS (block 0
// also stores them in B.start); otherwise B.start holds the scan already.
// Then tail(i) for every particle i < n.  The grid covers max(n, tiles)
// threads.
template <bool kLds, class Tiles, class Tail>
__global__ __launch_bounds__(kT) void k_bin_place(Tiles tiles_of, int n, Bins B, int tiles, int par, Tail tail) {
    __shared__ int sstart[kLds ? kLdsTiles + 1 : 1];
    __shared__ int wsum[kT / 64];
    const int32_t *start = B.start;
    // the particle's tiles and its entries' ranks, loaded before the counts
    // are scanned so that the two waits overlap (past n: particle 0, unused)
    const int i = blockIdx.x * kT + threadIdx.x;
    const int ic = i < n ? i : 0;
    int tx0, tx1, ty0, ty1;
    const bool binned = tiles_of(ic, tx0, tx1, ty0, ty1) && i < n;
    int rank[4];
    for (int j = 0; j < 4; ++j) rank[j] = B.rank[int64_t(j) * n + ic];
    if (kLds) {
        block_scan_tiles(tiles, [&](int k) { return B.count[par][k]; }, sstart, wsum);
        if (blockIdx.x == 0)
            for (int k = threadIdx.x; k <= tiles; k += kT) B.start[k] = sstart[k];
        start = sstart;
    }
    for (int k = blockIdx.x * kT + threadIdx.x; k < tiles; k += gridDim.x * kT) B.count[par ^ 1][k] = 0;
    if (binned)
        for (int j = 0; j < 4; ++j) {
            const int t = entry_tile(j, tx0, tx1, ty0, ty1, B.nty);
            if (t >= 0) B.list[start[t] + rank[j]] = i | (j == 0 ? kHome : 0);
        }
    if (i < n) tail(i);
}

// Work items per tile from its count and the step before's (grids beyond
// kLdsTiles tiles, before the device scan of B.items into B.start).
__global__ __launch_bounds__(kT) void k_tile_items(Bins B, int tiles, int par) {
    const int k = blockIdx.x * kT + threadIdx.x;
    if (k < tiles) B.items[k] = work_items(B.count[par][k], B.count[par ^ 1][k]);
}

// The fluid steps' bins: tile t's entries are laid out from list[kT *
// first[t]], kT per work item (the last item of a tile partly filled), so
// that item b's entries sit at list[kT * b ...].  first[] is the scan of
// the tiles' item counts: with kLds every block scans them in LDS (block 0
// also writes the item descriptors and their number); otherwise B.start
// holds it already (k_tile_items and the device scan).  The step before's
// counts, which every placing block reads here, are cleared by k_p2g.  The
// grid covers max(n, tiles) threads.
template <bool kLds>
__global__ __launch_bounds__(kT) void k_item_place(StencilTiles tiles_of, int n, Bins B, int tiles, int par) {
    __shared__ int sfirst[kLds ? kLdsTiles + 1 : 1];
    __shared__ int scount[kLds ? kLdsTiles : 1];
    __shared__ int wsum[kT / 64];
    const int32_t *first = B.start;
    // the particle and its entries' ranks first, as in k_bin_place
    const int i = blockIdx.x * kT + threadIdx.x;
    const int ic = i < n ? i : 0;
    int tx0, tx1, ty0, ty1;
    const bool binned = tiles_of(ic, tx0, tx1, ty0, ty1) && i < n;
    int rank[4];
    for (int j = 0; j < 4; ++j) rank[j] = B.rank[int64_t(j) * n + ic];
    if (kLds) {
        block_scan_tiles(tiles, [&](int k) {
            const int c = B.count[par][k], before = B.count[par ^ 1][k];
            scount[k] = c;
            return work_items(c, before);
        }, sfirst, wsum);
        first = sfirst;
        if (blockIdx.x == 0) {
            for (int k = threadIdx.x; k < tiles; k += kT) put_items(B, sfirst[k], sfirst[k + 1] - sfirst[k], k, scount[k]);
            if (threadIdx.x == 0) *B.nchunks = sfirst[tiles];
        }
        // nodes of tiles split over several items start from zero
        for (int t = blockIdx.x; t < tiles; t += gridDim.x)
            if (scount[t] > kT) tiles_of.clear_tile(t, B.nty);
    } else {
        for (int k = blockIdx.x * kT + threadIdx.x; k < tiles; k += gridDim.x * kT)
            put_items(B, first[k], first[k + 1] - first[k], k, B.count[par][k]);
        if (blockIdx.x == 0 && threadIdx.x == 0) *B.nchunks = first[tiles];
        for (int t = blockIdx.x; t < tiles; t += gridDim.x)
            if (B.count[par][t] > kT) tiles_of.clear_tile(t, B.nty);
    }
    if (!binned) return;
    for (int j = 0; j < 4; ++j) {
        const int t = entry_tile(j, tx0, tx1, ty0, ty1, B.nty);
        if (t >= 0) B.list[int64_t(first[t]) * kT + rank[j]] = i | (j == 0 ? kHome : 0);
    }
}

// Calls body(slot, home, item) for the entries [s0, s1) of a whole bin (the
// splat's pixel tiles), kT at a time (slot -1 past the end), item =
// fetch(slot) holding the particle fields the body reads.  The next kT
// entries and their fields are loaded before the current ones are
// processed, so a dense bin overlaps its memory latency.  Trip counts are
// uniform over the block.
template <class Fetch, class Body>
__device__ __forceinline__ void for_bin(const Bins &B, int s0, int s1, const Fetch &fetch, const Body &body) {
    auto entry = [&](int base, int &slot, bool &home) {
        const int k = base + int(threadIdx.x);
        slot = -1;
        home = false;
        if (k < s1) {
            const int e = B.list[k];
            slot = e & 0x7FFFFFFF;
            home = e < 0;
        }
    };
    int slot, nslot = -1;
    bool home, nhome = false;
    entry(s0, slot, home);
    auto item = fetch(slot);
    for (int base = s0; base < s1; base += kT) {
        auto next = item;
        if (base + kT < s1) {
            entry(base + kT, nslot, nhome);
            next = fetch(nslot);
        }
        body(slot, home, item);
        slot = nslot;
        home = nhome;
        item = next;
    }
}

// Work item b of a tile kernel: tile t, whether it is the tile's only item
// (then it stores its nodes; otherwise it adds them atomically), and this
// thread's entry of it (slot -1 past the item's entries, home: the
// particle's home tile).  The item's entries are list[kT * b ...], so the
// entry is loaded together with the descriptor, not after it (both reads
// are in bounds for every b: b is clamped to the descriptors' capacity and
// the list holds kT entries per descriptor).  false: past the last item.
// Workgroups take items b = blockIdx.x, blockIdx.x + gridDim.x, ...;
// nchunks = *B.nchunks.
__device__ __forceinline__ bool tile_item(const Bins &B, int nchunks, int b, int &t, bool &single, int &slot,
                                          bool &home) {
    const int bc = min(b, B.max_chunks - 1);
    const int4 c = B.chunk[bc];
    const int e = B.list[int64_t(bc) * kT + threadIdx.x];
    if (b >= nchunks) return false;
    t = c.x;
    single = c.z != 0;
    const bool has = int(threadIdx.x) < c.y;
    slot = has ? (e & 0x7FFFFFFF) : -1;
    home = has && e < 0;
    return true;
}

// body(slot, home, item) on this thread's entry, item = fetch(slot) holding
// the particle fields the body reads; prologue() runs once those loads are
// issued and before they are used, so its own memory traffic (a node halo
// staged in LDS, with the barrier that publishes it) overlaps them.
template <class Fetch, class Body, class Prologue>
__device__ __forceinline__ void run_item(int slot, bool home, const Fetch &fetch, const Body &body,
                                         const Prologue &prologue) {
    const auto item = fetch(slot);
    prologue();
    body(slot, home, item);
}

__device__ __forceinline__ void put_node(float *f, int64_t k, float v, bool single) {
    if (single) f[k] = v;
    else if (v != 0.0f) atomicAdd(&f[k], v);
}

__device__ __forceinline__ bool in_reach(int cx, int cy, int x0, int y0) {
    return cx >= x0 - 2 && cx < x0 + kTile && cy >= y0 - 2 && cy < y0 + kTile;
}

// Node values a tile block reads, staged in LDS once per block: nodes
// x0-2 .. x0+kTile+1 by y0-2 .. y0+kTile+1 cover the stencil of every
// particle in reach of the tile (origins x0-2 .. x0+kTile-1, 3 nodes) and
// the bicubic density corners of particle_forces (int(x) <= origin + 1).
// Nodes outside the grid hold 0, as nd() reads them.
constexpr int kHalo = kTile + 4;

__device__ __forceinline__ int halo_index(int i, int j, int x0, int y0) {
    return (i - (x0 - 2)) * kHalo + (j - (y0 - 2));
}

// A node halo staged in LDS in two parts: load() issues this thread's loads
// of R raw fields (thread e < kHalo^2 takes halo entry e, the others repeat
// the last entry, so no load or LDS write sits in a branch and the loads
// overlap the bin list's), store() writes F fields computed from them once
// they arrive.  raw(f, k) = raw field f at grid index k; value(f, r) = staged
// field f from the raw values r.  Nodes outside the grid stage 0.
static_assert(kHalo * kHalo <= kT, "one halo entry per thread");
template <int R>
struct Halo {
    float r[R];
    int e;
    bool ok;
    template <class Raw>
    __device__ __forceinline__ void load(int x0, int y0, int gsx, int gsy, const Raw &raw) {
        e = min(int(threadIdx.x), kHalo * kHalo - 1);
        const int i = x0 - 2 + e / kHalo, j = y0 - 2 + e % kHalo;
        ok = node_ok(i, j, gsx, gsy);
        const int64_t k = ok ? int64_t(i) * gsy + j : 0;
        for (int f = 0; f < R; ++f) r[f] = raw(f, k);
    }
    template <int F, class Value>
    __device__ __forceinline__ void store(float (&sm)[F][kHalo * kHalo], const Value &value) const {
        for (int f = 0; f < F; ++f) sm[f][e] = ok ? value(f, r) : 0.0f;
    }
};

struct P2gItem {
    double x, y, s, r, g, b, a;
    int mature;
};

// Mass (= density), density gradient and strength-weighted colour of the
// tile's nodes from the moments of the cells that reach them.
//
// The block first counting-sorts its chunk by cell (the kReach x
// kReach cells in reach, then the entries out of reach) through LDS, sums
// the moments in that order, and leaves the chunk of the bin list in that
// order for k_forces and k_g2p.  Particles move about two cells per step,
// so the frame's cell order is gone after one step; without the sort, runs
// of lanes on one cell (summed with DPP before one LDS atomic per run)
// shrink to about one lane and every lane pays its own atomics.
//
// Six waves per SIMD (79 VGPRs): the LDS (18 KB) allows eight blocks per
// CU; seven waves spill 12 bytes per lane, which measured slower, and
// without a bound the node assembly's scheduling takes 95 VGPRs (five).
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(6))) void k_p2g(Particles P, Grid Gd, int gsx,
                                                                                   int gsy, Bins B) {
    __shared__ float cs[kCellSums][kReach * kReach];
    constexpr int kCells = kReach * kReach;  // local cells; bucket kCells: out of reach
    __shared__ int cellpos[kCells + 2];  // per bucket: count, then first place
    __shared__ int2 mine[kT];            // per thread: its list entry, (bucket << 8) | rank (-1: no entry)
    __shared__ int perm[kT];             // sorted place -> thread
    // The chunk's items (position and colour terms), one array per field so
    // that consecutive lanes write consecutive words; dead once the moments
    // are summed, so they share their space with the node assembly's partials.
    union Stage {
        struct {
            double x[kT], y[kT];
            float t[5][kT];
        } item;
        float part[3][8][kTile * kTile];
    };
    __shared__ Stage stage;
    static_assert(kCells + 2 <= 2 * 64, "one wave scans the buckets, two per lane");
    const int nchunks = *B.nchunks;
    // the step before's counts, read by every placing block, are cleared
    // here for the next step's counting
    for (int k = blockIdx.x * kT + threadIdx.x; k < B.ntx * B.nty; k += gridDim.x * kT) B.count[B.par ^ 1][k] = 0;
    for (int item = int(blockIdx.x);; item += int(gridDim.x)) {
        int tile, slot0;
        bool single, home0;
        if (!tile_item(B, nchunks, item, tile, single, slot0, home0)) break;
        const int64_t s0 = int64_t(item) * kT;  // the item's entries in the list
        const int tx = tile / B.nty, ty = tile % B.nty;
        const int x0 = tx * kTile, y0 = ty * kTile;
        // every field load is unconditional (slot -1, past the end of the
        // item, reads particle 0 and is masked), so the eight loads are issued
        // together, and the sums are cleared while they are in flight
        auto fetch = [&](int slot) {
            const int q = slot >= 0 ? slot : 0;
            return P2gItem{P.x[q], P.y[q], P.strength[q], P.r[q], P.g[q], P.b[q], P.a[q], P.mature[q]};
        };
        const auto clear = [&] {
            for (int k = threadIdx.x; k < kCellSums * kReach * kReach; k += kT) (&cs[0][0])[k] = 0.0f;
            if (threadIdx.x < kCells + 2) cellpos[threadIdx.x] = 0;
            // an empty item (a tile that has just emptied) never runs the body
            mine[threadIdx.x] = make_int2(0, -1);
            __syncthreads();
        };
        // A particle's colour terms: strength-weighted running mean ==
        // sum(s c) / sum(s), over mature particles of positive strength.
        const auto tints = [](const P2gItem &it, float (&t)[5]) {
            const bool tint = it.mature && it.s > 0.0;
            t[0] = tint ? float(it.s * it.r) : 0.0f;
            t[1] = tint ? float(it.s * it.g) : 0.0f;
            t[2] = tint ? float(it.s * it.b) : 0.0f;
            t[3] = tint ? float(it.s * it.a) : 0.0f;
            t[4] = tint ? float(it.s) : 0.0f;
        };
        // The moments of one particle at (x, y) with colour terms t (valid:
        // in reach of the tile) added to its cell's LDS sums; every lane of
        // the block calls it.  Branch-free: selects instead of conditional
        // blocks, so no load is deferred into a branch.
        const auto moments = [&](bool valid, double x, double y, const float (&t)[5]) {
            const int cx = static_cast<int>(x - 0.5), cy = static_cast<int>(y - 0.5);
            float c[kCellSums];
            const double e = x - cx - 1.0, h = y - cy - 1.0;
            const double ep[3] = {1.0, e, e * e}, hp[3] = {1.0, h, h * h};
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) c[p * 3 + q] = valid ? float(ep[p] * hp[q]) : 0.0f;
            for (int k = 0; k < 5; ++k) c[9 + k] = valid ? t[k] : 0.0f;
            const Seg S = segments(valid, cx, cy);
            for (int k = 0; k < kCellSums; ++k) c[k] = seg_sum(c[k], S);
            if (valid && S.tail) {
                const int idx = (cx - x0 + 2) * kReach + (cy - y0 + 2);
                for (int k = 0; k < kCellSums; ++k)
                    if (c[k] != 0.0f) atomicAdd(&cs[k][idx], c[k]);
            }
        };
        {
            // at most kT entries: one pass, ranks fit in 8 bits
            run_item(slot0, home0, fetch, [&](int slot, bool home, const P2gItem &it) {
                const int cx = static_cast<int>(it.x - 0.5), cy = static_cast<int>(it.y - 0.5);
                const int bucket = in_reach(cx, cy, x0, y0) ? (cx - x0 + 2) * kReach + (cy - y0 + 2) : kCells;
                if (slot >= 0) {
                    const int l = threadIdx.x;
                    mine[l] = make_int2(slot | (home ? kHome : 0), (bucket << 8) | atomicAdd(&cellpos[bucket], 1));
                    float t[5];
                    tints(it, t);
                    stage.item.x[l] = it.x;
                    stage.item.y[l] = it.y;
                    for (int k = 0; k < 5; ++k) stage.item.t[k][l] = t[k];
                }
            }, clear);
            __syncthreads();
            if (threadIdx.x < 64) {  // exclusive scan of the bucket counts into first places
                const int lane = int(threadIdx.x);
                const int v0 = 2 * lane < kCells + 1 ? cellpos[2 * lane] : 0;
                const int v1 = 2 * lane + 1 < kCells + 1 ? cellpos[2 * lane + 1] : 0;
                int incl = v0 + v1;
                for (int off = 1; off < 64; off <<= 1) {
                    const int v = __shfl_up(incl, off, 64);
                    if (lane >= off) incl += v;
                }
                const int excl = incl - v0 - v1;
                if (2 * lane < kCells + 1) cellpos[2 * lane] = excl;
                if (2 * lane + 1 < kCells + 1) cellpos[2 * lane + 1] = excl + v0;
            }
            __syncthreads();
            {
                const int2 m = mine[threadIdx.x];
                if (m.y >= 0) {
                    const int dest = cellpos[m.y >> 8] + (m.y & 255);
                    perm[dest] = int(threadIdx.x);
                    B.list[s0 + dest] = m.x;
                }
            }
            __syncthreads();
            // sorted place threadIdx.x; the entries in reach come first
            const bool in = int(threadIdx.x) < cellpos[kCells];
            const int src = in ? perm[threadIdx.x] : 0;
            // place 0's item may be stale or unset: keep the position finite
            const double x = in ? stage.item.x[src] : 0.0, y = in ? stage.item.y[src] : 0.0;
            float t[5];
            for (int k = 0; k < 5; ++k) t[k] = stage.item.t[k][src];
            moments(in, x, y, t);
        }
        __syncthreads();
        // Nodes from the cells: thread (a, node) sums the three cells
        // (node - (a, b)), then one thread per node adds the three partials.
        auto &part = stage.part;
        static_assert(3 * kTile * kTile <= kT, "one thread per (a, node)");
        {
            const int t = threadIdx.x, node = t % (kTile * kTile), a = t / (kTile * kTile);
            if (a < 3) {
                const int ni = x0 + node / kTile, nj = y0 + node % kTile;
                float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                for (int b = 0; b < 3; ++b) {
                    // an empty cell's sums are zero and add nothing; no branch on
                    // it, so the three cells' LDS loads are in flight together
                    const int cell = (ni - a - x0 + 2) * kReach + (nj - b - y0 + 2);
                    float m[kCellSums];
                    for (int k = 0; k < kCellSums; ++k) m[k] = cs[k][cell];
                    float v[8];
                    cell_to_node(m, a, b, v);
                    for (int f = 0; f < 8; ++f) acc[f] += v[f];
                }
                for (int f = 0; f < 8; ++f) part[a][f][node] = acc[f];
            }
        }
        __syncthreads();
        if (threadIdx.x < kTile * kTile) {
            const int node = threadIdx.x, ni = x0 + node / kTile, nj = y0 + node % kTile;
            if (ni < gsx && nj < gsy) {
                const int64_t k = int64_t(ni) * gsy + nj;
                // fields m, gx, gy, r, g, b, a, w: their arrays and indices
                float *f[8] = {Gd.muvw, Gd.gx, Gd.gy, Gd.col, Gd.col, Gd.col, Gd.col, Gd.muvw};
                const int64_t at[8] = {4 * k, k, k, 4 * k, 4 * k + 1, 4 * k + 2, 4 * k + 3, 4 * k + 3};
                for (int q = 0; q < 8; ++q)
                    put_node(f[q], at[q], part[0][q][node] + part[1][q][node] + part[2][q][node], single);
            }
        }
        __syncthreads();  // the LDS sums are reused by the next item
    }
}

// Sums v (zeros for lanes that do not contribute) over runs of lanes with
// the same stencil origin and adds the run totals to node (cx + a, cy + b)
// of the tile's LDS sums when that node belongs to the tile.
__device__ __forceinline__ void tile_add2(float (&acc)[2][kTile * kTile], const Seg &S, bool valid, int cx, int cy,
                                          int x0, int y0, int a, int b, float v0, float v1) {
    v0 = seg_sum(v0, S);
    v1 = seg_sum(v1, S);
    const int li = cx + a - x0, lj = cy + b - y0;
    if (valid && S.tail && li >= 0 && li < kTile && lj >= 0 && lj < kTile) {
        if (v0 != 0.0f) atomicAdd(&acc[0][li * kTile + lj], v0);
        if (v1 != 0.0f) atomicAdd(&acc[1][li * kTile + lj], v1);
    }
}

// Stores (or adds) the tile's two LDS sums into fields f0, f1: node k's
// values at f0[stride*k], f1[stride*k].
__device__ __forceinline__ void store_tile2(const float (&acc)[2][kTile * kTile], float *f0, float *f1, int stride,
                                            int x0, int y0, int gsx, int gsy, bool single) {
    for (int t = threadIdx.x; t < kTile * kTile; t += kT) {
        const int ni = x0 + t / kTile, nj = y0 + t % kTile;
        if (ni >= gsx || nj >= gsy) continue;
        const int64_t k = int64_t(ni) * gsy + nj;
        put_node(f0, stride * k, acc[0][t], single);
        put_node(f1, stride * k, acc[1][t], single);
    }
}

// Node accelerations: pressure and wall forces (fluidmodel.cpp:251-357).
__global__ __launch_bounds__(kT) void k_forces(Particles P, Grid Gd, int gsx, int gsy, Bins B) {
    __shared__ float acc[2][kTile * kTile];
    __shared__ float nodes[3][kHalo * kHalo];  // m, gx, gy
    const int nchunks = *B.nchunks;
    for (int item = int(blockIdx.x);; item += int(gridDim.x)) {
        int tile, slot0;
        bool single, home0;
        if (!tile_item(B, nchunks, item, tile, single, slot0, home0)) break;
        const int tx = tile / B.nty, ty = tile % B.nty;
        const int x0 = tx * kTile, y0 = ty * kTile;
        for (int k = threadIdx.x; k < 2 * kTile * kTile; k += kT) (&acc[0][0])[k] = 0.0f;
        // the halo's loads are issued beside the particles', which are
        // unconditional (slot -1 reads particle 0)
        Halo<3> halo;
        halo.load(x0, y0, gsx, gsy,
                  [&](int f, int64_t k) { return f == 0 ? Gd.muvw[4 * k] : f == 1 ? Gd.gx[k] : Gd.gy[k]; });
        const auto stage = [&] {
            halo.store(nodes, [](int f, const float *r) { return r[f]; });
            __syncthreads();
        };
        const auto node = [&](int f, int i, int j) { return nodes[f][halo_index(i, j, x0, y0)]; };
        auto fetch = [&](int slot) {
            const int q = slot >= 0 ? slot : 0;
            return double2{P.x[q], P.y[q]};
        };
        run_item(slot0, home0, fetch, [&](int slot, bool, const double2 &it) {
            int cx, cy;
            double px[3], py[3], gx[3], gy[3];
            double pressure = 0.0, fx = 0.0, fy = 0.0;
            const double x = it.x, y = it.y;
            weights(x, y, cx, cy, px, py, gx, gy);
            const bool valid = slot >= 0 && in_reach(cx, cy, x0, y0);
            if (valid) particle_forces(x, y, node, gsx, gsy, pressure, fx, fy);  // LDS reads only
            const Seg S = segments(valid, cx, cy);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double phi = px[a] * py[b];
                    const float v0 = valid ? float(-((gx[a] * py[b]) * pressure) + fx * phi) : 0.0f;
                    const float v1 = valid ? float(-((px[a] * gy[b]) * pressure) + fy * phi) : 0.0f;
                    tile_add2(acc, S, valid, cx, cy, x0, y0, a, b, v0, v1);
                }
        }, stage);
        __syncthreads();
        store_tile2(acc, Gd.ax, Gd.ay, 1, x0, y0, gsx, gsy, single);
        __syncthreads();  // the LDS sums are reused by the next item
    }
}

struct G2pItem {
    double x, y, gx, gy, u, v;
    int mature;
};

// Particle velocity from node accelerations plus the attractor (written by
// the particle's home tile), then the momentum of the tile's nodes
// (fluidmodel.cpp:359-437).
__global__ __launch_bounds__(kT) void k_g2p(Particles P, Grid Gd, int gsx, int gsy, Bins B) {
    __shared__ float acc[2][kTile * kTile];
    __shared__ float accel[2][kHalo * kHalo];  // node accelerations per mass
    const int nchunks = *B.nchunks;
    for (int item = int(blockIdx.x);; item += int(gridDim.x)) {
        int tile, slot0;
        bool single, home0;
        if (!tile_item(B, nchunks, item, tile, single, slot0, home0)) break;
        const int tx = tile / B.nty, ty = tile % B.nty;
        const int x0 = tx * kTile, y0 = ty * kTile;
        for (int k = threadIdx.x; k < 2 * kTile * kTile; k += kT) (&acc[0][0])[k] = 0.0f;
        // as in k_forces: halo loads beside the unconditional particle loads
        Halo<3> halo;
        halo.load(x0, y0, gsx, gsy,
                  [&](int f, int64_t k) { return f == 0 ? Gd.ax[k] : f == 1 ? Gd.ay[k] : Gd.muvw[4 * k]; });
        const auto stage = [&] {
            halo.store(accel, [](int f, const float *r) { return per_mass(r[f], r[2]); });
            __syncthreads();
        };
        auto fetch = [&](int slot) {
            const int q = slot >= 0 ? slot : 0;
            return G2pItem{P.x[q], P.y[q], P.gx[q], P.gy[q], P.u[q], P.v[q], P.mature[q]};
        };
        run_item(slot0, home0, fetch, [&](int slot, bool home, const G2pItem &it) {
            int cx, cy;
            double px[3], py[3], gx[3], gy[3];
            const double x = it.x, y = it.y;
            weights(x, y, cx, cy, px, py, gx, gy);
            const bool valid = slot >= 0 && in_reach(cx, cy, x0, y0);
            double cax, cay;  // attractor, 0.03 per step (fluidmodel.cpp:383-410)
            toward(x, y, it.gx, it.gy, 0.03, cax, cay);
            double u = it.u, v = it.v;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double phi = px[a] * py[b];
                    double ax = 0.0, ay = 0.0;
                    if (valid) {  // in reach: the node is staged
                        const int e = halo_index(cx + a, cy + b, x0, y0);
                        ax = accel[0][e];
                        ay = accel[1][e];
                    } else if (home) {
                        ax = nd_m(Gd.ax, Gd.muvw, cx + a, cy + b, gsx, gsy);
                        ay = nd_m(Gd.ay, Gd.muvw, cx + a, cy + b, gsx, gsy);
                    }
                    u += phi * (ax + cax);
                    v += phi * (ay + cay);
                }
            if (home) {  // other tiles still read the old velocity
                P.u2[slot] = u;
                P.v2[slot] = v;
            }
            const bool act = valid || home;
            const double mu = act ? (it.mature ? u : u * 0.0) : 0.0;
            const double mv = act ? (it.mature ? v : v * 0.0) : 0.0;
            const Seg S = segments(valid, cx, cy);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double phi = px[a] * py[b];
                    const float v0 = valid ? float(phi * mu) : 0.0f;
                    const float v1 = valid ? float(phi * mv) : 0.0f;
                    tile_add2(acc, S, valid, cx, cy, x0, y0, a, b, v0, v1);
                }
        }, stage);
        __syncthreads();
        store_tile2(acc, Gd.muvw + 1, Gd.muvw + 2, 4, x0, y0, gsx, gsy, single);
        __syncthreads();  // the LDS sums are reused by the next item
    }
}

// The values that change from frame to frame, read by the frame's kernels
// from device memory, so that one captured graph replays every frame.
struct FrameParams {
    Frame F;           // time of the frame
    double radius;     // freedom radius of its steps
    uint64_t frame;    // frames rendered before (the wall jitter's seed)
    uint32_t *out;     // the frame's image
};

__global__ void k_set_params(FrameParams *dst, FrameParams v) { *dst = v; }

// An atom's PERLIN lag and slope as last computed (atom_lag_slope), for the
// source point p1 under the tables and bounding box with fingerprint tag
// (never 0, so zeroed memory holds nothing).  They stay the same for all
// the frames of a key segment.
struct LagSlopeMemo {
    double lag, slope;
    uint64_t p1, tag;
};

// Wall jitter seed of step k of a frame (never 0, which means none).
__device__ __forceinline__ uint64_t step_seed(uint64_t frame, int k) { return mix64(frame * 131 + uint64_t(k)) | 1u; }

__device__ __forceinline__ double jitter(uint64_t seed, int i, int k) {
    // replaces rand() / RAND_MAX * 0.01 of the wall bounce (fluidmodel.cpp:554-564);
    // seed 0 bounces without jitter
    if (seed == 0) return 0.0;
    return double(mix64(seed ^ (uint64_t(i) << 3) ^ uint64_t(k)) >> 11) * (1.0 / 9007199254740992.0) * 0.01;
}

// v where ok, else zeros: selects, not a branch, so the load of v is not
// waited for before the next loads are issued.
__device__ __forceinline__ float4 keep_if(bool ok, float4 v) {
    return float4{ok ? v.x : 0.0f, ok ? v.y : 0.0f, ok ? v.z : 0.0f, ok ? v.w : 0.0f};
}

// The frame's splat arguments.
struct SplatArgs {
    int W, H;
    double sp;
    int bx1, by1, bx2, by2;  // the blob bounding box of the frame
    uint32_t *xy;            // [n] X | Y << 16 of the top-left tap, kNoTap = not drawn
    double *taps;            // [n][4] tap weights
    uint32_t *cols;          // [n] packed colour
};

// An active particle's splat from its values (strength s, position, colour
// targets R0..A0 and colours r0..a0): the colour update, top-left tap pixel
// (xy[i], kNoTap when nothing is drawn), four tap weights and packed
// colour.  Returns xy[i].
__device__ __forceinline__ uint32_t splat_core(Particles P, int i, double t, const SplatArgs &a, double s, double x,
                                               double y, double R0, double G0, double B0, double A0, double r0,
                                               double g0, double b0, double a0) {
    const int W = a.W, H = a.H, bx1 = a.bx1, by1 = a.by1, bx2 = a.bx2, by2 = a.by2;
    const double sp = a.sp;
    uint32_t *__restrict__ xy = a.xy;
    double *__restrict__ taps = a.taps;
    uint32_t *__restrict__ cols = a.cols;
    const double w = (1.0 / (100.0 * pow((t - (1.0 * (1.0 - s))), 2.0) + 1.0)) * pow(s, 2.0);
    const double r = (w * R0) + (1.0 - w) * r0, g = (w * G0) + (1.0 - w) * g0;
    const double b = (w * B0) + (1.0 - w) * b0, al = (w * A0) + (1.0 - w) * a0;
    P.r[i] = r; P.g[i] = g; P.b[i] = b; P.a[i] = al;
    double px, py;
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
    if ((ix >= W || iy >= H) && (ix > bx2 || ix < bx1 || iy > by2 || iy < by1)) {
        xy[i] = kNoTap;
        return kNoTap;
    }
    const double w11 = ((1.0 - xf) * (1.0 - yf)), w21 = (xf * (1.0 - yf));
    const double w12 = ((1.0 - xf) * yf), w22 = (xf * yf);
    // a tap is drawn when its weight is positive, its guard holds and it lies
    // inside the image; taps (X,Y) (X+1,Y) (X,Y+1) (X+1,Y+1)
    double wt[4] = {0.0, 0.0, 0.0, 0.0};
    if (w11 > 0.0) wt[0] = w11;
    if ((X < bx2 || X + 1 < W) && w21 > 0.0 && X + 1 < W) wt[1] = w21;
    if ((Y < by2 || Y + 1 < H) && w12 > 0.0 && Y + 1 < H) wt[2] = w12;
    if (w22 > 0.0 && ((Y < by2 && X < bx2) || (Y + 1 < H && X + 1 < W)) && X + 1 < W && Y + 1 < H) wt[3] = w22;
    const uint32_t v = uint32_t(X) | (uint32_t(Y) << 16);
    xy[i] = v;
    for (int q = 0; q < 4; ++q) taps[int64_t(i) * 4 + q] = wt[q];
    cols[i] = rgba_pack(narrow8(round(r * 255.0)), narrow8(round(g * 255.0)), narrow8(round(b * 255.0)),
                        narrow8(round(al * 255.0)));
    return v;
}

// Particle i's splat (splat_core from its stored values; xy[i] = kNoTap
// when it is not active).  Returns xy[i].
__device__ __forceinline__ uint32_t splat_one(Particles P, int i, double t, const SplatArgs &a) {
    // every load before the first store (see k_fluid_update)
    const int act = P.active[i];
    const double s = P.strength[i], x = P.x[i], y = P.y[i];
    const double R0 = P.R[i], G0 = P.G[i], B0 = P.B[i], A0 = P.A[i];
    const double r0 = P.r[i], g0 = P.g[i], b0 = P.b[i], a0 = P.a[i];
    if (!act) {
        a.xy[i] = kNoTap;
        return kNoTap;
    }
    return splat_core(P, i, t, a, s, x, y, R0, G0, B0, A0, r0, g0, b0, a0);
}

// One particle's advection (fluidmodel.cpp:443-564); returns its new
// position.  atom: the particle held in each slot when P is the cell-ordered
// copy (NULL: slot i holds particle i); the wall jitter is keyed by particle.
// Every load is unconditional and issued up front (nodes outside the grid
// read node 0 and count as zero), so the particle costs two memory round
// trips -- its fields, then its nine nodes -- instead of one per node.
// The values of an advected particle that its splat takes (the frame's last
// step splats each particle right after moving it).
struct Advected {
    double s;               // strength
    double R, G, B, A;      // colour targets
    double r, g, b, a;      // colours after the diffusion
};

__device__ __forceinline__ double2 advect_one(Particles P, int i, const Grid &Gd, int gsx, int gsy, int steps_left,
                                              double freedom_radius, uint64_t seed, const int32_t *__restrict__ atom,
                                              Advected *out = nullptr) {
    const int id = atom ? atom[i] : i;
    double x = P.x[i], y = P.y[i];
    const int mature = P.mature[i];
    const double R0 = P.R[i], G0 = P.G[i], B0 = P.B[i], A0 = P.A[i];
    const double r0 = P.r[i], g0 = P.g[i], b0 = P.b[i], a0 = P.a[i];
    const double tgx = P.gx[i], tgy = P.gy[i], fr = P.fr[i], u2 = P.u2[i], v2 = P.v2[i];
    if (out) {
        out->s = P.strength[i];
        out->R = R0; out->G = G0; out->B = B0; out->A = A0;
    }
    int cx, cy;
    double px[3], py[3], gx[3], gy[3];
    weights(x, y, cx, cy, px, py, gx, gy);
    float4 q[3][3], c[3][3];  // m, u, v, w and r, g, b, a of the stencil nodes
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const bool ok = node_ok(cx + a, cy + b, gsx, gsy);
            const int64_t k = ok ? int64_t(cx + a) * gsy + cy + b : 0;
            q[a][b] = keep_if(ok, reinterpret_cast<const float4 *>(Gd.muvw)[k]);
            c[a][b] = keep_if(ok, reinterpret_cast<const float4 *>(Gd.col)[k]);
        }
    __builtin_amdgcn_sched_barrier(0);  // all 18 node loads are in flight before the first use
    double gu = 0, gv = 0, nR = 0, nG = 0, nB = 0, nA = 0, wsum = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const double phi = px[a] * py[b];
            gu += phi * per_mass(q[a][b].y, q[a][b].x);
            gv += phi * per_mass(q[a][b].z, q[a][b].x);
            // node colour = sum(s c)/sum(s), weighted by sum(s), over the
            // nodes with w > 0 (selects rather than a branch per node)
            const double w = q[a][b].w;
            const bool wpos = w > 0.0;
            wsum += wpos ? w : 0.0;
            const float4 cw = keep_if(wpos, c[a][b]);
            nR += cw.x;
            nG += cw.y;
            nB += cw.z;
            nA += cw.w;
        }
    {   // colour diffusion (fluidmodel.cpp:470-495) where wsum > 0; computed
        // and stored unconditionally, so no load waits inside a branch
        const bool diffuse = wsum > 0.0;
        const double ws = diffuse ? wsum : 1.0;
        nR /= ws; nG /= ws; nB /= ws; nA /= ws;
        const double wr = fabs(nR - R0), wg = fabs(nG - G0), wb = fabs(nB - B0), wa = fabs(nA - A0);
        const double r1 = mature ? (1.0 - wr) * r0 + wr * nR : nR, g1 = mature ? (1.0 - wg) * g0 + wg * nG : nG;
        const double b1 = mature ? (1.0 - wb) * b0 + wb * nB : nB, a1 = mature ? (1.0 - wa) * a0 + wa * nA : nA;
        const double r = diffuse ? r1 : r0, g = diffuse ? g1 : g0, b = diffuse ? b1 : b0, a = diffuse ? a1 : a0;
        if (out) {  // the splat stores the colours it relaxes from these
            out->r = r; out->g = g; out->b = b; out->a = a;
        } else {
            P.r[i] = r; P.g[i] = g; P.b[i] = b; P.a[i] = a;
        }
    }
    x += gu;
    y += gv;
    {   // freedom radius (fluidmodel.cpp:501-539)
        const double A = fabs(y - tgy), B = fabs(x - tgx), Cd = sqrt(A * A + B * B);
        const double r = freedom_radius * fr;
        if (Cd > r) {
            double mx, my;
            toward(x, y, tgx, tgy, Cd - r, mx, my);
            const double w = 1.0 / (steps_left + 1);
            x += mx * w;
            y += my * w;
        }
    }
    double u = u2, v = v2;
    u += gu - u;
    v += gv - v;
    if (x < 1.0) { x = 1.0 + jitter(seed, id, 0); u = 0.0; }
    else if (x > gsx - 2) { x = gsx - 2 - jitter(seed, id, 1); u = 0.0; }
    if (y < 1.0) { y = 1.0 + jitter(seed, id, 2); v = 0.0; }
    else if (y > gsy - 2) { y = gsy - 2 - jitter(seed, id, 3); v = 0.0; }
    P.x[i] = x; P.y[i] = y; P.u[i] = u; P.v[i] = v;
    return double2{x, y};
}

// Advection, one thread per particle slot.  (A per-tile form over the home
// entries with the node halo in LDS measured slower: the particle fields it
// gathers through the bin list cost more than the node reads it saves.)
// With count_par >= 0 the moved particles are also counted into the next
// step's bins (B.count[count_par]), which saves that step's counting pass.
// kSplat (a frame's last step, fp set): each particle is splatted once
// moved (k_splat_prep's work, counting its taps into the pixel-tile bins
// B.count[count_par]).
template <bool kSplat>
__global__ __launch_bounds__(kT) void k_advect(Particles P, int n, Grid Gd, int gsx, int gsy, int steps_left,
                                               double freedom_radius, uint64_t seed,
                                               const int32_t *__restrict__ atom, Bins B, int count_par,
                                               const FrameParams *__restrict__ fp, int step_k, SplatArgs sa) {
    if (fp) {  // a frame's step: radius and jitter seed of the frame
        freedom_radius = fp->radius;
        seed = step_seed(fp->frame, step_k);
    }
    const int i = blockIdx.x * kT + threadIdx.x;
    const bool act = i < n && P.active[i];
    int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
    bool counted = act;
    if (act) {
        if (kSplat) {
            Advected o;
            const double2 q = advect_one(P, i, Gd, gsx, gsy, steps_left, freedom_radius, seed, atom, &o);
            const uint32_t v = splat_core(P, i, fp->F.tl, sa, o.s, q.x, q.y, o.R, o.G, o.B, o.A, o.r, o.g, o.b, o.a);
            counted = v != kNoTap;
            if (counted) splat_tiles(v, sa.W, sa.H, tx0, tx1, ty0, ty1);
        } else {
            const double2 q = advect_one(P, i, Gd, gsx, gsy, steps_left, freedom_radius, seed, atom);
            stencil_tiles(q.x, q.y, gsx, gsy, tx0, tx1, ty0, ty1);
        }
    } else if (kSplat && i < n) {
        sa.xy[i] = kNoTap;
    }
    if (count_par >= 0) count_entries(counted, tx0, tx1, ty0, ty1, i, n, B, count_par);
}


__global__ void k_fill_i32(int32_t *p, int64_t n) {
    const int64_t k = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (k < n) p[k] = 0;
}

// --------------------------------------------------------- cell order
// Once per frame the particles are put in cell order (a counting sort by
// the Morton rank of the cell of their stencil origin), so a block's
// particles fill a compact patch of the grid: the LDS tiles of the scatters
// stay small and the node reads of a wave share cache lines.  Inactive
// particles go to one bucket after the last cell.
__device__ __forceinline__ int cell_of(double x, double y, int gsx, int gsy) {
    const int cx = min(max(static_cast<int>(x - 0.5), 0), gsx - 1);
    const int cy = min(max(static_cast<int>(y - 0.5), 0), gsy - 1);
    return cx * gsy + cy;
}

// Inactive particles go after every cell, spread over this many buckets by
// wave: one bucket for all of them made every wave's atomic hit the same
// counter, and those serialise.
constexpr int kIdleBuckets = 256;

__device__ __forceinline__ int idle_bucket(int i, int nn) {
    return nn + int((uint32_t(i) >> 6) & uint32_t(kIdleBuckets - 1));
}

// Particle i's sort bucket `key` (in: i < n) counted into cnt[key], its rank
// among the bucket's particles into pos[i] and the key into cell[i].  Every
// lane of the wave calls it.
__device__ __forceinline__ void sort_rank(int key, bool in, int i, int32_t *__restrict__ cell,
                                          int32_t *__restrict__ pos, int32_t *__restrict__ cnt) {
    // group the wave's lanes by key: one atomic per distinct key, all in flight together
    const int lane = threadIdx.x & 63;
    int lead = lane, pre = 0, total = 0;
    uint64_t todo = __ballot(in);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const int kl = __shfl(key, leader, 64);
        const uint64_t same = __ballot(key == kl);
        if (key == kl) {
            lead = leader;
            pre = int(__popcll(same & ((1ull << lane) - 1ull)));
        }
        if (lane == leader) total = int(__popcll(same));
        todo &= ~same;
    }
    const int base = total > 0 ? atomicAdd(&cnt[key], total) : 0;
    const int b = __shfl(base, lead, 64);
    if (!in) return;
    cell[i] = key;
    pos[i] = b + pre;
}

// The cell sort's counting pass done by another kernel (k_sort_count's work): the
// buckets, ranks and keys it writes.  cnt == nullptr: not counted.
struct SortCount {
    const int32_t *cell_rank;
    int32_t *cell, *pos, *cnt;
    int gsx, gsy;
};

__global__ __launch_bounds__(kT) void k_sort_count(Particles P, int n, int gsx, int gsy,
                                                  const int32_t *__restrict__ cell_rank, int32_t *__restrict__ cell,
                                                  int32_t *__restrict__ pos, int32_t *__restrict__ cnt) {
    const int i = blockIdx.x * kT + threadIdx.x;
    const bool in = i < n;
    const int key = !in ? -1 : P.active[i] ? cell_rank[cell_of(P.x[i], P.y[i], gsx, gsy)] : idle_bucket(i, gsx * gsy);
    sort_rank(key, in, i, cell, pos, cnt);
}

// The new slot t of the particle in slot i of P: S.atom[t] = P.atom[i]
// and prev[t] = i.  The buckets' counts, scanned already, are cleared for
// the next frame's sort.
__global__ __launch_bounds__(kT) void k_sort_slots(int n, const int32_t *__restrict__ cell,
                                                  const int32_t *__restrict__ pos, const int32_t *__restrict__ start,
                                                  const int32_t *__restrict__ atom, int32_t *__restrict__ prev,
                                                  Particles S, int32_t *__restrict__ cnt, int64_t nk) {
    const int i = blockIdx.x * kT + threadIdx.x;
    for (int64_t k = i; k < nk; k += int64_t(gridDim.x) * kT) cnt[k] = 0;
    if (i >= n) return;
    const int t = start[cell[i]] + pos[i];
    S.atom[t] = atom[i];
    prev[t] = i;
}

// S = P in cell order, one thread per new slot t reading slot prev[t] of P:
// the fields are written coalesced, and since P is in the last frame's cell
// order the reads of neighbouring slots mostly share cache lines.  Inactive
// particles are copied too (S becomes the particle state, and their values
// are kept while they rest).  par >= 0: the frame's first step's bin
// entries are counted too (as k_bin_count does, into B.count[par] and
// B.rank), from the position just gathered.
__global__ __launch_bounds__(kT) void k_sort_gather(Particles P, int n, const int32_t *__restrict__ prev, Particles S,
                                                   int gsx, int gsy, Bins B, int par) {
    const int t = blockIdx.x * kT + threadIdx.x;
    const bool in = t < n;
    int act = 0;
    double x = 0.0, y = 0.0;
    if (in) {
        const int i = prev[t];
        act = P.active[i];
        S.active[t] = act;
        // one field at a time: with all 18 random gathers of a wave in flight at
        // once the L2 held fewer of their lines
        x = P.x[i];
        y = P.y[i];
        S.x[t] = x; S.y[t] = y; S.u[t] = P.u[i]; S.v[t] = P.v[i];
        S.gx[t] = P.gx[i]; S.gy[t] = P.gy[i]; S.fr[t] = P.fr[i];
        S.R[t] = P.R[i]; S.G[t] = P.G[i]; S.B[t] = P.B[i]; S.A[t] = P.A[i];
        S.r[t] = P.r[i]; S.g[t] = P.g[i]; S.b[t] = P.b[i]; S.a[t] = P.a[i];
        S.strength[t] = P.strength[i];
        S.mature[t] = P.mature[i];
        S.owner[t] = P.owner[i];
        S.key[t] = P.key[i];
    }
    if (par < 0) return;  // uniform: no step follows
    int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
    if (in) stencil_tiles(x, y, gsx, gsy, tx0, tx1, ty0, ty1);
    count_entries(in && act != 0, tx0, tx1, ty0, ty1, t, n, B, par);  // every lane of the wave
}



}  // namespace

// ------------------------------------------------------------- the model
struct amh_fluid_s {
    int32_t gsx = 0, gsy = 0, count = 0;
    void *buf = nullptr;       // one allocation for particles + grid + render
    // The particle state is kept in the order of the last cell sort: slot i
    // of P holds atom P.atom[i] (the identity until the first sort).  A
    // frame sorts P into S, steps and draws S, and the two swap roles.
    Particles P;
    Particles S;
    int32_t swapped = 0;       // P and S have swapped roles an odd number of times
    Grid G;
    // cell ordering: per particle its cell and rank in it; per cell counts
    // and starts ([nn + kIdleBuckets + 1]: the inactive particles last)
    int32_t *cell = nullptr, *pos = nullptr, *cnt = nullptr, *start = nullptr;
    int32_t *prev = nullptr;   // [n] per slot of the sorted copy: its slot in P before
    int32_t *scan_sums = nullptr, *scan_offs = nullptr;
    int32_t *cell_rank = nullptr;  // [nn] sort key of each cell (Morton order)
    Bins bins;                     // per-step binning by node tile
    Bins pbins;                    // the splat's binning by pixel tile (rank, list shared with bins)
    int32_t pbin_par = 0;
    uint32_t *sxy = nullptr;       // [n] splat: top-left tap X | Y << 16, kNoTap = not drawn
    double *staps = nullptr;       // [n][4] splat tap weights
    uint32_t *scol = nullptr;      // [n] splat colours
    // render buffers
    int32_t *mask = nullptr;   // [npx] blob membership / feather state
    uint32_t *img = nullptr;   // [npx] colours during feather
    int32_t W = 0, H = 0;
    uint64_t frame = 0;        // frames rendered (jitter seed)
    int32_t bin_par = 0;       // parity of the next step's bin counts
    FrameParams *fp = nullptr; // the current frame's values (device)
    LagSlopeMemo *memo = nullptr;  // [n] the atoms' PERLIN lag and slope, kept between frames
    // Captured frames: the launch sequence of one frame as a hipGraph, keyed
    // by everything its launches bake in (render arguments, steps, parities).
    struct Captured {
        std::vector<unsigned char> key;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    Captured graphs[4];
    uint64_t graph_clock = 0;
    bool use_graphs = true;
};

namespace {
// Upper bound of the chunks: every tile has at least one, and a particle
// has at most 4 entries.
size_t max_chunks(size_t tiles, size_t n) { return tiles + (4 * n + kT - 1) / kT + 1; }
// Entries of the list: kT per work item (the steps' items start on kT
// boundaries), at least the 4n of whole bins (the splat shares the list).
size_t list_entries(size_t tiles, size_t n) { return std::max(4 * n, size_t(kT) * max_chunks(tiles, n)); }

size_t fluid_bytes(int gsx, int gsy, int count, int W, int H, int nkeys) {
    const size_t n = size_t(count), nn = size_t(gsx) * gsy, npx = size_t(W) * H;
    const size_t nb = size_t(scan::blocks(int64_t(nn) + kIdleBuckets));
    const size_t tiles = size_t((gsx + kTile - 1) / kTile) * size_t((gsy + kTile - 1) / kTile);
    const size_t ptiles = size_t((W + kPxTile - 1) / kPxTile) * size_t((H + kPxTile - 1) / kPxTile);
    (void)nkeys;
    // every carved array is rounded up to 256 bytes: the last term covers that
    return 2 * (n * 8 * kParticleDoubles + n * 4 * kParticleInts) + n * 4 * 3 + (nn + kIdleBuckets + 1) * 4 * 2 +
           (2 * nb + 1) * 4 + nn * 4 + (3 * tiles + 1) * 4 + n * 4 * 4 + list_entries(tiles, n) * 4 + tiles * 4 +
           (3 * ptiles + 1) * 4 + n * (4 + 32 + 4) + max_chunks(tiles, n) * 16 + 4 + nn * 4 * kGridFields +
           npx * 4 * 2 + 256 + n * sizeof(LagSlopeMemo) + 128 * 256;
}

void carve_particles(Particles &Q, size_t n, const std::function<char *(size_t)> &take) {
    double **d[kParticleDoubles] = {&Q.x, &Q.y, &Q.u, &Q.v, &Q.gx, &Q.gy, &Q.fr, &Q.R, &Q.G,
                                    &Q.B, &Q.A, &Q.r, &Q.g, &Q.b, &Q.a, &Q.strength, &Q.u2, &Q.v2};
    for (double **q : d) *q = reinterpret_cast<double *>(take(n * 8));
    int32_t **ints[kParticleInts] = {&Q.active, &Q.mature, &Q.owner, &Q.key, &Q.atom};
    for (int32_t **q : ints) *q = reinterpret_cast<int32_t *>(take(n * 4));
}
}  // namespace

extern "C" int amh_fluid_create(void **fluid, int32_t gsx, int32_t gsy, int32_t count, int32_t width, int32_t height,
                                int32_t nkeys, void *stream) {
    // The splat packs a tap pixel as x | y << 16 with 0xFFFFFFFF as "no tap",
    // and cell keys pack (cx << 16) | cy: both need coordinates below 65535.
    if (!fluid || gsx < 8 || gsy < 8 || gsx >= 65535 || gsy >= 65535 || count < 0 || width <= 0 ||
        height <= 0 || width > 65535 || height > 65535 || nkeys <= 0) {
        set_error("amh_fluid_create: invalid argument (grid %dx%d, frame %dx%d, %d particles, %d keys)", gsx, gsy,
                  width, height, count, nkeys);
        return AMH_E_ARG;
    }
    amh_fluid_s *f = new amh_fluid_s;
    f->gsx = gsx; f->gsy = gsy; f->count = count; f->W = width; f->H = height;
    // AMH_FLUID_GRAPHS=0 in the environment starts the handle with graph
    // capture off (for tools that follow only eager launches)
    if (const char *g = std::getenv("AMH_FLUID_GRAPHS")) f->use_graphs = std::strcmp(g, "0") != 0;
    const size_t bytes = fluid_bytes(gsx, gsy, count, width, height, nkeys);
    if (int r = check_hip(hipMalloc(&f->buf, bytes), "amh_fluid_create")) {
        delete f;
        return r;
    }
    char *p = static_cast<char *>(f->buf);
    auto take = [&](size_t b) { char *q = p; p += (b + 255) & ~size_t(255); return q; };
    const size_t n = size_t(count), nn = size_t(gsx) * gsy, npx = size_t(width) * height;
    carve_particles(f->P, n, take);
    carve_particles(f->S, n, take);
    f->cell = reinterpret_cast<int32_t *>(take(n * 4));
    f->pos = reinterpret_cast<int32_t *>(take(n * 4));
    f->prev = reinterpret_cast<int32_t *>(take(n * 4));
    f->cnt = reinterpret_cast<int32_t *>(take((nn + kIdleBuckets + 1) * 4));
    f->start = reinterpret_cast<int32_t *>(take((nn + kIdleBuckets + 1) * 4));
    const size_t nb = size_t(scan::blocks(int64_t(nn) + kIdleBuckets));
    f->scan_sums = reinterpret_cast<int32_t *>(take(nb * 4));
    f->scan_offs = reinterpret_cast<int32_t *>(take((nb + 1) * 4));
    f->cell_rank = reinterpret_cast<int32_t *>(take(nn * 4));
    f->bins.ntx = (gsx + kTile - 1) / kTile;
    f->bins.nty = (gsy + kTile - 1) / kTile;
    const size_t tiles = size_t(f->bins.ntx) * size_t(f->bins.nty);
    f->bins.count[0] = reinterpret_cast<int32_t *>(take(tiles * 4));
    f->bins.count[1] = reinterpret_cast<int32_t *>(take(tiles * 4));
    f->bins.start = reinterpret_cast<int32_t *>(take((tiles + 1) * 4));
    f->bins.rank = reinterpret_cast<int32_t *>(take(n * 4 * 4));
    f->bins.max_chunks = int32_t(max_chunks(tiles, n));
    f->bins.list = reinterpret_cast<int32_t *>(take(list_entries(tiles, n) * 4));
    f->bins.chunk = reinterpret_cast<int4 *>(take(size_t(f->bins.max_chunks) * 16));
    f->bins.nchunks = reinterpret_cast<int32_t *>(take(4));
    f->bins.items = reinterpret_cast<int32_t *>(take(tiles * 4));
    f->pbins.ntx = (width + kPxTile - 1) / kPxTile;
    f->pbins.nty = (height + kPxTile - 1) / kPxTile;
    const size_t ptiles = size_t(f->pbins.ntx) * size_t(f->pbins.nty);
    f->pbins.count[0] = reinterpret_cast<int32_t *>(take(ptiles * 4));
    f->pbins.count[1] = reinterpret_cast<int32_t *>(take(ptiles * 4));
    f->pbins.start = reinterpret_cast<int32_t *>(take((ptiles + 1) * 4));
    f->pbins.rank = f->bins.rank;  // the splat bins after the last step of a frame
    f->pbins.list = f->bins.list;
    f->pbins.chunk = nullptr;  // one workgroup per pixel tile
    f->pbins.nchunks = nullptr;
    f->pbins.items = nullptr;
    f->pbins.max_chunks = 0;
    f->sxy = reinterpret_cast<uint32_t *>(take(n * 4));
    f->staps = reinterpret_cast<double *>(take(n * 32));
    f->scol = reinterpret_cast<uint32_t *>(take(n * 4));
    f->G.muvw = reinterpret_cast<float *>(take(nn * 16));
    f->G.col = reinterpret_cast<float *>(take(nn * 16));
    float **g[4] = {&f->G.gx, &f->G.gy, &f->G.ax, &f->G.ay};
    for (float **q : g) *q = reinterpret_cast<float *>(take(nn * 4));
    f->mask = reinterpret_cast<int32_t *>(take(npx * 4));
    f->img = reinterpret_cast<uint32_t *>(take(npx * 4));
    f->fp = reinterpret_cast<FrameParams *>(take(sizeof(FrameParams)));
    f->memo = reinterpret_cast<LagSlopeMemo *>(take(n * sizeof(LagSlopeMemo)));
    if (size_t(p - static_cast<char *>(f->buf)) > bytes) {
        set_error("amh_fluid_create: layout of %zu bytes exceeds its allocation of %zu",
                  size_t(p - static_cast<char *>(f->buf)), bytes);
        (void)hipFree(f->buf);
        delete f;
        return AMH_E_ARG;
    }
    // cells ranked in Morton (Z) order of (x, y): a run of consecutive ranks
    // is a compact patch of the grid
    std::vector<int32_t> by_code(nn), rank(nn);
    std::vector<uint32_t> code(nn);
    auto spread = [](uint32_t v) {
        v &= 0xFFFFu;
        v = (v | (v << 8)) & 0x00FF00FFu;
        v = (v | (v << 4)) & 0x0F0F0F0Fu;
        v = (v | (v << 2)) & 0x33333333u;
        v = (v | (v << 1)) & 0x55555555u;
        return v;
    };
    for (size_t c = 0; c < nn; ++c) {
        code[c] = spread(uint32_t(c / size_t(gsy))) | (spread(uint32_t(c % size_t(gsy))) << 1);
        by_code[c] = int32_t(c);
    }
    std::sort(by_code.begin(), by_code.end(), [&](int32_t a, int32_t b) { return code[a] < code[b]; });
    for (size_t r = 0; r < nn; ++r) rank[by_code[r]] = int32_t(r);
    hipStream_t st = static_cast<hipStream_t>(stream);
    int r = check_hip(hipMemsetAsync(f->buf, 0, bytes, st), "amh_fluid_create");
    if (!r) r = check_hip(hipMemcpyAsync(f->cell_rank, rank.data(), nn * 4, hipMemcpyHostToDevice, st),
                          "amh_fluid_create");
    std::vector<int32_t> iota(n);
    for (size_t i = 0; i < n; ++i) iota[i] = int32_t(i);
    if (!r && n) r = check_hip(hipMemcpyAsync(f->P.atom, iota.data(), n * 4, hipMemcpyHostToDevice, st),
                               "amh_fluid_create");
    if (!r) r = check_hip(hipStreamSynchronize(st), "amh_fluid_create");
    if (r) {
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
    (void)hipDeviceSynchronize();  // frames may still be queued on the caller's stream
    for (auto &g : f->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
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
    // the state is now in atom order: slot i holds atom i
    for (size_t i = 0; i < n; ++i) ic[i] = int32_t(i);
    if (int r = check_hip(hipMemcpyAsync(f->P.atom, ic.data(), n * 4, hipMemcpyHostToDevice, s), "set")) return r;
    if (int r = check_hip(hipStreamSynchronize(s), "set")) return r;
    return AMH_OK;
}

extern "C" int amh_fluid_get_particles(void *fluid, double *host, void *stream) {
    amh_fluid_s *f = static_cast<amh_fluid_s *>(fluid);
    if (!f || !host) { set_error("amh_fluid_get_particles: NULL"); return AMH_E_ARG; }
    const size_t n = size_t(f->count);
    std::vector<double> col(n);
    std::vector<int32_t> ic(n);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the state is in slot order: slot i holds atom P.atom[i]
    std::vector<int32_t> atom(n);
    if (int r = check_hip(hipMemcpyAsync(atom.data(), f->P.atom, n * 4, hipMemcpyDeviceToHost, s), "get")) return r;
    if (int r = check_hip(hipStreamSynchronize(s), "get")) return r;
    for (size_t i = 0; i < n; ++i)
        if (atom[i] < 0 || size_t(atom[i]) >= n) {
            set_error("amh_fluid_get_particles: slot %zu holds atom %d of %zu", i, atom[i], n);
            return AMH_E_ARG;
        }
    const double *src[] = {f->P.x, f->P.y, f->P.u, f->P.v, f->P.gx, f->P.gy, f->P.fr, f->P.R, f->P.G, f->P.B,
                           f->P.A, f->P.r, f->P.g, f->P.b, f->P.a, f->P.strength};
    for (int k = 0; k < 16; ++k) {
        if (int r = check_hip(hipMemcpyAsync(col.data(), src[k], n * 8, hipMemcpyDeviceToHost, s), "get")) return r;
        if (int r = check_hip(hipStreamSynchronize(s), "get")) return r;
        for (size_t i = 0; i < n; ++i) host[20 * size_t(atom[i]) + k] = col[i];
    }
    const int32_t *isrc[] = {f->P.active, f->P.mature};
    for (int k = 0; k < 2; ++k) {
        if (int r = check_hip(hipMemcpyAsync(ic.data(), isrc[k], n * 4, hipMemcpyDeviceToHost, s), "get")) return r;
        if (int r = check_hip(hipStreamSynchronize(s), "get")) return r;
        for (size_t i = 0; i < n; ++i) host[20 * size_t(atom[i]) + 16 + k] = ic[i];
    }
    return AMH_OK;
}

namespace {
// Counting sort of the particles' entries into the bins of B (tiles bins,
// counts of parity par; the other parity is cleared for the next use).
// counted: the counts and ranks are in place already.  Whole bins (the
// splat's pixel tiles).  Beyond kLdsTiles tiles the counts are scanned
// first by one workgroup.  tail(i) runs for every particle i after its
// placing (k_bin_place).
template <class Tiles, class Tail = NoTail>
void bin_particles(const Tiles &tiles_of, int n, const Bins &B, int tiles, int par, hipStream_t s,
                   bool counted = false, const Tail &tail = Tail()) {
    const uint32_t gp = uint32_t((n + kT - 1) / kT);
    if (!counted) hipLaunchKernelGGL(k_bin_count<Tiles>, dim3(gp), dim3(kT), 0, s, tiles_of, n, B, par);
    const uint32_t gb = uint32_t((std::max(n, tiles) + kT - 1) / kT);
    if (tiles <= kLdsTiles) {
        hipLaunchKernelGGL((k_bin_place<true, Tiles, Tail>), dim3(gb), dim3(kT), 0, s, tiles_of, n, B, tiles, par,
                           tail);
    } else {
        hipLaunchKernelGGL(scan::k_scan_sums, dim3(1), dim3(1024), 0, s, B.count[par], tiles, B.start);
        hipLaunchKernelGGL((k_bin_place<false, Tiles, Tail>), dim3(gb), dim3(kT), 0, s, tiles_of, n, B, tiles, par,
                           tail);
    }
}

// The fluid steps' bins and work items (counted: the counts and ranks are in
// place already).  Beyond kLdsTiles tiles the tiles' item counts are
// scanned first, by the device scan (scan::exclusive) on the workspace
// sums, offs (scan::blocks(tiles) and one more).
void place_items(const StencilTiles &tiles_of, int n, const Bins &B, int tiles, int par, hipStream_t s,
                 bool counted, int *sums, int *offs) {
    const uint32_t gp = uint32_t((n + kT - 1) / kT);
    if (!counted) hipLaunchKernelGGL(k_bin_count<StencilTiles>, dim3(gp), dim3(kT), 0, s, tiles_of, n, B, par);
    const uint32_t gb = uint32_t((std::max(n, tiles) + kT - 1) / kT);
    if (tiles <= kLdsTiles) {
        hipLaunchKernelGGL(k_item_place<true>, dim3(gb), dim3(kT), 0, s, tiles_of, n, B, tiles, par);
    } else {
        hipLaunchKernelGGL(k_tile_items, dim3(uint32_t((tiles + kT - 1) / kT)), dim3(kT), 0, s, B, tiles, par);
        scan::exclusive(B.items, tiles, sums, offs, B.start, s);
        hipLaunchKernelGGL(k_item_place<false>, dim3(gb), dim3(kT), 0, s, tiles_of, n, B, tiles, par);
    }
}

// One FluidModel::step (fluidmodel.cpp:165-580) on the particles Q, atom[i]
// the atom in slot i (the wall jitter is keyed by atom).  counted: the previous step's
// advection counted this step's bins; count_next: this step's advection
// counts the next step's (when another step follows with unchanged
// particles).
int fluid_step_on(amh_fluid_s *f, const Particles &Q, const int32_t *atom, int32_t steps_left,
                  double freedom_radius, uint64_t seed, hipStream_t s, bool counted = false,
                  bool count_next = false, const FrameParams *fp = nullptr, int step_k = 0,
                  const SplatArgs *splat = nullptr, int splat_par = 0) {
    amh::ProfScope step_("fluid_step", s);
    if (f->count == 0) return AMH_OK;
    const int tiles = f->bins.ntx * f->bins.nty;
    // this step's count parity: the counts alternate from step to step
    Bins B = f->bins;
    const int par = f->bin_par;
    f->bin_par ^= 1;
    B.par = par;
    // about one work item per workgroup (a bin averages under one chunk of
    // kT entries, and empty tiles get none); a workgroup takes further items
    // when there are more
    const uint32_t grid = uint32_t(std::min(B.max_chunks, 2 * std::min(tiles, kLdsTiles)));
    {
        amh::ProfScope ps_("fluid_bins", s);
        place_items(StencilTiles{Q, f->gsx, f->gsy, f->G}, f->count, B, tiles, par, s, counted, f->scan_sums,
                    f->scan_offs);
    }
    amh::launch_timed("k_p2g", k_p2g, dim3(grid), dim3(kT), 0, s, Q, f->G, f->gsx, f->gsy, B);
    amh::launch_timed("k_forces", k_forces, dim3(grid), dim3(kT), 0, s, Q, f->G, f->gsx, f->gsy, B);
    amh::launch_timed("k_g2p", k_g2p, dim3(grid), dim3(kT), 0, s, Q, f->G, f->gsx, f->gsy, B);
    const uint32_t gp = uint32_t((f->count + kT - 1) / kT);
    if (splat)  // the frame's last step: the splat's per-particle half too
        amh::launch_timed("k_advect", k_advect<true>, dim3(gp), dim3(kT), 0, s, Q, f->count, f->G, f->gsx, f->gsy,
                          int(steps_left), freedom_radius, seed, atom, f->pbins, splat_par, fp, step_k, *splat);
    else
        amh::launch_timed("k_advect", k_advect<false>, dim3(gp), dim3(kT), 0, s, Q, f->count, f->G, f->gsx, f->gsy,
                          int(steps_left), freedom_radius, seed, atom, B, count_next ? f->bin_par : -1, fp, step_k,
                          SplatArgs{});
    AMH_LAUNCH_CHECK("amh_fluid_step");
    return AMH_OK;
}

// The counting pass of sort_particles done by another kernel on the way.
SortCount sort_count_of(amh_fluid_s *f) {
    return SortCount{f->cell_rank, f->cell, f->pos, f->cnt, f->gsx, f->gsy};
}

// f->S = f->P in cell order (f->S.atom: the atom in each slot).  The
// bucket counts f->cnt are zero on entry (from the handle's creation, then
// from the last sort's k_sort_slots).  count_first: the first step's bins
// are counted on the way (its parity f->bin_par).  counted: the buckets are
// counted already (k_fluid_update with sort_count_of(f)).
int sort_particles(amh_fluid_s *f, hipStream_t s, bool count_first, bool counted) {
    amh::ProfScope ps_("fluid_sort", s);
    const int64_t nn = int64_t(f->gsx) * f->gsy;
    const uint32_t gp = uint32_t((f->count + kT - 1) / kT);
    const int64_t nk = nn + kIdleBuckets;
    if (!counted)
        hipLaunchKernelGGL(k_sort_count, dim3(gp), dim3(kT), 0, s, f->P, f->count, f->gsx, f->gsy, f->cell_rank,
                           f->cell, f->pos, f->cnt);
    scan::exclusive(f->cnt, nk, f->scan_sums, f->scan_offs, f->start, s);
    hipLaunchKernelGGL(k_sort_slots, dim3(gp), dim3(kT), 0, s, f->count, f->cell, f->pos, f->start, f->P.atom,
                       f->prev, f->S, f->cnt, nk);
    hipLaunchKernelGGL(k_sort_gather, dim3(gp), dim3(kT), 0, s, f->P, f->count, f->prev, f->S, f->gsx, f->gsy,
                       f->bins, count_first ? f->bin_par : -1);
    AMH_LAUNCH_CHECK("fluid sort");
    return AMH_OK;
}
}  // namespace

// One FluidModel::step (fluidmodel.cpp:165-580) on the particles as they are
// stored (in the order of the last cell sort; the wall jitter is keyed by
// atom).
extern "C" int amh_fluid_step(void *fluid, int32_t steps_left, double freedom_radius, uint64_t seed, void *stream) {
    amh_fluid_s *f = static_cast<amh_fluid_s *>(fluid);
    if (!f) { set_error("amh_fluid_step: NULL"); return AMH_E_ARG; }
    return fluid_step_on(f, f->P, f->P.atom, steps_left, freedom_radius, seed, static_cast<hipStream_t>(stream));
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

// Fingerprint of the PERLIN tables and the bounding box, the same in every
// block: each thread adds a mix of each table entry it staged (hsum), the
// block sums them in wsum[kT / 64] (after a barrier, by every thread).
__device__ __forceinline__ uint64_t lag_slope_tag(const amh_render &R, uint64_t hsum, uint64_t *wsum) {
    hsum = wave_sum_u64(hsum);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = hsum;
    __syncthreads();
    uint64_t h = 0;
    for (int w = 0; w < kT / 64; ++w) h += wsum[w];
    const uint64_t box = mix64((uint64_t(uint32_t(R.bbox_x1)) << 32) | uint32_t(R.bbox_y1)) ^
                         mix64(((uint64_t(uint32_t(R.bbox_x2)) << 32) | uint32_t(R.bbox_y2)) + 0x9E3779B97F4A7C15ull);
    return mix64(h ^ box) | 1u;
}

// update_particle for the particle in slot s of P, atom a of chain c
// (morph.cpp:680-1050); returns its sort bucket when sc.cnt is set (its
// cell's rank, or an idle bucket when it is not active).  ka: the atom's
// key points in one record (atom-major copy), else NULL (read from
// R.pts_dev, one row per key).  lag, slope: the atom's PERLIN lag and
// slope (atom_lag_slope) when the fade is PERLIN.
__device__ __forceinline__ int update_one(const amh_render &R, const Frame &F, Particles P,
                                          const int32_t *__restrict__ rank, const int32_t *__restrict__ surf,
                                          const double *__restrict__ bcol, double sp, int s, int64_t a, int c,
                                          const SortCount &sc, double lag, double slope,
                                          const uint64_t *__restrict__ ka) {
    // every load first: the compiler must assume the stores to P below may
    // alias the other arrays, and would wait for each load in turn
    const int nA = surf[int64_t(F.fk) * R.nchains + c], nB = surf[int64_t(F.fnext) * R.nchains + c];
    const int rk = rank[int64_t(F.fk) * R.wtot + a];
    const uint64_t p1 = ka ? ka[F.fk] : R.pts_dev[int64_t(F.fk) * R.wtot + a];
    const uint64_t p2 = ka ? ka[F.fnext] : R.pts_dev[int64_t(F.fnext) * R.wtot + a];
    const int act0 = P.active[s], key0 = P.key[s], mat0 = P.mature[s], own0 = P.owner[s];
    const double r0 = P.r[s], g0 = P.g[s], b0 = P.b[s], a0 = P.a[s];
    const double x0 = sc.cnt ? P.x[s] : 0.0, y0 = sc.cnt ? P.y[s] : 0.0;
    uint64_t tk[4] = {p1, p1, p2, p2};
    if (ka) traj_keys_of(R, F, ka, tk);
    else traj_keys(R, F, a, tk);
    const int limit = int(round((1.0 - F.tl) * double(nA) + F.tl * double(nB)));
    if (rk >= limit) {
        P.active[s] = 0;
        return idle_bucket(s, sc.gsx * sc.gsy);
    }
    // a new particle (Particle::clear, fluidmodel.cpp:80-95)
    const bool fresh = !(act0 && key0 == F.fk);
    const int mature = fresh ? 0 : mat0, owner = fresh ? 1 : own0;
    const bool h1 = pt_flags(p1) & HAS_PIXEL, h2 = pt_flags(p2) & HAS_PIXEL;
    const int W = R.width, H = R.height;
    // both key pixels are read whatever the case (selects, not branches)
    const uint32_t k1 = key_px(R.keys_dev, F.fk, W, H, pt_x(p1), pt_y(p1));
    const uint32_t k2 = key_px(R.keys_dev, F.fnext, W, H, pt_x(p2), pt_y(p2));
    uint32_t c1 = k1, c2 = k2;
    if (h1 && !h2) c2 = nB > 0 ? (k1 & 0x00FFFFFFu) : k1;
    else if (h2 && !h1) c1 = nA > 0 ? (k2 & 0x00FFFFFFu) : k2;
    uint32_t qx, qy, qxf, qyf;
    traj_point(R, F, a, p1, p2, qx, qy, qxf, qyf, tk);
    const uint32_t col = atom_colour_ls(R, F, c1, c2, lag, slope);
    // point2xy works in float (atomorph.h:316-319)
    const float ptx = float(qx) + float(qxf) / float(256), pty = float(qy) + float(qyf) / float(256);
    const double gxv = fmin(fmax((double(ptx) * sp) + 10.0, 1.0), double(W * sp + 10.0));
    const double gyv = fmin(fmax((double(pty) * sp) + 10.0, 1.0), double(H * sp + 10.0));
    double Rr = ch(col, 0) / 255.0, Gg = ch(col, 1) / 255.0, Bb = ch(col, 2) / 255.0, Aa = ch(col, 3) / 255.0;
    if (R.show_blobs == 2 && bcol) {  // DISTINCT
        const double *q = bcol + (int64_t(F.fk) * R.nchains + c) * 4;
        Rr = q[0]; Gg = q[1]; Bb = q[2]; Aa = 1.0;
    } else if (R.show_blobs == 1 && bcol) {  // AVERAGE: the blob's mean colour
        const double *q = bcol + (int64_t(F.fk) * R.nchains + c) * 4;
        Rr = q[0]; Gg = q[1]; Bb = q[2]; Aa = q[3];
    }
    const double t = F.tl;
    double r = t * Rr + (1.0 - t) * r0, g = t * Gg + (1.0 - t) * g0;
    double b = t * Bb + (1.0 - t) * b0, al = t * Aa + (1.0 - t) * a0;
    if (fresh) {
        P.u[s] = 0.0; P.v[s] = 0.0;
        P.mature[s] = 0;
        P.owner[s] = 1;
        P.key[s] = F.fk;
    }
    P.gx[s] = gxv;
    P.gy[s] = gyv;
    P.R[s] = Rr; P.G[s] = Gg; P.B[s] = Bb; P.A[s] = Aa;
    const bool placed = !mature && owner;  // the particle moves to its target
    if (!mature) {  // (a fresh particle's strength 1.0 is replaced here)
        if (owner) {
            P.x[s] = gxv;
            P.y[s] = gyv;
            r = Rr; g = Gg; b = Bb; al = Aa;
        }
        P.strength[s] = nA == 0 ? 0.1 : 1.0;
        P.fr[s] = 1.0;
    }
    P.r[s] = r; P.g[s] = g; P.b[s] = b; P.a[s] = al;
    P.active[s] = 1;
    return sc.cnt ? sc.cell_rank[cell_of(placed ? gxv : x0, placed ? gyv : y0, sc.gsx, sc.gsy)] : 0;
}

// Chains whose first atoms the update stages in LDS (more: searched in
// global memory).
constexpr int kChainLds = 4096;

// The last k with v[k] <= x (v ascending, v[0] <= x).
__device__ __forceinline__ int last_at_or_below(const int32_t *v, int n, int32_t x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (v[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// step_fluid's bookkeeping and update_particle (morph.cpp:680-1050), one
// thread per slot of P (the particles in the last cell order; P.atom the
// atom in each); with sc.cnt set, also the counting pass of the cell sort
// that follows (as k_sort_count).
__global__ __launch_bounds__(kT) void k_fluid_update(amh_render R, const FrameParams *__restrict__ fp, Particles P,
                                                    const int32_t *__restrict__ rank,
                                                    const int32_t *__restrict__ surf,
                                                    const double *__restrict__ bcol, double sp, SortCount sc,
                                                    LagSlopeMemo *__restrict__ memo,
                                                    const uint64_t *__restrict__ atom_keys) {
    const Frame F = fp->F;
    __shared__ int lagp[512], slopep[512];
    __shared__ uint64_t wsum[kT / 64];
    __shared__ int32_t choff[kChainLds];
    uint64_t tag = 0;
    if (R.fading == kPerlin) {  // (uniform)
        uint64_t hsum = 0;
        for (int i = threadIdx.x; i < 512; i += kT) {
            const int l = R.lag_perm_dev[i], q = R.slope_perm_dev[i];
            lagp[i] = l;
            slopep[i] = q;
            hsum += mix64((uint64_t(i) << 32) | uint32_t(l)) + mix64((uint64_t(i + 512) << 32) | uint32_t(q));
        }
        tag = lag_slope_tag(R, hsum, wsum);  // (its barrier also publishes the tables)
    }
    // the chains' first atoms, for the chain of each slot's atom (the slots
    // are in the last cell order, so a block's atoms are not consecutive)
    const bool lds_chains = R.nchains <= kChainLds;
    if (lds_chains)
        for (int k = threadIdx.x; k < R.nchains; k += kT) choff[k] = int32_t(R.chain_off_dev[k]);
    __syncthreads();
    const int s = int(blockIdx.x) * kT + int(threadIdx.x);
    const bool in = s < R.wtot;
    const int64_t a = in ? P.atom[s] : 0;
    const uint64_t *ka = atom_keys ? atom_keys + a * R.nkeys : nullptr;
    const int c = lds_chains ? last_at_or_below(choff, R.nchains, int32_t(a)) : find_chain(R.chain_off_dev, R.nchains, a);
    // PERLIN: the atom's lag and slope from memo[a] when it holds them for
    // this source point and tag, else computed and kept there (on its own,
    // before the update's loads, so few values are live around the noise)
    double lag = 0.5, slope = 0.5;
    if (tag != 0 && in) {
        const uint64_t p1 = ka ? ka[F.fk] : R.pts_dev[int64_t(F.fk) * R.wtot + a];
        const LagSlopeMemo mm = memo[a];
        lag = mm.lag;
        slope = mm.slope;
        if (!(mm.p1 == p1 && mm.tag == tag)) {
            atom_lag_slope(R, p1, lagp, slopep, lag, slope);
            memo[a] = LagSlopeMemo{lag, slope, p1, tag};
        }
    }
    const int key = in ? update_one(R, F, P, rank, surf, bcol, sp, s, a, c, sc, lag, slope, ka) : -1;
    if (sc.cnt) sort_rank(key, in, s, sc.cell, sc.pos, sc.cnt);  // uniform; every lane of the wave
}

// draw_fluid, splat (morph.cpp:1080-1167), in two passes without atomics
// on global memory.  k_splat_prep: per particle the colour update, the
// pixel of its top-left tap and the four bilinear tap weights (0 = not
// drawn); then the particles are binned by 32 x 32 pixel tile; then
// k_splat_tile gathers a tile's taps in LDS (FP64 sums: weight, weight *
// channel) and resolves its pixels (morph.cpp:1169-1209) straight into the
// image.
// The splat of every particle, which also counts the pixel-tile bins of
// its taps into B.count[par] (saving the binning's counting pass).
__global__ __launch_bounds__(kT) void k_splat_prep(Particles P, int n, const FrameParams *__restrict__ fp,
                                                  SplatArgs a, Bins B, int par) {
    const double t = fp->F.tl;
    const int i = blockIdx.x * kT + threadIdx.x;
    const uint32_t v = i < n ? splat_one(P, i, t, a) : kNoTap;
    int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
    if (v != kNoTap) splat_tiles(v, a.W, a.H, tx0, tx1, ty0, ty1);
    count_entries(v != kNoTap, tx0, tx1, ty0, ty1, i, n, B, par);  // every lane of the wave
}

struct SplatItem {
    uint32_t xy;
    uint32_t col;
    double w[4];
};

__global__ __launch_bounds__(kT) void k_splat_tile(Bins B, const uint32_t *__restrict__ xy,
                                                  const double *__restrict__ taps, const uint32_t *__restrict__ cols,
                                                  int W, int H, const FrameParams *__restrict__ fp, int keep_bg,
                                                  uint32_t *__restrict__ img, int32_t *__restrict__ mask) {
    const double t = fp->F.tl;
    __shared__ double sums[5][kPxTile * kPxTile];  // 40 KB
    const int tx = int(blockIdx.x) / B.nty, ty = int(blockIdx.x) % B.nty;
    const int x0 = tx * kPxTile, y0 = ty * kPxTile;
    for (int k = threadIdx.x; k < 5 * kPxTile * kPxTile; k += kT) (&sums[0][0])[k] = 0.0;
    __syncthreads();
    auto fetch = [&](int slot) {
        SplatItem it = {kNoTap, 0u, {0.0, 0.0, 0.0, 0.0}};
        if (slot >= 0) {
            it.xy = xy[slot];
            it.col = cols[slot];
            for (int q = 0; q < 4; ++q) it.w[q] = taps[int64_t(slot) * 4 + q];
        }
        return it;
    };
    for_bin(B, B.start[blockIdx.x], B.start[blockIdx.x + 1], fetch, [&](int slot, bool, const SplatItem &it) {
        if (slot < 0 || it.xy == kNoTap) return;
        const int X = int(it.xy & 0xFFFFu), Y = int(it.xy >> 16);
        for (int q = 0; q < 4; ++q) {
            const double wq = it.w[q];
            const int lx = X + (q & 1) - x0, ly = Y + (q >> 1) - y0;
            if (wq <= 0.0 || lx < 0 || lx >= kPxTile || ly < 0 || ly >= kPxTile) continue;
            const int idx = ly * kPxTile + lx;
            atomicAdd(&sums[0][idx], wq);
            for (int k = 0; k < 4; ++k) atomicAdd(&sums[1 + k][idx], double(ch(it.col, k)) * wq);
        }
    });
    __syncthreads();
    for (int idx = threadIdx.x; idx < kPxTile * kPxTile; idx += kT) {
        const int px = x0 + idx % kPxTile, py = y0 + idx / kPxTile;
        if (px >= W || py >= H) continue;
        const int64_t p = int64_t(py) * W + px;
        const double ws = sums[0][idx];
        if (ws > 0.0) {
            const double r = round(sums[1][idx] / ws), g = round(sums[2][idx] / ws), b = round(sums[3][idx] / ws);
            double a = round(sums[4][idx] / ws);
            if (keep_bg) a *= 1.0 - pow(t, 24.0);
            img[p] = rgba_pack(narrow8(r), narrow8(g), narrow8(b), narrow8(a));
            mask[p] = 1;
        } else {
            mask[p] = 0;
        }
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
__global__ __launch_bounds__(kT) void k_fluid_compose(amh_render R, const FrameParams *__restrict__ fp,
                                                      const uint32_t *img, const int32_t *mask) {
    const Frame F = fp->F;
    uint32_t *out = fp->out;
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

// ------------------------------------------------------ activation ranks
// Which atoms of a chain carry the round((1-t)|A| + t|B|) active particles
// of key segment y -> y+1: slots ordered by class (HAS_FLUID in both rows,
// only in row y, only in row y+1, in neither), then by a pseudo-random
// permutation of the slots keyed by (seed, chain, y).  One block per
// (y, chain).
constexpr int kRankT = 1024;  // the longest chain's block sets the time: its Feistel walks are ALU work
constexpr int kRankWaves = kRankT / 64;

// The classes of up to kRankBatch consecutive permutation places [p, p1)
// (4 past the end): every place's two loads are issued before the first is
// used.
constexpr int kRankBatch = 8;
__device__ __forceinline__ void rank_batch(const uint64_t *__restrict__ pts, int64_t wtot, int y, int yn,
                                           int64_t o, uint32_t w, uint64_t key, int p, int p1,
                                           uint32_t (&x)[kRankBatch], int (&cls)[kRankBatch]) {
    uint64_t q1[kRankBatch], q2[kRankBatch];
    for (int j = 0; j < kRankBatch; ++j) {
        x[j] = p + j < p1 ? feistel_perm(uint32_t(p + j), w, key) : 0u;
        q1[j] = pts[int64_t(y) * wtot + o + x[j]];
        q2[j] = pts[int64_t(yn) * wtot + o + x[j]];
    }
    for (int j = 0; j < kRankBatch; ++j) {
        const bool f1 = pt_flags(q1[j]) & HAS_FLUID, f2 = pt_flags(q2[j]) & HAS_FLUID;
        cls[j] = p + j >= p1 ? 4 : f1 && f2 ? 0 : f1 ? 1 : f2 ? 2 : 3;
    }
}

// Thread t takes the run of permutation places [t J, (t + 1) J), J =
// ceil(w / kRankT): it counts its run's classes (loads in batches), a block
// scan gives each class's first rank in the run, and a second walk hands
// the ranks out in place order.
__global__ __launch_bounds__(kRankT) void k_fluid_ranks(const uint64_t *__restrict__ pts, int h, int64_t wtot,
                                                        const int64_t *__restrict__ off,
                                                        const int32_t *__restrict__ cw, int nchains, uint64_t seed,
                                                        int32_t *__restrict__ rank) {
    const int c = int(blockIdx.x % uint32_t(nchains)), y = int(blockIdx.x / uint32_t(nchains));
    const int yn = (y + 1) % h, w = cw[c];
    const int64_t o = off[c];
    if (w <= 0) return;
    const uint64_t key = mix64(seed ^ (uint64_t(uint32_t(c)) << 32) ^ uint64_t(uint32_t(y)));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ int wtot4[kRankWaves][4];
    const int J = (w + kRankT - 1) / kRankT;
    const int p0 = min(w, int(threadIdx.x) * J), p1 = min(w, p0 + J);
    uint32_t x[kRankBatch];
    int cls[kRankBatch];
    int cnt[4] = {0, 0, 0, 0};
    for (int p = p0; p < p1; p += kRankBatch) {
        rank_batch(pts, wtot, y, yn, o, uint32_t(w), key, p, p1, x, cls);
        for (int j = 0; j < kRankBatch; ++j)
            for (int k = 0; k < 4; ++k) cnt[k] += cls[j] == k;
    }
    // exclusive scan of the four counts over the threads, in thread order
    int incl[4];
    for (int k = 0; k < 4; ++k) {
        incl[k] = cnt[k];
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl[k], d, 64);
            if (lane >= d) incl[k] += v;
        }
        if (lane == 63) wtot4[wave][k] = incl[k];
    }
    __syncthreads();
    int next[4];
    int classbase = 0;
    for (int k = 0; k < 4; ++k) {
        int before = 0, total = 0;
        for (int v = 0; v < kRankWaves; ++v) {
            before += v < wave ? wtot4[v][k] : 0;
            total += wtot4[v][k];
        }
        next[k] = classbase + before + incl[k] - cnt[k];
        classbase += total;
    }
    for (int p = p0; p < p1; p += kRankBatch) {
        rank_batch(pts, wtot, y, yn, o, uint32_t(w), key, p, p1, x, cls);
        for (int j = 0; j < kRankBatch; ++j)
            for (int k = 0; k < 4; ++k)
                if (cls[j] == k) rank[int64_t(y) * wtot + o + x[j]] = next[k]++;
    }
}

}  // namespace

namespace {
__global__ __launch_bounds__(kT) void k_atom_keys(const uint64_t *__restrict__ pts, int h, int64_t wtot,
                                                 uint64_t *__restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (i >= wtot * h) return;
    const int64_t a = i / h;
    const int y = int(i - a * h);
    out[i] = pts[int64_t(y) * wtot + a];
}
}  // namespace

extern "C" int amh_fluid_atom_keys(const uint64_t *pts_dev, int32_t nkeys, int64_t wtot, uint64_t *out_dev,
                                   void *stream) {
    if (!pts_dev || !out_dev || nkeys <= 0 || wtot < 0 || wtot > INT32_MAX) {
        set_error("amh_fluid_atom_keys: invalid argument");
        return AMH_E_ARG;
    }
    const int64_t total = wtot * nkeys;
    if (total == 0) return AMH_OK;
    hipLaunchKernelGGL(k_atom_keys, dim3(uint32_t((total + kT - 1) / kT)), dim3(kT), 0,
                       static_cast<hipStream_t>(stream), pts_dev, int(nkeys), wtot, out_dev);
    AMH_LAUNCH_CHECK("amh_fluid_atom_keys");
    return AMH_OK;
}

extern "C" int amh_fluid_ranks(const uint64_t *pts_dev, int32_t nkeys, int64_t wtot, const int64_t *chain_off_dev,
                               const int32_t *chain_w_dev, int32_t nchains, uint64_t seed, int32_t *rank_dev,
                               void *stream) {
    if (!pts_dev || !chain_off_dev || !chain_w_dev || !rank_dev || nkeys <= 0 || nchains <= 0 || wtot < 0 ||
        int64_t(nkeys) * nchains > INT32_MAX) {
        set_error("amh_fluid_ranks: invalid argument");
        return AMH_E_ARG;
    }
    if (wtot == 0) return AMH_OK;
    hipLaunchKernelGGL(k_fluid_ranks, dim3(uint32_t(nkeys) * uint32_t(nchains)), dim3(kRankT), 0,
                       static_cast<hipStream_t>(stream), pts_dev, int(nkeys), wtot, chain_off_dev, chain_w_dev,
                       int(nchains), seed, rank_dev);
    AMH_LAUNCH_CHECK("amh_fluid_ranks");
    return AMH_OK;
}

namespace {
// draw_fluid after step_fluid (morph.cpp:1082-1209) on the cell-ordered
// copy f->S: colour relax and splat, resolve into f->img / f->mask, then the
// particles (colours relaxed, mature set) back to f->P.  The frame's time
// comes from f->fp.
SplatArgs splat_args(const amh_fluid_s *f, const amh_render &R, double sparseness) {
    return SplatArgs{f->W, f->H, sparseness, R.bbox_x1, R.bbox_y1, R.bbox_x2, R.bbox_y2, f->sxy, f->staps, f->scol};
}

// prepped: the per-particle half (k_splat_prep) ran already, in the last
// step's advection, counting into f->pbins.count[f->pbin_par].
void draw_splat(amh_fluid_s *f, const amh_render &R, double sparseness, hipStream_t s, bool prepped) {
    amh::ProfScope ps_("k_fluid_splat", s);
    const uint32_t gp = uint32_t((f->count + kT - 1) / kT);
    const int ptiles = f->pbins.ntx * f->pbins.nty;
    const int par = f->pbin_par;
    f->pbin_par ^= 1;
    if (!prepped)
        hipLaunchKernelGGL(k_splat_prep, dim3(gp), dim3(kT), 0, s, f->S, f->count, f->fp,
                           splat_args(f, R, sparseness), f->pbins, par);
    bin_particles(SplatTiles{f->sxy, f->W, f->H}, f->count, f->pbins, ptiles, par, s, true);
    hipLaunchKernelGGL(k_splat_tile, dim3(uint32_t(ptiles)), dim3(kT), 0, s, f->pbins, f->sxy, f->staps, f->scol,
                       f->W, f->H, f->fp, R.keep_background, f->img, f->mask);
}

// The outward feather layers and the background blend (morph.cpp:1211-1298)
// into the frame's image f->fp->out.
void draw_finish(amh_fluid_s *f, const amh_render &R, hipStream_t s) {
    const int64_t npx = int64_t(f->W) * f->H;
    const uint32_t gpx = uint32_t((npx + kT - 1) / kT);
    for (int layer = 0; layer < R.feather; ++layer) {
        const double alpha = 1.0 - (double(layer + 1) / double(R.feather + 1));
        hipLaunchKernelGGL(k_fluid_feather, dim3(gpx), dim3(kT), 0, s, f->W, f->H, alpha, f->img, f->mask);
        hipLaunchKernelGGL(k_fluid_commit, dim3(gpx), dim3(kT), 0, s, f->mask, npx);
    }
    hipLaunchKernelGGL(k_fluid_compose, dim3(gpx), dim3(kT), 0, s, R, f->fp, f->img, f->mask);
}

// Once a frame's launches on f->S are enqueued, S holds the particle state.
void swap_state(amh_fluid_s *f) {
    std::swap(f->P, f->S);
    f->swapped ^= 1;
}

// One frame's launches after its values are in f->fp: step_fluid's
// bookkeeping, the model steps, the drawing.  `drawn`: the frame has a key
// segment and particles (otherwise every pixel is empty).
int enqueue_frame(amh_fluid_s *f, const amh_fluid_frame *fa, bool drawn, hipStream_t s) {
    const amh_render &R = *fa->r;
    if (drawn) {
        const uint32_t gp = uint32_t((f->count + kT - 1) / kT);
        // the update counts the cell sort's buckets on the way
        amh::launch_timed("k_fluid_update", k_fluid_update, dim3(gp), dim3(kT), 0, s, R, f->fp, f->P, fa->rank_dev,
                          fa->surface_dev, fa->blob_col_dev, fa->sparseness, sort_count_of(f), f->memo,
                          fa->atom_keys_dev);
        // the steps and the splat run on the cell-ordered copy
        // the sort counts the first step's bins; each step's advection the next's
        if (int r = sort_particles(f, s, fa->fluidsteps > 0, true)) return r;
        const SplatArgs sa = splat_args(f, R, fa->sparseness);
        for (int k = 0; k < fa->fluidsteps; ++k) {
            const bool last = k + 1 == fa->fluidsteps;
            const int r = fluid_step_on(f, f->S, f->S.atom, fa->fluidsteps - (k + 1), 0.0, 0, s, true, !last, f->fp,
                                        k, last ? &sa : nullptr, f->pbin_par);
            if (r) return r;
        }
        draw_splat(f, R, fa->sparseness, s, fa->fluidsteps > 0);
        swap_state(f);
    } else {
        const int64_t npx = int64_t(f->W) * f->H;
        hipLaunchKernelGGL(k_fill_i32, dim3(uint32_t((npx + kT - 1) / kT)), dim3(kT), 0, s, f->mask, npx);
    }
    draw_finish(f, R, s);
    AMH_LAUNCH_CHECK("fluid frame");
    return AMH_OK;
}

// Everything the launches of enqueue_frame bake into a captured graph.
std::vector<unsigned char> frame_key(const amh_fluid_s *f, const amh_fluid_frame *fa, bool drawn) {
    amh_fluid_frame fc = *fa;
    fc.r = nullptr;
    const int32_t state[5] = {drawn ? 1 : 0, f->pbin_par, f->bin_par, f->count, f->swapped};
    std::vector<unsigned char> k(sizeof(amh_render) + sizeof(fc) + sizeof(state));
    std::memcpy(k.data(), fa->r, sizeof(amh_render));
    std::memcpy(k.data() + sizeof(amh_render), &fc, sizeof(fc));
    std::memcpy(k.data() + sizeof(amh_render) + sizeof(fc), state, sizeof(state));
    return k;
}

// The frame from its captured graph (captured on first use for its key);
// false when the stream cannot be captured (then nothing was enqueued).
bool replay_frame(amh_fluid_s *f, const amh_fluid_frame *fa, bool drawn, hipStream_t s, int *rc) {
    *rc = AMH_OK;
    if (!f->use_graphs || !s || amh::profiling()) return false;
    const std::vector<unsigned char> key = frame_key(f, fa, drawn);
    const int32_t pbin_after = f->pbin_par ^ (drawn ? 1 : 0);
    const int32_t bin_after = f->bin_par ^ (drawn ? (fa->fluidsteps & 1) : 0);
    amh_fluid_s::Captured *hit = nullptr, *victim = &f->graphs[0];
    for (auto &g : f->graphs) {
        if (g.exec && g.key == key) hit = &g;
        if (!g.exec || g.used < victim->used) victim = &g;
    }
    if (!hit) {
        hipStreamCaptureStatus st;
        if (hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return false;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        const int32_t p0 = f->pbin_par, b0 = f->bin_par, w0 = f->swapped;
        const Particles P0 = f->P, S0 = f->S;
        const int r = enqueue_frame(f, fa, drawn, s);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &graph);
        f->pbin_par = p0;  // the replay below advances them
        f->bin_par = b0;
        f->P = P0;
        f->S = S0;
        f->swapped = w0;
        hipGraphExec_t exec = nullptr;
        if (r || e != hipSuccess || !graph || hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            f->use_graphs = false;  // run this handle's frames eagerly from now on
            return false;
        }
        (void)hipGraphDestroy(graph);
        if (victim->exec) (void)hipGraphExecDestroy(victim->exec);
        victim->exec = exec;
        victim->key = key;
        hit = victim;
    }
    hit->used = ++f->graph_clock;
    *rc = check_hip(hipGraphLaunch(hit->exec, s), "amh_fluid_render graph");
    f->pbin_par = pbin_after;
    f->bin_par = bin_after;
    if (drawn) swap_state(f);
    return true;
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
    amh::ProfScope whole_("fluid_render", s);
    FrameParams v;
    v.F = frame_of(t, R.nkeys);
    // morph.cpp:1050-1054
    const double freedom = sqrt(double(f->W * f->W + f->H * f->H)) * fa->sparseness / 4.0;
    v.radius = freedom * (1.0 - v.F.tl);
    v.frame = f->frame;
    v.out = out;
    const bool drawn = v.F.fk < R.nkeys && f->count > 0;
    hipLaunchKernelGGL(k_set_params, dim3(1), dim3(1), 0, s, f->fp, v);
    int r = AMH_OK;
    if (!replay_frame(f, fa, drawn, s, &r)) r = enqueue_frame(f, fa, drawn, s);
    if (r) return r;
    ++f->frame;
    AMH_LAUNCH_CHECK("amh_fluid_render");
    return AMH_OK;
}

extern "C" int amh_fluid_use_graphs(void *fluid, int32_t on) {
    amh_fluid_s *f = static_cast<amh_fluid_s *>(fluid);
    if (!f) {
        set_error("amh_fluid_use_graphs: NULL");
        return AMH_E_ARG;
    }
    f->use_graphs = on != 0;
    return AMH_OK;
}

// The drawing half on its own, on the particles as they are.
extern "C" int amh_fluid_draw(void *fluid, const amh_fluid_frame *fa, double t, uint32_t *out, void *stream) {
    amh_fluid_s *f = static_cast<amh_fluid_s *>(fluid);
    if (!f || !fa || !fa->r || !out || !(fa->sparseness > 0.0)) {
        set_error("amh_fluid_draw: invalid argument");
        return AMH_E_ARG;
    }
    const amh_render &R = *fa->r;
    if (R.width != f->W || R.height != f->H || R.nkeys <= 0 || !R.keys_dev ||
        (R.keep_background && R.fading == kPerlin && (!R.lag_perm_dev || !R.slope_perm_dev))) {
        set_error("amh_fluid_draw: render arguments do not match the model (%dx%d)", f->W, f->H);
        return AMH_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    FrameParams v;
    v.F = frame_of(t, R.nkeys);
    v.radius = 0.0;
    v.frame = f->frame;
    v.out = out;
    hipLaunchKernelGGL(k_set_params, dim3(1), dim3(1), 0, s, f->fp, v);
    if (v.F.fk < R.nkeys && f->count > 0) {
        if (int r = sort_particles(f, s, false, false)) return r;  // no step follows: nothing to count
        draw_splat(f, R, fa->sparseness, s, false);
        swap_state(f);
    } else {
        const int64_t npx = int64_t(f->W) * f->H;
        hipLaunchKernelGGL(k_fill_i32, dim3(uint32_t((npx + kT - 1) / kT)), dim3(kT), 0, s, f->mask, npx);
    }
    draw_finish(f, R, s);
    AMH_LAUNCH_CHECK("amh_fluid_draw");
    return AMH_OK;
}
