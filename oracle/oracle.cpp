/*
 * CPU oracle for the atomorph morph-engine hot path -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, as the checker and as the CPU timing baseline.  The product
 * (atomorph_amd/) never links or calls it.
 *
 * Every function restates the reference algorithm; citations are
 * file:line in the atomorph source tree.  Floating-point expressions keep the
 * reference's operation order wherever a rounding boundary depends on it
 * (HSP, Perlin, spline, colour fades, composite), so results can be compared
 * bit-for-bit with the compiled reference.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <random>
#include <set>
#include <sstream>
#include <vector>

namespace {

/* ------------------------------------------------------------------ points */
inline uint32_t px_(uint64_t w) { return uint32_t(w & 0xFFFFu); }
inline uint32_t py_(uint64_t w) { return uint32_t((w >> 16) & 0xFFFFu); }
inline uint32_t pxf_(uint64_t w) { return uint32_t((w >> 32) & 0xFFu); }
inline uint32_t pyf_(uint64_t w) { return uint32_t((w >> 40) & 0xFFu); }
inline uint32_t pfl_(uint64_t w) { return uint32_t((w >> 48) & 0xFFu); }
inline int64_t fixx_(uint64_t w) { return int64_t(px_(w)) * 256 + pxf_(w); }
inline int64_t fixy_(uint64_t w) { return int64_t(py_(w)) * 256 + pyf_(w); }
inline uint64_t mkpt_(uint32_t x, uint32_t y, uint32_t xf, uint32_t yf, uint32_t fl) {
    return uint64_t(x & 0xFFFF) | (uint64_t(y & 0xFFFF) << 16) |
           (uint64_t(xf & 0xFF) << 32) | (uint64_t(yf & 0xFF) << 40) |
           (uint64_t(fl & 0xFF) << 48);
}
const unsigned HAS_PIXEL = 1;

/* atomorph.h:334-339: squared distance in 1/256-pixel fixed point. */
inline uint64_t dist_(uint64_t a, uint64_t b) {
    int64_t dx = fixx_(a) - fixx_(b);
    int64_t dy = fixy_(a) - fixy_(b);
    return uint64_t(dx * dx) + uint64_t(dy * dy);
}

/* x86-64 conversion of a double to a narrow unsigned field, as g++ emits it
 * for the reference's implicit conversions (cvttsd2si to int32, keep the low
 * bits).  Used where the reference assigns a double to uint16_t / uint8_t. */
inline uint32_t narrow16_(double v) { return uint32_t(int32_t(v)) & 0xFFFFu; }
inline uint32_t narrow8_(double v) { return uint32_t(int32_t(v)) & 0xFFu; }

/* ------------------------------------------------------ hash + permutation */
inline uint64_t mix64_(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline uint32_t ceil_log2_(uint32_t n) {
    uint32_t k = 0;
    while ((1u << k) < n) ++k;
    return k;
}

/* Balanced Feistel permutation over [0, 2^k) (k even), cycle-walked to [0,n). */
uint32_t feistel_(uint32_t i, uint32_t n, uint64_t key) {
    uint32_t k = ceil_log2_(n < 2 ? 2 : n);
    if (k & 1) ++k;
    uint32_t half = k / 2, mask = (1u << half) - 1u;
    uint32_t v = i;
    do {
        uint32_t L = v >> half, R = v & mask;
        for (uint32_t r = 0; r < 4; ++r) {
            uint32_t f = uint32_t(mix64_(key ^ (uint64_t(R) << 8) ^ r)) & mask;
            uint32_t nl = R;
            R = L ^ f;
            L = nl;
        }
        v = (L << half) | R;
    } while (v >= n);
    return v;
}

/* The GPU schedule's round: the fold matching {i, c - i mod n}, c drawn from
 * the round's hash (atomorph_amd/csrc/common.h, round_params).  Enumerated
 * here independently of the kernel's pair numbering: slot i meets
 * j = (c - i) mod n when i < j. */
inline uint32_t fold_centre_(uint64_t seed, uint64_t stream, uint64_t round, uint32_t n) {
    uint64_t h = mix64_(seed ^ mix64_(stream ^ (round * 0xD1B54A32D192ED03ull)));
    return uint32_t(((h >> 32) * uint64_t(n)) >> 32);
}
template <class F>
inline void fold_pairs_(uint32_t c, uint32_t n, F &&f) {
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t j = (c + n - i) % n;
        if (i < j) f(i, j);
    }
}

/* ---------------------------------------------------------------- colour */
const double kPr = 0.299, kPg = 0.587, kPb = 0.114;

/* color.cpp:71-109 (Finley's RGB->HSP); same operation order. */
void rgb2hsp_(double R, double G, double B, double *H, double *S, double *P) {
    *P = std::sqrt(R * R * kPr + G * G * kPg + B * B * kPb);
    if (R == G && R == B) { *H = 0.0; *S = 0.0; return; }
    /* Largest channel L, with the two others taken in cyclic order after L:
     * `nx` follows L, `pv` precedes it. */
    double L, nx, pv, base_lo, base_hi;
    if (R >= G && R >= B)      { L = R; nx = G; pv = B; base_hi = 6. / 6.; base_lo = 0. / 6.; }
    else if (G >= R && G >= B) { L = G; nx = B; pv = R; base_hi = 2. / 6.; base_lo = 2. / 6.; }
    else                       { L = B; nx = R; pv = G; base_hi = 4. / 6.; base_lo = 4. / 6.; }
    if (pv >= nx) { *H = base_hi - 1. / 6. * (pv - nx) / (L - nx); *S = 1. - nx / L; }
    else          { *H = base_lo + 1. / 6. * (nx - pv) / (L - pv); *S = 1. - pv / L; }
}

/* color.cpp:124-178 (Finley's HSP->RGB); six hue sectors, roles of the
 * channels per sector: largest, middle, smallest. */
void hsp2rgb_(double H, double S, double P, double *R, double *G, double *B) {
    static const int role[6][3] = {{0, 1, 2}, {1, 0, 2}, {1, 2, 0},
                                   {2, 1, 0}, {2, 0, 1}, {0, 2, 1}};
    static const double wt[3] = {kPr, kPg, kPb};
    int sec;
    if (H < 1. / 6.) sec = 0; else if (H < 2. / 6.) sec = 1; else if (H < 3. / 6.) sec = 2;
    else if (H < 4. / 6.) sec = 3; else if (H < 5. / 6.) sec = 4; else sec = 5;
    /* Even sectors measure hue forward from k/6, odd ones backward from (k+1)/6. */
    double h6 = (sec & 1) ? 6. * (-H + double(sec + 1) / 6.) : 6. * (H - double(sec) / 6.);
    int cl = role[sec][0], cm = role[sec][1], cs = role[sec][2];
    double out[3];
    double mom = 1. - S;
    if (mom > 0.) {
        double part = 1. + h6 * (1. / mom - 1.);
        double sm = P / std::sqrt(wt[cl] / mom / mom + wt[cm] * part * part + wt[cs]);
        double lg = sm / mom;
        out[cs] = sm;
        out[cl] = lg;
        out[cm] = sm + h6 * (lg - sm);
    } else {
        double lg = std::sqrt(P * P / (wt[cl] + wt[cm] * h6 * h6));
        out[cl] = lg;
        out[cm] = lg * h6;
        out[cs] = 0.;
    }
    *R = out[0]; *G = out[1]; *B = out[2];
}

/* ---------------------------------------------------------------- perlin */
inline double fade_(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
inline double lerp_(double t, double a, double b) { return a + t * (b - a); }
inline double grad_(int hash, double x, double y, double z) {
    const int h = hash & 15;
    const double u = h < 8 ? x : y, v = h < 4 ? y : h == 12 || h == 14 ? x : z;
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}
/* perlin.cpp:19-43 */
double noise3_(const int32_t *p, double x, double y, double z) {
    const int X = static_cast<int>(std::floor(x)) & 255;
    const int Y = static_cast<int>(std::floor(y)) & 255;
    const int Z = static_cast<int>(std::floor(z)) & 255;
    x -= std::floor(x);
    y -= std::floor(y);
    z -= std::floor(z);
    const double u = fade_(x), v = fade_(y), w = fade_(z);
    const int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
    const int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
    double lo = lerp_(v, lerp_(u, grad_(p[AA], x, y, z), grad_(p[BA], x - 1, y, z)),
                      lerp_(u, grad_(p[AB], x, y - 1, z), grad_(p[BB], x - 1, y - 1, z)));
    double hi = lerp_(v, lerp_(u, grad_(p[AA + 1], x, y, z - 1), grad_(p[BA + 1], x - 1, y, z - 1)),
                      lerp_(u, grad_(p[AB + 1], x, y - 1, z - 1), grad_(p[BB + 1], x - 1, y - 1, z - 1)));
    return lerp_(w, lo, hi);
}
/* perlin.cpp:57-69 */
double octave2_(const int32_t *p, double x, double y, int octaves) {
    double result = 0.0, amp = 1.0;
    for (int i = 0; i < octaves; ++i) {
        result += noise3_(p, x, y, 0.0) * amp;
        x *= 2.0;
        y *= 2.0;
        amp *= 0.5;
    }
    return result;
}

/* --------------------------------------------------------------- colours */
struct rgba_ { uint32_t r, g, b, a; };
inline rgba_ unpack_(uint32_t c) { return {c & 255u, (c >> 8) & 255u, (c >> 16) & 255u, c >> 24}; }
inline uint32_t pack_(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    return (r & 255u) | ((g & 255u) << 8) | ((b & 255u) << 16) | ((a & 255u) << 24);
}
/* morph.cpp:1481-1488 */
rgba_ mix_(rgba_ c1, rgba_ c2, double w1) {
    rgba_ c;
    c.r = narrow8_(std::round(w1 * double(c1.r) + (1.0 - w1) * double(c2.r)));
    c.g = narrow8_(std::round(w1 * double(c1.g) + (1.0 - w1) * double(c2.g)));
    c.b = narrow8_(std::round(w1 * double(c1.b) + (1.0 - w1) * double(c2.b)));
    c.a = narrow8_(std::round(w1 * double(c1.a) + (1.0 - w1) * double(c2.a)));
    return c;
}
/* morph.cpp:1467-1479; the reference applies the cosine ramp when the
 * engine's fading is COSINE or PERLIN. */
rgba_ fade_mix_(rgba_ c1, rgba_ c2, double lag, double slope, double str, int fading) {
    if (fading == 5 || fading == 6) {
        const double pi = 3.14159265358;
        double s = (slope + 0.1) / 1.1;
        double l = (1.0 - s) * lag;
        if (str <= l) str = 0.0;
        else if (str >= (l + s)) str = 1.0;
        else str = ((-std::cos((str - l) * (pi / s)) + 1.0) / 2.0);
    }
    return mix_(c1, c2, str);
}

}  // namespace

/* ======================================================================= */
extern "C" {

uint64_t or_hash64(uint64_t x) { return mix64_(x); }
uint32_t or_perm_index(uint32_t i, uint32_t n, uint64_t key) { return feistel_(i, n, key); }

uint64_t or_point_distance(uint64_t p1, uint64_t p2) { return dist_(p1, p2); }

/* thread.cpp:1109-1125, summed exactly in 128 bits. */
uint64_t or_chain_cost(const uint64_t *pts, size_t w, size_t h, uint64_t *hi) {
    unsigned __int128 e = 0;
    if (w > 1 && h > 0) {
        for (size_t x = 0; x < w; ++x)
            for (size_t j = 0; j < h; ++j)
                e += dist_(pts[x * h + j], pts[x * h + (j + 1) % h]);
    }
    if (hi) *hi = uint64_t(e >> 64);
    return uint64_t(e);
}

/* thread.cpp:989-1041 with thread_count == 0: the mutex and the word re-check
 * are no-ops in a single thread. */
uint64_t or_morph_asynch(uint64_t *pts, size_t w, size_t h, size_t repeat,
                         unsigned rng_seed, double *gain) {
    std::mt19937 gen(rng_seed);
    std::uniform_int_distribution<size_t> ux(0, w - 1);
    std::uniform_int_distribution<size_t> uy(0, h - 1);
    size_t y = uy(gen);
    size_t yn = (y + 1) % h;
    size_t yp = (y > 0 ? y - 1 : h - 1);
    uint64_t accepted = 0;
    for (size_t i = 0; i < repeat; ++i) {
        size_t x1 = ux(gen), x2;
        do { x2 = ux(gen); } while (x1 == x2);
        uint64_t a_p = pts[x1 * h + yp], b_p = pts[x2 * h + yp];
        uint64_t a_c = pts[x1 * h + y], b_c = pts[x2 * h + y];
        uint64_t a_n = pts[x1 * h + yn], b_n = pts[x2 * h + yn];
        double c1 = double(dist_(a_p, a_c) + dist_(a_c, a_n)) + double(dist_(b_p, b_c) + dist_(b_c, b_n));
        double c2 = double(dist_(b_p, a_c) + dist_(a_c, b_n)) + double(dist_(a_p, b_c) + dist_(b_c, a_n));
        if (c1 >= c2) {
            pts[x1 * h + y] = b_c;
            pts[x2 * h + y] = a_c;
            *gain += c1 - c2;
            ++accepted;
        }
    }
    return accepted;
}

/* thread.cpp:1043-1068 (threads == 0): one chain visit. */
uint64_t or_morph_visit(uint64_t *pts, size_t w, size_t h, size_t cycle_length,
                        uint64_t *e1_state, double *gain) {
    if (w <= 1) return 0;
    std::minstd_rand0 e1;
    e1.seed(static_cast<std::minstd_rand0::result_type>(*e1_state));
    std::uniform_int_distribution<unsigned> useed(0, std::numeric_limits<unsigned>::max());
    unsigned s = useed(e1);
    std::stringstream ss;
    ss << e1;
    unsigned long long st = 0;
    ss >> st;
    *e1_state = st;
    return or_morph_asynch(pts, w, h, cycle_length, s, gain);
}

/* The kernels' schedule: rounds of a fold matching over the w slots of row y
 * (atomorph_amd/csrc/refine.hip).  Each pair is judged with the reference's
 * own four distances (thread.cpp:1014-1022), exactly, in int64. */
uint64_t or_refine_rounds(uint64_t *pts, size_t w, size_t h, size_t y,
                          uint64_t rounds, uint64_t seed, uint64_t stream,
                          int64_t *gain_out, uint64_t *attempts_out) {
    uint64_t acc = 0, att = 0;
    int64_t gain = 0;
    if (w >= 2 && h >= 2) {
        size_t yn = (y + 1) % h, yp = (y > 0 ? y - 1 : h - 1);
        for (uint64_t r = 0; r < rounds; ++r) {
            uint32_t c = fold_centre_(seed, stream, r, uint32_t(w));
            fold_pairs_(c, uint32_t(w), [&](uint32_t i, uint32_t j) {
                uint64_t a_p = pts[i * h + yp], b_p = pts[j * h + yp];
                uint64_t a_c = pts[i * h + y], b_c = pts[j * h + y];
                uint64_t a_n = pts[i * h + yn], b_n = pts[j * h + yn];
                int64_t c1 = int64_t(dist_(a_p, a_c) + dist_(a_c, a_n) + dist_(b_p, b_c) + dist_(b_c, b_n));
                int64_t c2 = int64_t(dist_(b_p, a_c) + dist_(a_c, b_n) + dist_(a_p, b_c) + dist_(b_c, a_n));
                ++att;
                if (c1 >= c2) {
                    pts[i * h + y] = b_c;
                    pts[j * h + y] = a_c;
                    gain += c1 - c2;
                    ++acc;
                }
            });
        }
    }
    if (gain_out) *gain_out = gain;
    if (attempts_out) *attempts_out = att;
    return acc;
}

/* Batched pass: same grouping, row choice and stream ids as
 * atomorph_amd/csrc/refine.hip (k_refine_pass). */
uint64_t or_refine_pass(uint64_t *pts, const int64_t *off, const int32_t *widths, int32_t nchains,
                        int32_t h, int32_t cap, uint64_t pass, int32_t rounds, uint64_t seed,
                        int64_t *gain_out, uint64_t *attempts_out) {
    auto pass_stream = [](uint64_t c, uint64_t p, uint64_t g) { return mix64_((c << 32) ^ p) ^ g; };
    uint64_t acc = 0, att = 0;
    int64_t gain = 0;
    if (h >= 2)
        for (int32_t c = 0; c < nchains; ++c) {
            const int32_t w = widths[c];
            if (w < 2) continue;
            uint64_t *cp = pts + size_t(off[c]) * h;
            const uint32_t y = uint32_t(mix64_(seed ^ 0x5851F42D4C957F2Dull ^ mix64_((uint64_t(c) << 32) ^ pass)) % uint32_t(h));
            const uint32_t yp = y > 0 ? y - 1 : h - 1, yn = (y + 1) % h;
            const int ng = (w + cap - 1) / cap, gs = (w + ng - 1) / ng;
            const uint64_t pkey = pass_stream(uint64_t(c), pass, 0xFFFFFFFFull);
            for (int g = 0; g < ng; ++g) {
                const int g0 = g * gs, n = std::min(gs, w - g0);
                std::vector<uint32_t> slot(n);
                std::vector<uint64_t> C(n), Pv(n), Nx(n);
                for (int l = 0; l < n; ++l) {
                    slot[l] = ng > 1 ? feistel_(uint32_t(g0 + l), uint32_t(w), pkey) : uint32_t(l);
                    Pv[l] = cp[size_t(slot[l]) * h + yp];
                    Nx[l] = cp[size_t(slot[l]) * h + yn];
                    C[l] = cp[size_t(slot[l]) * h + y];
                }
                if (n >= 2) {
                    const uint64_t sid = pass_stream(uint64_t(c), pass, uint64_t(g));
                    for (int r = 0; r < rounds; ++r) {
                        uint32_t ctr = fold_centre_(seed, sid, uint64_t(r), uint32_t(n));
                        fold_pairs_(ctr, uint32_t(n), [&](uint32_t i, uint32_t j) {
                            int64_t c1 = int64_t(dist_(Pv[i], C[i]) + dist_(C[i], Nx[i]) +
                                                 dist_(Pv[j], C[j]) + dist_(C[j], Nx[j]));
                            int64_t c2 = int64_t(dist_(Pv[j], C[i]) + dist_(C[i], Nx[j]) +
                                                 dist_(Pv[i], C[j]) + dist_(C[j], Nx[i]));
                            ++att;
                            if (c1 >= c2) {
                                std::swap(C[i], C[j]);
                                gain += c1 - c2;
                                ++acc;
                            }
                        });
                    }
                }
                for (int l = 0; l < n; ++l) cp[size_t(slot[l]) * h + y] = C[l];
            }
        }
    if (gain_out) *gain_out = gain;
    if (attempts_out) *attempts_out = att;
    return acc;
}

/* kNN-guided batched pass (amh_refine_knn) replayed on the CPU.  The
 * proposal rule is the reference's OpenCV build of thread::morph
 * (thread.cpp:893-986): for slot x1, the two row-(y+1) points nearest to
 * x1's point in row y, the two nearest to x1's own row-(y+1) point, and one
 * random slot, one of them drawn at random.  "Nearest" is searched among
 * the four entries around a point's place in the Z-order (Morton) sort of
 * row y+1, ties to the earlier entry.  All slots propose at once each round;
 * a proposal runs only when it holds the smallest claim on both of its
 * slots, so the pairs that run are disjoint. */
namespace {
uint32_t zorder_(uint32_t x, uint32_t y) {
    uint32_t m = 0;
    for (int b = 0; b < 16; ++b) m |= (((x >> b) & 1u) << (2 * b)) | (((y >> b) & 1u) << (2 * b + 1));
    return m;
}
}  // namespace

uint64_t or_refine_knn_pass(uint64_t *pts, const int64_t *off, const int32_t *widths, int32_t nchains,
                            int32_t h, int32_t cap, uint64_t pass, int32_t rounds, uint64_t seed,
                            int64_t *gain_out, uint64_t *attempts_out) {
    auto pass_stream = [](uint64_t c, uint64_t p, uint64_t g) { return mix64_((c << 32) ^ p) ^ g; };
    uint64_t acc = 0, att = 0;
    int64_t gain = 0;
    if (h >= 2)
        for (int32_t c = 0; c < nchains; ++c) {
            const int32_t w = widths[c];
            if (w < 2) continue;
            uint64_t *cp = pts + size_t(off[c]) * h;
            const uint32_t y = uint32_t(mix64_(seed ^ 0x5851F42D4C957F2Dull ^ mix64_((uint64_t(c) << 32) ^ pass)) % uint32_t(h));
            const uint32_t yp = y > 0 ? y - 1 : h - 1, yn = (y + 1) % h;
            const int ng = (w + cap - 1) / cap, gs = (w + ng - 1) / ng;
            const uint64_t pkey = pass_stream(uint64_t(c), pass, 0xFFFFFFFFull);
            for (int g = 0; g < ng; ++g) {
                const int g0 = g * gs, n = std::min(gs, w - g0);
                std::vector<uint32_t> slot(n);
                std::vector<uint64_t> C(n), Pv(n), Nx(n);
                for (int l = 0; l < n; ++l) {
                    slot[l] = ng > 1 ? feistel_(uint32_t(g0 + l), uint32_t(w), pkey) : uint32_t(l);
                    Pv[l] = cp[size_t(slot[l]) * h + yp];
                    Nx[l] = cp[size_t(slot[l]) * h + yn];
                    C[l] = cp[size_t(slot[l]) * h + y];
                }
                if (n >= 2) {
                    // the Z-order of row y+1: entries (key << 16 | slot)
                    std::vector<uint64_t> order(n);
                    for (int l = 0; l < n; ++l)
                        order[l] = (uint64_t(zorder_(px_(Nx[l]), py_(Nx[l]))) << 16) | uint64_t(l);
                    std::sort(order.begin(), order.end());
                    std::vector<int> place(n);
                    for (int k = 0; k < n; ++k) place[order[k] & 0xFFFFu] = k;
                    auto nearest2 = [&](int lo, int hi, int self, uint64_t at, std::vector<uint32_t> &out) {
                        int b1 = -1, b2 = -1;
                        uint64_t d1 = 0, d2 = 0;
                        for (int k = std::max(lo, 0); k < std::min(hi, n); ++k) {
                            const int s = int(order[k] & 0xFFFFu);
                            if (s == self) continue;
                            const int64_t dx = int64_t(px_(Nx[s])) - px_(at), dy = int64_t(py_(Nx[s])) - py_(at);
                            const uint64_t dd = uint64_t(dx * dx + dy * dy);
                            if (b1 < 0 || dd < d1) {
                                b2 = b1; d2 = d1;
                                b1 = s; d1 = dd;
                            } else if (b2 < 0 || dd < d2) {
                                b2 = s; d2 = dd;
                            }
                        }
                        if (b1 >= 0) out.push_back(uint32_t(b1));
                        if (b2 >= 0) out.push_back(uint32_t(b2));
                    };
                    const uint64_t sid = pass_stream(uint64_t(c), pass, uint64_t(g));
                    std::vector<uint32_t> part(n), pri(n), claim(n);
                    for (int r = 0; r < rounds; ++r) {
                        std::fill(claim.begin(), claim.end(), 0xFFFFFFFFu);
                        for (int l = 0; l < n; ++l) {
                            const uint64_t d =
                                mix64_(seed ^ 0x2545F4914F6CDD1Dull ^ mix64_(sid ^ (uint64_t(r) << 32) ^ uint64_t(l)));
                            std::vector<uint32_t> cand;
                            const uint64_t q = C[l];
                            const uint64_t kq = uint64_t(zorder_(px_(q), py_(q))) << 16;
                            const int lo = int(std::lower_bound(order.begin(), order.end(), kq) - order.begin());
                            nearest2(lo - 2, lo + 2, l, q, cand);
                            const int kd = place[l];
                            nearest2(kd - 2, kd + 3, l, Nx[l], cand);
                            const uint32_t u = uint32_t((((d >> 32) & 0xFFFFu) * uint64_t(n - 1)) >> 16);
                            cand.push_back(u < uint32_t(l) ? u : u + 1u);
                            part[l] = cand[((d >> 48) * uint64_t(cand.size())) >> 16];
                            pri[l] = (uint32_t(d) & 0xFFFFF000u) | uint32_t(l);
                            claim[l] = std::min(claim[l], pri[l]);
                            claim[part[l]] = std::min(claim[part[l]], pri[l]);
                        }
                        for (int i = 0; i < n; ++i) {
                            const int j = int(part[i]);
                            if (claim[i] != pri[i] || claim[j] != pri[i]) continue;
                            int64_t c1 = int64_t(dist_(Pv[i], C[i]) + dist_(C[i], Nx[i]) +
                                                 dist_(Pv[j], C[j]) + dist_(C[j], Nx[j]));
                            int64_t c2 = int64_t(dist_(Pv[j], C[i]) + dist_(C[i], Nx[j]) +
                                                 dist_(Pv[i], C[j]) + dist_(C[j], Nx[i]));
                            ++att;
                            if (c1 >= c2) {
                                std::swap(C[i], C[j]);
                                gain += c1 - c2;
                                ++acc;
                            }
                        }
                    }
                }
                for (int l = 0; l < n; ++l) cp[size_t(slot[l]) * h + y] = C[l];
            }
        }
    if (gain_out) *gain_out = gain;
    if (attempts_out) *attempts_out = att;
    return acc;
}

/* color.cpp:31-44 */
uint32_t or_rgb_to_hsp(uint32_t c) {
    rgba_ v = unpack_(c);
    double h, s, p;
    rgb2hsp_(v.r / 255.0, v.g / 255.0, v.b / 255.0, &h, &s, &p);
    return pack_(narrow8_(std::round(h * 255.0)), narrow8_(std::round(s * 255.0)),
                 narrow8_(std::round(p * 255.0)), v.a);
}
/* color.cpp:46-59 (clamps only the upper bound). */
uint32_t or_hsp_to_rgb(uint32_t c) {
    rgba_ v = unpack_(c);
    double r, g, b;
    hsp2rgb_(v.r / 255.0, v.g / 255.0, v.b / 255.0, &r, &g, &b);
    return pack_(narrow8_(std::min(std::round(r * 255.0), 255.0)),
                 narrow8_(std::min(std::round(g * 255.0), 255.0)),
                 narrow8_(std::min(std::round(b * 255.0), 255.0)), v.a);
}

/* morph.cpp:298-301 (add_pixel) and 378-392 (get_pixel). */
void or_ingest(const uint32_t *rgba, size_t n, int hsp, uint32_t *stored, uint32_t *shown) {
    for (size_t i = 0; i < n; ++i) {
        uint32_t c = rgba[i], st = 0, sh = 0;
        if ((c >> 24) != 0) {
            st = hsp ? or_rgb_to_hsp(c) : c;
            sh = hsp ? or_hsp_to_rgb(st) : st;
        }
        if (stored) stored[i] = st;
        if (shown) shown[i] = sh;
    }
}

/* perlin.cpp:11-17 */
void or_perlin_perm(unsigned seed, int32_t *p) {
    if (seed == 0) seed = std::mt19937::default_seed;
    int tmp[256];
    std::iota(tmp, tmp + 256, 0);
    std::shuffle(tmp, tmp + 256, std::mt19937(seed));
    for (int i = 0; i < 256; ++i) { p[i] = tmp[i]; p[256 + i] = tmp[i]; }
}
double or_perlin_noise3(const int32_t *p, double x, double y, double z) { return noise3_(p, x, y, z); }
double or_perlin_octave2(const int32_t *p, double x, double y, int octaves) { return octave2_(p, x, y, octaves); }

/* spline.cpp:29-57; control points are the key positions in pixels. */
void or_spline_eval(const double *vx, const double *vy, int n, double t, double *ox, double *oy) {
    double delta_t = 1.0 / n;
    int p = (int)(t / delta_t);
    auto wrap = [n](int i) { return i < 0 ? n - 1 : (i >= n ? i - n : i); };
    int i0 = wrap(p - 1), i1 = wrap(p), i2 = wrap(p + 1), i3 = wrap(p + 2);
    double lt = (t - delta_t * (double)p) / delta_t;
    double t2 = lt * lt, t3 = t2 * lt;
    double b1 = 0.5 * (-t3 + 2.0 * t2 - lt);
    double b2 = 0.5 * (3.0 * t3 - 5.0 * t2 + 2.0);
    double b3 = 0.5 * (-3.0 * t3 + 4.0 * t2 + lt);
    double b4 = 0.5 * (t3 - t2);
    *ox = vx[i0] * b1 + vx[i1] * b2 + vx[i2] * b3 + vx[i3] * b4;
    *oy = vy[i0] * b1 + vy[i1] * b2 + vy[i2] * b3 + vy[i3] * b4;
}

/* thread.cpp:225-412 at blob_threshold 1.0: every 4-neighbour pair of present
 * pixels is mergeable (color_distance <= 1, color.h:17-24), so blobify ends
 * with one blob per 4-connected component.  Breadth-first flood fill. */
int or_segment(const uint8_t *rgba, int W, int H, int32_t *labels) {
    const size_t n = size_t(W) * H;
    for (size_t i = 0; i < n; ++i) labels[i] = -1;
    std::vector<int32_t> queue;
    int comps = 0;
    for (size_t s = 0; s < n; ++s) {
        if (rgba[4 * s + 3] == 0 || labels[s] >= 0) continue;
        ++comps;
        queue.clear();
        queue.push_back(int32_t(s));
        labels[s] = int32_t(s);  /* s is the smallest index: row-major scan */
        for (size_t q = 0; q < queue.size(); ++q) {
            int32_t v = queue[q];
            int x = v % W, y = v / W;
            const int nb[4][2] = {{x + 1, y}, {x - 1, y}, {x, y + 1}, {x, y - 1}};
            for (auto &c : nb) {
                if (c[0] < 0 || c[0] >= W || c[1] < 0 || c[1] >= H) continue;
                int32_t u = c[1] * W + c[0];
                if (rgba[4 * size_t(u) + 3] == 0 || labels[u] >= 0) continue;
                labels[u] = int32_t(s);
                queue.push_back(u);
            }
        }
    }
    return comps;
}

uint32_t or_create_color(double r, double g, double b, double a) {
    return pack_(narrow8_(std::round(r * 255.0)), narrow8_(std::round(g * 255.0)),
                 narrow8_(std::round(b * 255.0)), narrow8_(std::round(a * 255.0)));
}

uint32_t or_distinct_color(size_t group) {
    std::mt19937 gen(static_cast<std::mt19937::result_type>(group));
    std::uniform_int_distribution<unsigned char> ub(0, 255);
    unsigned r = ub(gen), g = ub(gen), b = ub(gen);
    return pack_(r, g, b, 255);
}

double or_get_time(size_t fi, size_t total, int finite, size_t keys) {
    /* morph.cpp:1522-1538 */
    if (total == 0) return 0.0;
    if (finite) {
        if (total == 1) {
            if (keys == 0) return 0.0;
            return (1.0 - (1.0 / double(keys))) / 2.0;
        }
        double t = double(fi) / double(total - 1);
        t *= 1.0 - (1.0 / double(keys));
        return t;
    }
    return double(fi) / double(total);
}

/* morph.cpp:1405-1421 (get_pixels(t)), 1302-1403 (draw_atoms),
 * 452-678 (get_pixels(blob)), 1431-1465 (get_background). */
void or_render_atoms(const or_render_cfg *cfg, const uint8_t *keys,
                     const uint64_t *pts, const int64_t *offsets,
                     const int64_t *widths, int32_t nchains,
                     const int32_t *order, int32_t nblobs,
                     const uint32_t *blob_rgba, double time, uint8_t *out) {
    const int W = cfg->width, Hh = cfg->height, h = cfg->nkeys;
    const size_t npx = size_t(W) * Hh;
    const double bx1 = cfg->bbox_x1, by1 = cfg->bbox_y1;
    const double bbox_w = cfg->bbox_x2 - cfg->bbox_x1 + 1.0;
    const double bbox_h = cfg->bbox_y2 - cfg->bbox_y1 + 1.0;
    (void)bx1; (void)by1;

    double integ;
    double tn = std::modf(time, &integ);
    if (tn < 0.0) tn += 1.0;
    size_t fk = size_t(tn * h);
    const bool have_key = h > 0 && fk < size_t(h);
    double tl = 0.0;
    if (have_key) {
        double dt = 1.0 / double(h);
        tl = std::max(0.0, (tn - (fk * dt)) / dt);
    }
    size_t fnext = have_key ? (fk + 1) % h : 0;

    auto key_px = [&](size_t f, size_t x, size_t y) -> rgba_ {
        if (x >= size_t(W) || y >= size_t(Hh)) return {0, 0, 0, 0};
        const uint8_t *q = keys + (f * npx + y * W + x) * 4;
        if (q[3] == 0) return {0, 0, 0, 0};
        return {q[0], q[1], q[2], q[3]};
    };

    /* background (morph.cpp:1410-1417, 1431-1465) */
    for (int y = 0; y < Hh; ++y)
        for (int x = 0; x < W; ++x) {
            rgba_ c = {0, 0, 0, 0};
            if (cfg->keep_background && have_key) {
                rgba_ c1 = key_px(fk, x, y), c2 = key_px(fnext, x, y);
                if (cfg->fading == 6) {
                    double f = 8.0;
                    double perlin_x = ((x - cfg->bbox_x1) / double(bbox_w)) * f;
                    double perlin_y = ((y - cfg->bbox_y1) / double(bbox_h)) * f;
                    double lag = octave2_(cfg->lag_perm, perlin_x, perlin_y, 8) * 0.5 + 0.5;
                    double slope = octave2_(cfg->slope_perm, perlin_x, perlin_y, 8) * 0.5 + 0.5;
                    c = fade_mix_(c1, c2, lag, slope, 1.0 - tl, cfg->fading);
                } else if (cfg->fading == 5) {
                    c = fade_mix_(c1, c2, 0.5, 0.5, 1.0 - tl, cfg->fading);
                } else {
                    c = mix_(c1, c2, 1.0 - tl);
                }
            }
            uint8_t *o = out + (size_t(y) * W + x) * 4;
            o[0] = uint8_t(c.r); o[1] = uint8_t(c.g); o[2] = uint8_t(c.b); o[3] = uint8_t(c.a);
        }
    if (!have_key) return;

    std::map<size_t, std::vector<rgba_>> comp;  /* draw_atoms' colors map */
    for (int32_t b = 0; b < nblobs; ++b) {
        int32_t grp = order[fk * nblobs + b];
        if (grp < 0 || grp >= nchains) break;  /* get_pixels returns nullptr */
        const size_t w = size_t(widths[grp]);
        const uint64_t *cp = pts + size_t(offsets[grp]) * h;
        const size_t y = fk, yn = (fk + 1) % h;

        std::map<size_t, std::vector<rgba_>> colors;
        std::map<size_t, std::vector<double>> weights;
        std::vector<double> vx(h), vy(h);
        for (size_t x = 0; x < w; ++x) {
            uint64_t p1 = cp[x * h + y], p2 = cp[x * h + yn];
            bool h1 = pfl_(p1) & HAS_PIXEL, h2 = pfl_(p2) & HAS_PIXEL;
            if (!h1 && !h2) continue;
            rgba_ c1, c2;
            if (h1 && !h2) { c1 = key_px(fk, px_(p1), py_(p1)); c2 = c1; c2.a = 0; }
            else if (h2 && !h1) { c2 = key_px(fnext, px_(p2), py_(p2)); c1 = c2; c1.a = 0; }
            else { c1 = key_px(fk, px_(p1), py_(p1)); c2 = key_px(fnext, px_(p2), py_(p2)); }

            uint32_t qx = px_(p1), qy = py_(p1), qxf = pxf_(p1), qyf = pyf_(p1);
            if (cfg->motion == 3) {         /* LINEAR, morph.cpp:1501-1515 */
                double w1 = 1.0 - tl;
                double x1 = (256.0 * px_(p1) + pxf_(p1)), y1 = (256.0 * py_(p1) + pyf_(p1));
                double x2 = (256.0 * px_(p2) + pxf_(p2)), y2 = (256.0 * py_(p2) + pyf_(p2));
                double xx = (w1 * x1 + (1.0 - w1) * x2);
                double yy = (w1 * y1 + (1.0 - w1) * y2);
                qx = narrow16_(xx / 256.0); qxf = narrow8_(xx - (int(qx) * 256));
                qy = narrow16_(yy / 256.0); qyf = narrow8_(yy - (int(qy) * 256));
            } else if (cfg->motion == 4) {  /* SPLINE, morph.cpp:525-531 */
                for (int j = 0; j < h; ++j) {
                    uint64_t q = cp[x * h + j];
                    vx[j] = px_(q) + pxf_(q) / double(256);
                    vy[j] = py_(q) + pyf_(q) / double(256);
                }
                double sx, sy, fr, ig;
                or_spline_eval(vx.data(), vy.data(), h, tn, &sx, &sy);
                fr = std::modf(sx, &ig); qx = narrow16_(ig); qxf = narrow8_(std::round(fr * 255));
                fr = std::modf(sy, &ig); qy = narrow16_(ig); qyf = narrow8_(std::round(fr * 255));
            }

            rgba_ c;
            if (cfg->fading == 6) {
                double f = 8.0;
                double perlin_x = (((int(px_(p1)) - cfg->bbox_x1) * 256 + int(pxf_(p1))) / double(bbox_w * 256)) * f;
                double perlin_y = (((int(py_(p1)) - cfg->bbox_y1) * 256 + int(pyf_(p1))) / double(bbox_h * 256)) * f;
                double lag = octave2_(cfg->lag_perm, perlin_x, perlin_y, 8) * 0.5 + 0.5;
                double slope = octave2_(cfg->slope_perm, perlin_x, perlin_y, 8) * 0.5 + 0.5;
                c = fade_mix_(c1, c2, lag, slope, 1.0 - tl, cfg->fading);
            } else if (cfg->fading == 5) {
                c = fade_mix_(c1, c2, 0.5, 0.5, 1.0 - tl, cfg->fading);
            } else {
                c = mix_(c1, c2, 1.0 - tl);
            }

            if (int(qx) >= W || int(qy) >= Hh) {
                if (int(qx) > cfg->bbox_x2 || int(qx) < cfg->bbox_x1 ||
                    int(qy) > cfg->bbox_y2 || int(qy) < cfg->bbox_y1) continue;
            }
            const double tw = 255 * 255;
            double w11 = ((255 - int(qxf)) * (255 - int(qyf))) / tw;
            double w21 = (int(qxf) * (255 - int(qyf))) / tw;
            double w12 = ((255 - int(qxf)) * int(qyf)) / tw;
            double w22 = (int(qxf) * int(qyf)) / tw;
            auto pos = [](uint32_t xx, uint32_t yy) { return size_t(yy & 0xFFFF) * 65536 + (xx & 0xFFFF); };
            int X = int(qx), Y = int(qy);
            if (w11 > 0.0) { size_t k = pos(X, Y); colors[k].push_back(c); weights[k].push_back(w11); }
            if ((X < cfg->bbox_x2 || X + 1 < W) && w21 > 0.0) {
                size_t k = pos(X + 1, Y); colors[k].push_back(c); weights[k].push_back(w21);
            }
            if ((Y < cfg->bbox_y2 || Y + 1 < Hh) && w12 > 0.0) {
                size_t k = pos(X, Y + 1); colors[k].push_back(c); weights[k].push_back(w12);
            }
            if (w22 > 0.0) {
                if ((Y < cfg->bbox_y2 && X < cfg->bbox_x2) || (Y + 1 < Hh && X + 1 < W)) {
                    size_t k = pos(X + 1, Y + 1); colors[k].push_back(c); weights[k].push_back(w22);
                }
            }
        }

        /* resolve (morph.cpp:591-623) */
        std::map<size_t, rgba_> blobpx;
        std::vector<std::pair<size_t, rgba_>> pixels;
        for (auto &kv : colors) {
            const auto &cs = kv.second;
            const auto &ws = weights[kv.first];
            double r = 0, g = 0, bb = 0, a = 0, wsum = 0;
            for (size_t i = 0; i < cs.size(); ++i) {
                wsum += ws[i];
                r += cs[i].r * ws[i];
                g += cs[i].g * ws[i];
                bb += cs[i].b * ws[i];
                a += cs[i].a * ws[i];
            }
            double wd = cfg->density > 0 ? double(cs.size()) / double(cfg->density) : 0.0;
            if (wd > 1.0) wd = 1.0;
            rgba_ q;
            q.r = narrow8_(std::round(r / wsum));
            q.g = narrow8_(std::round(g / wsum));
            q.b = narrow8_(std::round(bb / wsum));
            q.a = narrow8_(std::round(wd * (a / wsum)));
            if (cfg->feather > 0) blobpx[kv.first] = q;
            else pixels.push_back({kv.first, q});
        }
        if (cfg->feather > 0) {   /* inward peel, morph.cpp:625-674 */
            std::vector<std::set<size_t>> layers;
            std::map<size_t, rgba_> peeled = blobpx;
            while (!peeled.empty() && layers.size() < size_t(cfg->feather)) {
                std::set<size_t> border;
                for (auto &kv : peeled) {
                    size_t p = kv.first;
                    uint32_t x = p % 65536, y = p / 65536;
                    if (x == 0 || x == 65535 || y == 0 || y == 65535) { border.insert(p); continue; }
                    if (!peeled.count(size_t(y) * 65536 + x + 1) || !peeled.count(size_t(y) * 65536 + x - 1) ||
                        !peeled.count(size_t(y + 1) * 65536 + x) || !peeled.count(size_t(y - 1) * 65536 + x))
                        border.insert(p);
                }
                for (size_t p : border) peeled.erase(p);
                layers.push_back(border);
            }
            for (size_t l = 0; l < layers.size(); ++l)
                for (size_t p : layers[l]) {
                    rgba_ q = blobpx[p];
                    q.a = narrow8_(std::round(double(q.a) * (double(l + 1) / double(cfg->feather + 1))));
                    pixels.push_back({p, q});
                }
            for (auto &kv : peeled) pixels.push_back(kv);
        }
        /* draw_atoms consumes the list from the back (morph.cpp:1312-1345). */
        for (size_t i = pixels.size(); i-- > 0;) {
            size_t p = pixels[i].first;
            uint32_t x = p % 65536, y = p / 65536;
            size_t ip = size_t(y) * W + x;
            if (ip >= npx) continue;
            rgba_ q = pixels[i].second;
            if (cfg->show_blobs != 0 && blob_rgba) q = unpack_(blob_rgba[fk * nblobs + b]);
            if (q.a == 0) continue;
            comp[ip].push_back(q);
        }
    }

    /* composite, morph.cpp:1348-1401 */
    for (auto &kv : comp) {
        const auto &cs = kv.second;
        double r = 0, g = 0, b = 0, a = 0, rs = 0, gs = 0, bs = 0, as = 0;
        for (size_t i = 0; i < cs.size(); ++i) {
            double sr = cs[i].r / 255.0, sg = cs[i].g / 255.0, sb = cs[i].b / 255.0, sa = cs[i].a / 255.0;
            as += sa; rs += sr * sa; gs += sg * sa; bs += sb * sa;
            if (i == 0) { r = sr; g = sg; b = sb; a = sa; }
            else {
                r = sa * sr + (1.0 - sa) * r;
                g = sa * sg + (1.0 - sa) * g;
                b = sa * sb + (1.0 - sa) * b;
                a = a + (1.0 - a) * (cs[i].a / 255.0);
            }
        }
        if (cfg->blend_blobs) { r = rs / as; g = gs / as; b = bs / as; }
        uint8_t *o = out + kv.first * 4;
        if (cfg->keep_background) {
            double bgr = double(o[0]) / 255.0, bgg = double(o[1]) / 255.0;
            double bgb = double(o[2]) / 255.0, bga = double(o[3]) / 255.0;
            r = a * r + (1.0 - a) * bgr;
            g = a * g + (1.0 - a) * bgg;
            b = a * b + (1.0 - a) * bgb;
            a = bga + (1.0 - bga) * a;
        }
        o[0] = uint8_t(narrow8_(std::round(r * 255.0)));
        o[1] = uint8_t(narrow8_(std::round(g * 255.0)));
        o[2] = uint8_t(narrow8_(std::round(b * 255.0)));
        o[3] = uint8_t(narrow8_(std::round(a * 255.0)));
    }
}

/* fluidmodel.cpp:165-580, FP64, the mouse-drag branch omitted (never
 * pressed in a morph). */
void or_fluid_step(double *P, int32_t count, int32_t gsx, int32_t gsy,
                   int32_t steps_left, double freedom_radius, unsigned jitter_seed) {
    struct node_ { double m, d, gx, gy, u, v, ax, ay, r, g, b, a, weight; };
    std::vector<node_> grid(size_t(gsx) * gsy);
    std::memset(grid.data(), 0, grid.size() * sizeof(node_));
    auto N = [&](unsigned cx, unsigned cy) -> node_ & { return grid[size_t(cx) * gsy + cy]; };
    struct pk_ { unsigned cx, cy; double px[3], py[3], gx[3], gy[3]; };
    std::vector<pk_> K(count);
    const double mass = 1.0;
    if (jitter_seed) srand(jitter_seed);
    auto jit = [&]() { return jitter_seed ? static_cast<double>(rand()) / RAND_MAX * 0.01 : 0.0; };
    enum { X, Y, U, V, GX, GY, FR, R_, G_, B_, A_, r_, g_, b_, a_, STR, ACT, MAT };

    for (int i = 0; i < count; ++i) {
        double *p = P + 20 * size_t(i);
        if (p[ACT] == 0.0) continue;
        pk_ &k = K[i];
        k.cx = static_cast<int>(p[X] - 0.5);
        k.cy = static_cast<int>(p[Y] - 0.5);
        double x = k.cx - p[X];
        k.px[0] = (0.5 * x * x + 1.5 * x + 1.125); k.gx[0] = (x + 1.5);
        x += 1.0;
        k.px[1] = (-x * x + 0.75); k.gx[1] = (-2.0 * x);
        x += 1.0;
        k.px[2] = (0.5 * x * x - 1.5 * x + 1.125); k.gx[2] = (x - 1.5);
        double y = k.cy - p[Y];
        k.py[0] = (0.5 * y * y + 1.5 * y + 1.125); k.gy[0] = (y + 1.5);
        y += 1.0;
        k.py[1] = (-y * y + 0.75); k.gy[1] = (-2.0 * y);
        y += 1.0;
        k.py[2] = (0.5 * y * y - 1.5 * y + 1.125); k.gy[2] = (y - 1.5);
        for (unsigned a = 0; a < 3; ++a)
            for (unsigned b = 0; b < 3; ++b) {
                node_ &n = N(k.cx + a, k.cy + b);
                double phi = k.px[a] * k.py[b];
                n.m += phi * mass;
                n.d += phi;
                n.gx += k.gx[a] * k.py[b];
                n.gy += k.px[a] * k.gy[b];
                if (p[MAT] != 0.0 && p[STR] > 0.0) {
                    double nw = p[STR], ow = n.weight, sw = nw + ow;
                    n.r = (ow * n.r + nw * p[r_]) / sw;
                    n.g = (ow * n.g + nw * p[g_]) / sw;
                    n.b = (ow * n.b + nw * p[b_]) / sw;
                    n.a = (ow * n.a + nw * p[a_]) / sw;
                    n.weight = sw;
                }
            }
    }
    for (int i = 0; i < count; ++i) {
        double *p = P + 20 * size_t(i);
        if (p[ACT] == 0.0) continue;
        pk_ &k = K[i];
        unsigned cx = static_cast<int>(p[X]), cy = static_cast<int>(p[Y]);
        node_ &n01 = N(cx, cy), &n02 = N(cx, cy + 1), &n11 = N(cx + 1, cy), &n12 = N(cx + 1, cy + 1);
        double pdx = n11.d - n01.d, pdy = n02.d - n01.d;
        double C20 = 3.0 * pdx - n11.gx - 2.0 * n01.gx;
        double C02 = 3.0 * pdy - n02.gy - 2.0 * n01.gy;
        double C30 = -2.0 * pdx + n11.gx + n01.gx;
        double C03 = -2.0 * pdy + n02.gy + n01.gy;
        double csum1 = n01.d + n01.gy + C02 + C03;
        double csum2 = n01.d + n01.gx + C20 + C30;
        double C21 = 3.0 * n12.d - 2.0 * n02.gx - n12.gx - 3.0 * csum1 - C20;
        double C31 = -2.0 * n12.d + n02.gx + n12.gx + 2.0 * csum1 - C30;
        double C12 = 3.0 * n12.d - 2.0 * n11.gy - n12.gy - 3.0 * csum2 - C02;
        double C13 = -2.0 * n12.d + n11.gy + n12.gy + 2.0 * csum2 - C03;
        double C11 = n02.gx - C13 - C12 - n01.gx;
        double u = p[X] - cx, u2 = u * u, u3 = u * u2;
        double v = p[Y] - cy, v2 = v * v, v3 = v * v2;
        double density = n01.d + n01.gx * u + n01.gy * v + C20 * u2 + C02 * v2 + C30 * u3 +
                         C03 * v3 + C21 * u2 * v + C31 * u3 * v + C12 * u * v2 + C13 * u * v3 + C11 * u * v;
        double pressure = density - 1.0;
        if (pressure > 2.0) pressure = 2.0;
        double fx = 0.0, fy = 0.0;
        if (p[X] < 4.0) fx += mass * (4.0 - p[X]);
        else if (p[X] > gsx - 5) fx += mass * (gsx - 5 - p[X]);
        if (p[Y] < 4.0) fy += mass * (4.0 - p[Y]);
        else if (p[Y] > gsy - 5) fy += mass * (gsy - 5 - p[Y]);
        for (unsigned a = 0; a < 3; ++a)
            for (unsigned b = 0; b < 3; ++b) {
                node_ &n = N(k.cx + a, k.cy + b);
                double phi = k.px[a] * k.py[b];
                n.ax += -((k.gx[a] * k.py[b]) * pressure) + fx * phi;
                n.ay += -((k.px[a] * k.gy[b]) * pressure) + fy * phi;
            }
    }
    for (auto &n : grid)
        if (n.m > 0.0) { n.ax /= n.m; n.ay /= n.m; }

    for (int i = 0; i < count; ++i) {
        double *p = P + 20 * size_t(i);
        if (p[ACT] == 0.0) continue;
        pk_ &k = K[i];
        /* attractor towards the trajectory point (fluidmodel.cpp:383-410) */
        double cax = 0.0, cay = 0.0;
        {
            double x1 = p[X], y1 = p[Y], x2 = p[GX], y2 = p[GY], a = 0.03;
            double A = std::abs(y1 - y2), B = std::abs(x1 - x2), C = std::sqrt(A * A + B * B);
            if (a >= C) a = C;
            if (B <= 0.0) { if (y2 <= y1) cay -= a; else cay += a; }
            else if (C > 0.0) {
                double dx = (a * B) / C, dy = (A * dx) / B;
                if (x1 <= x2) cax += dx; else cax -= dx;
                if (y1 <= y2) cay += dy; else cay -= dy;
            }
        }
        for (unsigned a = 0; a < 3; ++a)
            for (unsigned b = 0; b < 3; ++b) {
                node_ &n = N(k.cx + a, k.cy + b);
                double phi = k.px[a] * k.py[b];
                p[U] += phi * (n.ax + cax);
                p[V] += phi * (n.ay + cay);
            }
        double mu = mass * p[U], mv = mass * p[V];
        if (p[MAT] == 0.0) { mu *= 0.0; mv *= 0.0; }
        for (unsigned a = 0; a < 3; ++a)
            for (unsigned b = 0; b < 3; ++b) {
                node_ &n = N(k.cx + a, k.cy + b);
                double phi = k.px[a] * k.py[b];
                n.u += phi * mu;
                n.v += phi * mv;
            }
    }
    for (auto &n : grid)
        if (n.m > 0.0) { n.u /= n.m; n.v /= n.m; }

    for (int i = 0; i < count; ++i) {
        double *p = P + 20 * size_t(i);
        if (p[ACT] == 0.0) continue;
        pk_ &k = K[i];
        double nR = 0, nG = 0, nB = 0, nA = 0, wsum = 0, gu = 0, gv = 0;
        for (unsigned a = 0; a < 3; ++a)
            for (unsigned b = 0; b < 3; ++b) {
                node_ &n = N(k.cx + a, k.cy + b);
                double phi = k.px[a] * k.py[b];
                gu += phi * n.u;
                gv += phi * n.v;
                if (n.weight > 0.0) {
                    wsum += n.weight;
                    nR += n.r * n.weight; nG += n.g * n.weight;
                    nB += n.b * n.weight; nA += n.a * n.weight;
                }
            }
        if (wsum > 0.0) {
            nR /= wsum; double wr = std::abs(nR - p[R_]);
            nG /= wsum; double wg = std::abs(nG - p[G_]);
            nB /= wsum; double wb = std::abs(nB - p[B_]);
            nA /= wsum; double wa = std::abs(nA - p[A_]);
            if (p[MAT] == 0.0) { p[r_] = nR; p[g_] = nG; p[b_] = nB; p[a_] = nA; }
            else {
                p[r_] = (1.0 - wr) * p[r_] + wr * nR;
                p[g_] = (1.0 - wg) * p[g_] + wg * nG;
                p[b_] = (1.0 - wb) * p[b_] + wb * nB;
                p[a_] = (1.0 - wa) * p[a_] + wa * nA;
            }
        }
        p[X] += gu;
        p[Y] += gv;
        {   /* freedom radius (fluidmodel.cpp:501-539) */
            double x1 = p[X], y1 = p[Y], x2 = p[GX], y2 = p[GY];
            double A = std::abs(y1 - y2), B = std::abs(x1 - x2), C = std::sqrt(A * A + B * B);
            double r = freedom_radius * p[FR];
            if (C > r) {
                double mx = 0.0, my = 0.0, a = C - r;
                if (a >= C) a = C;
                if (B <= 0.0) { if (y2 <= y1) my -= a; else my += a; }
                else if (C > 0.0) {
                    double dx = (a * B) / C, dy = (A * dx) / B;
                    if (x1 <= x2) mx += dx; else mx -= dx;
                    if (y1 <= y2) my += dy; else my -= dy;
                }
                double w = 1.0 / (steps_left + 1);
                mx *= w; my *= w;
                p[X] += mx; p[Y] += my;
            }
        }
        p[U] += gu - p[U];
        p[V] += gv - p[V];
        if (p[X] < 1.0) { p[X] = 1.0 + jit(); p[U] = 0.0; }
        else if (p[X] > gsx - 2) { p[X] = gsx - 2 - jit(); p[U] = 0.0; }
        if (p[Y] < 1.0) { p[Y] = 1.0 + jit(); p[V] = 0.0; }
        else if (p[Y] > gsy - 2) { p[Y] = gsy - 2 - jit(); p[V] = 0.0; }
    }
}

/* Hungarian method with potentials (shortest augmenting paths), int64. */
int64_t or_lap_min_cost(const int64_t *cost, int32_t n, int32_t *assign) {
    const int64_t INF = std::numeric_limits<int64_t>::max() / 4;
    std::vector<int64_t> u(n + 1, 0), v(n + 1, 0), minv(n + 1);
    std::vector<int32_t> p(n + 1, 0), way(n + 1, 0);
    std::vector<char> used(n + 1);
    for (int32_t i = 1; i <= n; ++i) {
        p[0] = i;
        int32_t j0 = 0;
        std::fill(minv.begin(), minv.end(), INF);
        std::fill(used.begin(), used.end(), 0);
        do {
            used[j0] = 1;
            int32_t i0 = p[j0], j1 = 0;
            int64_t delta = INF;
            for (int32_t j = 1; j <= n; ++j) {
                if (used[j]) continue;
                int64_t cur = cost[size_t(i0 - 1) * n + (j - 1)] - u[i0] - v[j];
                if (cur < minv[j]) { minv[j] = cur; way[j] = j0; }
                if (minv[j] < delta) { delta = minv[j]; j1 = j; }
            }
            for (int32_t j = 0; j <= n; ++j) {
                if (used[j]) { u[p[j]] += delta; v[j] -= delta; }
                else minv[j] -= delta;
            }
            j0 = j1;
        } while (p[j0] != 0);
        do {
            int32_t j1 = way[j0];
            p[j0] = p[j1];
            j0 = j1;
        } while (j0);
    }
    int64_t total = 0;
    for (int32_t j = 1; j <= n; ++j) {
        if (assign) assign[p[j] - 1] = j - 1;
        total += cost[size_t(p[j] - 1) * n + (j - 1)];
    }
    return total;
}

}  /* extern "C" */
