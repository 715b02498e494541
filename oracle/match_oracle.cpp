/*
 * Blob matching restated for the CPU oracle (TEST INFRASTRUCTURE ONLY, see
 * oracle.h).  Follows thread::match (thread.cpp:598-738), its energy
 * (get_energy(blob***), thread.cpp:1087-1107), blob_distance
 * (thread.cpp:1151-1174), set_blob_weights (thread.cpp:1141-1149), the
 * step counter of thread::step (thread.cpp:139-171) that drives the
 * degeneration, and init_morph's group refresh plus fix_volatiles
 * (thread.cpp:746-768, 1187-1233).
 *
 * Data layout: a blob is an index into its frame's list; the map is
 * map[x][y] = list index of the blob in column x of key frame y, and a
 * blob's group is the column it held at the last improvement of the best
 * energy, as the reference's blob->group is.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "oracle.h"

namespace {

struct MBlob {
    double x = 0, y = 0, r = 0, g = 0, b = 0, a = 0;
    int64_t size = 0;
    int32_t src = -1;  // index in the caller's list, -1 for a padding blob
    int32_t group = 0;
};

struct Matcher {
    int h = 0;
    size_t bw = 0;
    std::vector<std::vector<MBlob>> list;  // [y][list index]
    std::vector<int32_t> map;              // [x * h + y] -> list index
    uint32_t bbox_d = 0;
    double w_rgba = 0.33, w_size = 0.33, w_xy = 0.34;  // thread.h:107-109 until set

    // blob_distance (thread.cpp:1151-1174): create_pixel truncates the
    // means to uint16 and rounds the colours (color.cpp:22-29).
    double distance(const MBlob &b1, const MBlob &b2) const {
        const int64_t s1 = b1.size, s2 = b2.size, ss = s1 + s2;
        double pix = 0.0, col = 0.0, siz = 0.0;
        if (ss > 0) siz = std::fabs(double(s1) - double(s2)) / double(ss);
        if (s1 > 0 && s2 > 0) {
            const int32_t xd = int32_t(uint16_t(b1.x)) - int32_t(uint16_t(b2.x));
            const int32_t yd = int32_t(uint16_t(b1.y)) - int32_t(uint16_t(b2.y));
            pix = std::sqrt(double(uint32_t(xd * xd + yd * yd)) / bbox_d);
            auto ch = [](double v) { return int(uint8_t(std::round(v * 255.0))); };
            const int16_t rd = int16_t(ch(b1.r) - ch(b2.r)), gd = int16_t(ch(b1.g) - ch(b2.g));
            const int16_t bd = int16_t(ch(b1.b) - ch(b2.b)), ad = int16_t(ch(b1.a) - ch(b2.a));
            col = std::sqrt(double(rd * rd + gd * gd + bd * bd + ad * ad)) / 510.0;
        }
        return w_xy * pix + w_rgba * col + w_size * siz;
    }
    const MBlob &at(size_t x, int y) const { return list[y][map[x * h + y]]; }
    // get_energy(blob***) (thread.cpp:1087-1107)
    double energy() const {
        double e = 0.0;
        for (size_t x = 0; x < bw; ++x) {
            const MBlob *prev = &at(x, h - 1);
            for (int y = 0; y < h; ++y) {
                const MBlob *cur = &at(x, y);
                e += distance(*prev, *cur);
                prev = cur;
            }
        }
        return e;
    }
};

}  // namespace

extern "C" int32_t or_match(int32_t h, const int32_t *nblobs, const int64_t *sizes, const double *means,
                            const double *fmeans, const int32_t *bbox, const uint8_t *weights,
                            uint64_t degeneration, uint64_t rng_state, uint64_t steps, int fix, int32_t *order,
                            int32_t *group, double *energy, double *fixed, uint64_t *rng_out) {
    if (h <= 0 || !nblobs || !fmeans || !bbox || !weights) return -1;
    Matcher M;
    M.h = h;
    M.list.resize(h);
    size_t k = 0;
    for (int y = 0; y < h; ++y) {
        for (int32_t b = 0; b < nblobs[y]; ++b, ++k) {
            MBlob bl;
            const double *m = means + 6 * k;
            bl.x = m[0]; bl.y = m[1]; bl.r = m[2]; bl.g = m[3]; bl.b = m[4]; bl.a = m[5];
            bl.size = sizes[k];
            bl.src = b;
            M.list[y].push_back(bl);
        }
        M.bw = std::max(M.bw, size_t(nblobs[y]));
    }
    // thread::set_bbox / set_blob_weights as morph::synchronize calls them
    // (morph.cpp:119-120, thread.cpp:1127-1149)
    if (bbox[0] <= bbox[2] && bbox[1] <= bbox[3]) {
        const int dx = bbox[2] - bbox[0], dy = bbox[3] - bbox[1];
        M.bbox_d = uint32_t(dx * dx + dy * dy);
    }
    const double wsum = double(weights[0] * weights[0] + weights[1] * weights[1] + weights[2] * weights[2]);
    if (wsum != 0.0) {
        M.w_rgba = double(weights[0] * weights[0]) / wsum;
        M.w_size = double(weights[1] * weights[1]) / wsum;
        M.w_xy = 1.0 - (M.w_rgba + M.w_size);
    }
    std::default_random_engine e1(static_cast<std::default_random_engine::result_type>(rng_state));
    double map_e = 0.0, best_e = std::numeric_limits<double>::infinity(), best_report = 0.0;
    bool deviant = false;
    size_t counter = 1;  // thread::step leaves it at 1 entering BLOB_MATCHING
    bool built = false;
    for (uint64_t s = 0; s < steps; ++s, ++counter) {
        if (!built) {  // thread.cpp:599-665
            for (int y = 0; y < h; ++y) {
                const double *fm = fmeans + 6 * y;
                while (M.list[y].size() < M.bw) {
                    MBlob e;
                    e.x = fm[0]; e.y = fm[1]; e.r = fm[2]; e.g = fm[3]; e.b = fm[4]; e.a = fm[5];
                    M.list[y].push_back(e);
                }
                std::shuffle(M.list[y].begin(), M.list[y].end(), e1);
                for (size_t b = 0; b < M.bw; ++b) M.list[y][b].group = int32_t(b);
            }
            M.map.assign(M.bw * h, 0);
            for (size_t x = 0; x < M.bw; ++x)
                for (int y = 0; y < h; ++y) M.map[x * h + y] = int32_t(x);
            map_e = M.energy();
            built = true;
            continue;
        }
        if (M.bw <= 1) break;  // thread.cpp:667: matching is over
        std::uniform_int_distribution<size_t> ux(0, M.bw - 1), uy(0, size_t(h) - 1);
        const size_t y = uy(e1), yn = (y + 1) % h, yp = y > 0 ? y - 1 : h - 1;
        const size_t x1 = ux(e1);
        size_t x2;
        do {
            x2 = ux(e1);
        } while (x1 == x2);
        const MBlob &a_p = M.at(x1, int(yp)), &b_p = M.at(x2, int(yp));
        const MBlob &a_c = M.at(x1, int(y)), &b_c = M.at(x2, int(y));
        const MBlob &a_n = M.at(x1, int(yn)), &b_n = M.at(x2, int(yn));
        if (a_c.size == 0 && b_c.size == 0) continue;
        const double c1 = (M.distance(a_p, a_c) + M.distance(a_c, a_n)) + (M.distance(b_p, b_c) + M.distance(b_c, b_n));
        const double c2 = (M.distance(b_p, a_c) + M.distance(a_c, b_n)) + (M.distance(a_p, b_c) + M.distance(b_c, a_n));
        if (c1 >= c2 || (degeneration && (counter % degeneration) == 0)) {
            std::swap(M.map[x1 * h + y], M.map[x2 * h + y]);
            map_e -= c1 - c2;
            if (map_e < 0.0) map_e = 0.0;
            if (map_e < best_e) {
                best_e = map_e;
                best_report = best_e;
                counter = 0;
                if (deviant) {
                    for (int yy = 0; yy < h; ++yy)
                        for (size_t xx = 0; xx < M.bw; ++xx) M.list[yy][M.map[xx * h + yy]].group = int32_t(xx);
                } else {
                    M.list[y][M.map[x1 * h + y]].group = int32_t(x1);
                    M.list[y][M.map[x2 * h + y]].group = int32_t(x2);
                }
                deviant = false;
            } else {
                deviant = true;
            }
        }
        if (best_e == 0.0) break;  // matching is over
    }
    if (energy) {
        energy[0] = map_e;
        energy[1] = best_e;
        energy[2] = best_report;
    }
    if (!built) return int32_t(M.bw);
    for (int y = 0; y < h; ++y)
        for (size_t b = 0; b < M.bw; ++b) {
            if (order) order[y * M.bw + b] = M.list[y][b].src;
            if (group) group[y * M.bw + b] = M.list[y][b].group;
        }
    if (fix && fixed) {
        // init_morph: the map follows the groups (thread.cpp:746-758), then
        // every column's empty blobs take positions interpolated between
        // the real blobs around them, cyclically (thread.cpp:1187-1233).
        for (int y = 0; y < h; ++y)
            for (size_t b = 0; b < M.bw; ++b) M.map[size_t(M.list[y][b].group) * h + y] = int32_t(b);
        for (size_t x = 0; x < M.bw && h > 1; ++x) {
            size_t i = 0;
            bool started = false;
            std::vector<MBlob *> vol;
            MBlob *first = nullptr, *prev = nullptr;
            for (;;) {
                i = (i + 1) % size_t(h);
                MBlob *bl = &M.list[i][M.map[x * h + i]];
                const bool empty = bl->size == 0;
                if (!started) {
                    if (empty) {
                        if (i == 0) break;
                        continue;
                    }
                    started = true;
                    first = prev = bl;
                    continue;
                }
                if (empty) {
                    vol.push_back(bl);
                    continue;
                }
                for (size_t v = 0; v < vol.size(); ++v) {
                    const double t = (v + 1.0) / double(vol.size() + 1.0);
                    vol[v]->x = t * bl->x + (1.0 - t) * prev->x;
                    vol[v]->y = t * bl->y + (1.0 - t) * prev->y;
                }
                vol.clear();
                prev = bl;
                if (prev == first) break;
            }
        }
        for (int y = 0; y < h; ++y)
            for (size_t b = 0; b < M.bw; ++b) {
                fixed[2 * (y * M.bw + b)] = M.list[y][b].x;
                fixed[2 * (y * M.bw + b) + 1] = M.list[y][b].y;
            }
    }
    if (rng_out) {
        std::ostringstream os;
        os << e1;
        *rng_out = std::stoull(os.str());
    }
    return int32_t(M.bw);
}

// Energy of a map given as group[y][b] (column of list entry b of frame y)
// over frames already padded to bw blobs; the exhaustive minimum over every
// assignment of frames 1..h-1 to the columns (frame 0 fixed).
namespace {
void load_lists(Matcher &M, int32_t h, int32_t bw, const int64_t *sizes, const double *means, const int32_t *bbox,
                const uint8_t *weights) {
    M.h = h;
    M.bw = size_t(bw);
    M.list.assign(h, std::vector<MBlob>(bw));
    for (int y = 0; y < h; ++y)
        for (int b = 0; b < bw; ++b) {
            MBlob &bl = M.list[y][b];
            const double *m = means + 6 * (size_t(y) * bw + b);
            bl.x = m[0]; bl.y = m[1]; bl.r = m[2]; bl.g = m[3]; bl.b = m[4]; bl.a = m[5];
            bl.size = sizes[size_t(y) * bw + b];
        }
    if (bbox[0] <= bbox[2] && bbox[1] <= bbox[3]) {
        const int dx = bbox[2] - bbox[0], dy = bbox[3] - bbox[1];
        M.bbox_d = uint32_t(dx * dx + dy * dy);
    }
    const double wsum = double(weights[0] * weights[0] + weights[1] * weights[1] + weights[2] * weights[2]);
    if (wsum != 0.0) {
        M.w_rgba = double(weights[0] * weights[0]) / wsum;
        M.w_size = double(weights[1] * weights[1]) / wsum;
        M.w_xy = 1.0 - (M.w_rgba + M.w_size);
    }
    M.map.assign(size_t(bw) * h, 0);
}
}  // namespace

extern "C" double or_match_energy(int32_t h, int32_t bw, const int64_t *sizes, const double *means,
                                  const int32_t *bbox, const uint8_t *weights, const int32_t *group) {
    Matcher M;
    load_lists(M, h, bw, sizes, means, bbox, weights);
    for (int y = 0; y < h; ++y)
        for (int b = 0; b < bw; ++b) M.map[size_t(group[y * bw + b]) * h + y] = b;
    return M.energy();
}

extern "C" double or_match_brute(int32_t h, int32_t bw, const int64_t *sizes, const double *means,
                                 const int32_t *bbox, const uint8_t *weights, int32_t *best_group) {
    Matcher M;
    load_lists(M, h, bw, sizes, means, bbox, weights);
    if (h < 1 || bw < 1 || h > 3 || bw > 8) return -1.0;
    std::vector<std::vector<int32_t>> perm(h, std::vector<int32_t>(bw));
    for (int y = 0; y < h; ++y) std::iota(perm[y].begin(), perm[y].end(), 0);
    double best = std::numeric_limits<double>::infinity();
    // perm[y][x] = list index of frame y in column x; frame 0 is fixed
    auto eval = [&]() {
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < bw; ++x) M.map[size_t(x) * h + y] = perm[y][x];
        const double e = M.energy();
        if (e < best) {
            best = e;
            if (best_group)
                for (int y = 0; y < h; ++y)
                    for (int x = 0; x < bw; ++x) best_group[y * bw + perm[y][x]] = x;
        }
    };
    if (h == 1) {
        eval();
    } else if (h == 2) {
        do eval(); while (std::next_permutation(perm[1].begin(), perm[1].end()));
    } else {
        do {
            std::sort(perm[2].begin(), perm[2].end());
            do eval(); while (std::next_permutation(perm[2].begin(), perm[2].end()));
        } while (std::next_permutation(perm[1].begin(), perm[1].end()));
    }
    return best;
}
