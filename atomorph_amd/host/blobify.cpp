/*
 * Greedy blob detection and dust unification (thread.cpp:225-596).
 *
 * Detection starts every key frame from one blob per pixel, shuffled.  Each
 * worker step takes the last blob of each frame's list and merges it with
 * the neighbouring blob of the nearest mean colour within the threshold;
 * ties are broken by reservoir sampling.  A blob that finds no partner is
 * moved to the front region of blobs that cannot expand.  Unification then
 * merges undersized blobs ("dust") into the nearest of a few sampled blobs
 * whose rounded centres lie within the grip box.
 *
 * The random draws come from one std::default_random_engine seeded with the
 * morph's seed and shared by the key frames in step order, through the same
 * standard algorithms (std::shuffle, std::uniform_int_distribution), so for
 * a fresh morph the blobs -- membership, list order and running means --
 * are the reference's.  The merge arithmetic follows the reference's order
 * of operations.
 *
 * Layout: a pixel is its canvas index y*W + x, which orders like the
 * reference's y*65536 + x.  Membership is a union-find forest over pixels
 * (the reference keeps an owner map and a surface set per blob); a border is
 * an ordered set of pixel indices, iterated in ascending order as the
 * reference's std::set is, and merged small into large.
 */
#include "blobify.h"
#include "color.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <random>
#include <set>
#include <utility>

namespace am_blobify {
namespace {

using am::color;

struct Blob {
    double x = 0, y = 0, r = 0, g = 0, b = 0, a = 0;
    uint64_t group = 0;
    size_t size = 0;     // pixels
    size_t index = 0;    // place in the frame's blob list
    int32_t root = -1;   // union-find root of its pixels
    bool unified = false;
    uint32_t stamp = 0;  // == Frame::stamp: checked in the current pass
    std::set<int32_t> border;
};

struct Frame {
    int W = 0, H = 0;
    const uint32_t *px = nullptr;
    std::vector<int32_t> parent;   // union-find over pixels; -1: no pixel
    std::vector<int32_t> blob_of;  // blob id of a root pixel
    std::vector<Blob> pool;        // blobs by id
    std::vector<int32_t> blobs;    // the blob list, -1 = an empty slot
    size_t first_expansion = 0;    // [0, first_expansion): blobs that cannot expand
    size_t first_dust = 0;         // [0, first_dust): blobs unification is done with
    uint32_t stamp = 0;
    bool started = false;          // the one-pixel blobs were made
    std::vector<int32_t> erased;   // border entries to drop after a pass

    bool present(int i) const { return (px[i] >> 24) != 0; }
    int32_t find(int32_t i) {
        while (parent[i] != i) {
            parent[i] = parent[parent[i]];
            i = parent[i];
        }
        return i;
    }
    int32_t owner(int32_t pos) { return parent[pos] < 0 ? -1 : blob_of[find(pos)]; }
    // every pixel of blob `from` now belongs to blob `to`
    void take_pixels(int32_t to, int32_t from) {
        Blob &t = pool[to], &f = pool[from];
        int32_t ra = t.root, rb = f.root;
        if (t.size < f.size) std::swap(ra, rb);
        parent[rb] = ra;
        t.root = ra;
        blob_of[ra] = to;
        t.size += f.size;
        f.size = 0;
    }
    // a blob merged away: its border's memory back
    void kill(int32_t id) { std::set<int32_t>().swap(pool[id].border); }
};

color mean_color(const Blob &b) { return am::create_color(b.r, b.g, b.b, b.a); }

// The running means of `into` after absorbing `from` (thread.cpp:347-357).
void merge_means(Blob &into, const Blob &from) {
    const size_t s1 = into.size, s2 = from.size;
    const double w = s2 / double(s1 + s2);
    into.x = ((1.0 - w) * into.x + w * from.x);
    into.y = ((1.0 - w) * into.y + w * from.y);
    into.r = ((1.0 - w) * into.r + w * from.r);
    into.g = ((1.0 - w) * into.g + w * from.g);
    into.b = ((1.0 - w) * into.b + w * from.b);
    into.a = ((1.0 - w) * into.a + w * from.a);
}

// One blob per pixel in canvas order, then shuffled (thread.cpp:231-265).
bool start_frame(Frame &f, const options &o, std::default_random_engine &e1) {
    f.started = true;
    const int W = f.W, H = f.H;
    const int32_t npx = W * H;
    f.parent.assign(npx, -1);
    f.blob_of.assign(npx, -1);
    f.pool.clear();
    for (int32_t i = 0; i < npx; ++i) {
        if (!f.present(i)) continue;
        const int x = i % W, y = i / W;
        const int32_t id = int32_t(f.pool.size());
        f.pool.emplace_back();
        Blob &bl = f.pool.back();
        const color c = am::word_color(f.px[i]);
        bl.x = x;
        bl.y = y;
        bl.r = c.r / 255.0;
        bl.g = c.g / 255.0;
        bl.b = c.b / 255.0;
        bl.a = c.a / 255.0;
        bl.group = uint64_t(y) * 65536u + uint64_t(x);
        bl.size = 1;
        bl.root = i;
        f.parent[i] = i;
        f.blob_of[i] = id;
        if (x + 1 < W && f.present(i + 1)) bl.border.insert(i + 1);
        if (x > 0 && f.present(i - 1)) bl.border.insert(i - 1);
        if (y + 1 < H && f.present(i + W)) bl.border.insert(i + W);
        if (y > 0 && f.present(i - W)) bl.border.insert(i - W);
        f.blobs.push_back(id);
    }
    std::shuffle(f.blobs.begin(), f.blobs.end(), e1);
    for (size_t k = 0; k < f.blobs.size(); ++k) f.pool[f.blobs[k]].index = k;
    return f.blobs.empty() || o.max_size == 1;
}

// One detection step on a frame (thread.cpp:225-412).  true: done.
bool blobify_step(Frame &f, const options &o, std::default_random_engine &e1) {
    // the reference makes the blobs whenever the list is empty; for a frame
    // without pixels that makes nothing and draws nothing, every step
    if (f.blobs.empty()) return f.started ? true : start_frame(f, o, e1);
    size_t sz;
    int32_t id;
    for (;;) {  // drop empty slots and empty blobs from the end
        sz = f.blobs.size();
        if (sz <= 1) return true;
        id = f.blobs[sz - 1];
        if (id >= 0 && f.pool[id].size > 0) break;
        f.blobs.pop_back();
        if (id >= 0) f.kill(id);
    }
    const size_t b = sz - 1;
    if (f.blobs.size() <= o.number) return true;
    Blob &bl = f.pool[id];
    f.erased.clear();
    bool best_found = false;
    double best_dist = 0.0;
    int32_t best = -1;
    size_t best_equal = 0;
    if (bl.size >= o.min_size) bl.unified = true;
    bool check = true, scan = bl.size < o.max_size;
    ++f.stamp;  // nothing checked yet
    for (;;) {
        if (scan) {
            const color mine = mean_color(bl);
            for (const int32_t pos : bl.border) {
                const int32_t own = f.owner(pos);
                if (own == id || own < 0 || f.pool[own].stamp == f.stamp) {
                    // itself, or a neighbour already seen through another pixel
                    if (check && bl.size >= o.min_size) f.erased.push_back(pos);
                    continue;
                }
                Blob &nb = f.pool[own];
                nb.stamp = f.stamp;
                const double dist = am::color_distance(mean_color(nb), mine);
                if ((dist <= o.threshold || !check) && (!best_found || dist <= best_dist)) {
                    if (best_found && dist == best_dist) {
                        std::uniform_int_distribution<size_t> pick(0, best_equal);
                        best_equal++;
                        if (pick(e1) != 0) continue;
                    } else {
                        best_equal = 1;
                    }
                    best_found = true;
                    best = own;
                    best_dist = dist;
                }
            }
        }
        if (best_found) break;
        if (bl.size < o.min_size && !bl.border.empty() && check) {
            // an undersized outlier with neighbours: again without the threshold
            check = false;
            scan = true;
            ++f.stamp;
            continue;
        }
        break;
    }
    if (best_found) {
        Blob &bm = f.pool[best];
        merge_means(bl, bm);
        f.take_pixels(id, best);
        if (bl.border.size() < bm.border.size()) bl.border.swap(bm.border);
        bl.border.insert(bm.border.begin(), bm.border.end());
        const size_t bi = bm.index;
        f.blobs[bi] = id;
        bl.index = bi;
        bl.unified = bl.unified || bm.unified;
        f.kill(best);
        f.blobs.pop_back();
    } else {
        const size_t bb = f.first_expansion;
        std::set<int32_t>().swap(bl.border);  // it does not expand any more
        if (bb >= sz) return true;            // nothing left that can expand
        const int32_t other = f.blobs[bb];
        if (other >= 0) f.pool[other].index = b;
        f.blobs[b] = other;
        f.blobs[bb] = id;
        bl.index = bb;
        f.first_expansion++;
    }
    for (const int32_t pos : f.erased) bl.border.erase(pos);
    return false;
}

// Dust unification of a frame until it reports done (thread.cpp:426-585);
// unification draws nothing random, so frames run one after another.
bool unify_frame(Frame &f, const options &o) {
    while (!f.blobs.empty() && f.blobs.back() < 0) f.blobs.pop_back();
    if (o.box_samples == 0 || f.blobs.size() <= o.number) return true;
    bool done = true;
    int32_t ubl = -1, best = -1;
    size_t samples = 0, score = 0;
    uint16_t ux = 0, uy = 0;
    bool fresh = true;
    for (;;) {
        if (fresh) {
            ubl = best = -1;
            samples = 0;
            fresh = false;
        }
        size_t sz = f.blobs.size(), fexp = f.first_expansion;
        bool restart = false;
        for (size_t bb = f.first_dust; bb < fexp; ++bb) {
            const int32_t id = f.blobs[bb];
            if (id < 0) continue;
            Blob &bl = f.pool[id];
            bool decide = false;
            if (id == ubl) {
                decide = true;  // the scan came round to the dust blob again
            } else if (bl.unified) {
                // done with this one: swap it to the front
                const size_t fd = f.first_dust;
                if (bb != fd) {
                    const int32_t fdb = f.blobs[fd];
                    if (fdb >= 0) f.pool[fdb].index = bb;
                    bl.index = fd;
                    f.blobs[fd] = id;
                    f.blobs[bb] = fdb;
                }
                f.first_dust++;
                restart = true;
                break;
            } else if (samples >= o.box_samples) {
                decide = true;
            }
            if (decide) {
                samples = 0;
                Blob &u = f.pool[ubl];
                if (best < 0) {  // nothing to merge with
                    if (done) u.unified = true;
                    bb = u.index;
                    ubl = -1;
                    continue;
                }
                Blob &bm = f.pool[best];
                const size_t bi = bm.index;
                merge_means(u, bm);
                f.take_pixels(ubl, best);
                std::set<int32_t>().swap(u.border);
                f.blobs[bi] = -1;
                f.kill(best);
                best = -1;
                if (sz == fexp) {
                    if (f.blobs.back() >= 0) {  // fill the hole with the last blob
                        const int32_t last = f.blobs.back();
                        const size_t last_index = f.pool[last].index;
                        f.blobs[bi] = last;
                        f.pool[last].index = bi;
                        f.blobs[last_index] = -1;
                    }
                    while (!f.blobs.empty() && f.blobs.back() < 0) {
                        f.blobs.pop_back();
                        sz--;
                        fexp--;
                        f.first_expansion--;
                    }
                }
                done = false;
                if (sz <= o.number) {
                    f.pool[ubl].unified = true;
                    return false;
                }
                bb = f.pool[ubl].index;
                ubl = -1;
                continue;
            }
            if (ubl < 0) {
                ubl = id;
                ux = uint16_t(std::round(bl.x));
                uy = uint16_t(std::round(bl.y));
                continue;
            }
            samples++;
            const uint16_t vx = uint16_t(std::round(bl.x)), vy = uint16_t(std::round(bl.y));
            const size_t x0 = std::min(ux, vx), x1 = std::max(ux, vx);
            const size_t y0 = std::min(uy, vy), y1 = std::max(uy, vy);
            if (x0 + o.box_grip >= x1 && y0 + o.box_grip >= y1) {
                const size_t dx = x1 - x0, dy = y1 - y0, cur = dx * dx + dy * dy;
                if (best < 0 || cur < score) {
                    best = id;
                    score = cur;
                    continue;
                }
            }
        }
        if (restart) {
            fresh = true;
            continue;
        }
        if (samples > 0 && ubl >= 0) continue;  // not enough samples yet: scan again
        return done;
    }
}

}  // namespace

bool needed(const options &o) {
    return o.threshold < 1.0 || o.max_size != SIZE_MAX || o.min_size > 1 || o.number > 1;
}

bool segment(const std::vector<const uint32_t *> &stored, int W, int H, const options &o,
             std::default_random_engine &e1, std::vector<frame_result> &out) {
    out.clear();
    if (W <= 0 || H <= 0 || int64_t(W) * int64_t(H) > int64_t(INT32_MAX)) return false;
    std::vector<Frame> F(stored.size());
    for (size_t k = 0; k < F.size(); ++k) {
        F[k].W = W;
        F[k].H = H;
        F[k].px = stored[k];
    }
    // detection: one step of every frame per worker step (thread.cpp:414-423)
    for (bool done = false; !done;) {
        done = true;
        for (Frame &f : F) done &= blobify_step(f, o, e1);
    }
    for (Frame &f : F)
        while (!unify_frame(f, o)) {
        }
    out.assign(F.size(), {});
    const int32_t npx = W * H;
    for (size_t k = 0; k < F.size(); ++k) {
        Frame &f = F[k];
        frame_result &r = out[k];
        r.bop.assign(npx, -1);
        std::vector<int32_t> ordinal(f.pool.size(), -1);
        for (const int32_t id : f.blobs) {
            if (id < 0) continue;
            const Blob &bl = f.pool[id];
            ordinal[id] = int32_t(r.blobs.size());
            r.blobs.push_back({bl.x, bl.y, bl.r, bl.g, bl.b, bl.a, bl.group, bl.size});
        }
        if (f.parent.empty()) continue;  // a frame without pixels
        for (int32_t i = 0; i < npx; ++i)
            if (f.parent[i] >= 0) r.bop[i] = ordinal[f.owner(i)];
    }
    return true;
}

}  // namespace am_blobify

extern "C" int am_blobify_frames(const uint32_t *stored, int32_t nframes, int32_t W, int32_t H, double threshold,
                                 uint64_t max_size, uint64_t min_size, uint64_t number, uint64_t box_samples,
                                 uint32_t box_grip, unsigned seed, int32_t *bop, int32_t *nblobs, double *means,
                                 uint64_t *groups, uint64_t *sizes) {
    if (!stored || nframes < 0 || W <= 0 || H <= 0 || W > 65535 || H > 65535 || !bop || !nblobs || !means ||
        !groups || !sizes || box_grip > UINT16_MAX)
        return -1;
    am_blobify::options o;
    o.threshold = threshold;
    o.max_size = size_t(max_size);
    o.min_size = size_t(min_size);
    o.number = size_t(number);
    o.box_samples = size_t(box_samples);
    o.box_grip = uint16_t(box_grip);
    const size_t npx = size_t(W) * size_t(H);
    std::vector<const uint32_t *> frames(nframes);
    for (int32_t k = 0; k < nframes; ++k) frames[k] = stored + size_t(k) * npx;
    std::vector<am_blobify::frame_result> res;
    std::default_random_engine e1(seed);
    if (!am_blobify::segment(frames, W, H, o, e1, res)) return -1;
    for (int32_t k = 0; k < nframes; ++k) {
        const auto &r = res[k];
        std::copy(r.bop.begin(), r.bop.end(), bop + size_t(k) * npx);
        nblobs[k] = int32_t(r.blobs.size());
        for (size_t j = 0; j < r.blobs.size(); ++j) {
            const auto &bi = r.blobs[j];
            double *m = means + (size_t(k) * npx + j) * 6;
            m[0] = bi.x; m[1] = bi.y; m[2] = bi.r; m[3] = bi.g; m[4] = bi.b; m[5] = bi.a;
            groups[size_t(k) * npx + j] = bi.group;
            sizes[size_t(k) * npx + j] = bi.size;
        }
    }
    return 0;
}
