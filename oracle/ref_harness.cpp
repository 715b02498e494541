/*
 * C-ABI harness around the compiled atomorph reference, for writing golden
 * fixtures and cross-checking the oracle in the build container.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) from the
 * reference's own sources into oracle/_ref/libamref.so; never part of the
 * product.  The reference keeps chains and frames private, so the standard
 * headers are included first and the reference headers are then read with
 * `private` mapped to `public` (layout is unchanged: only access control).
 */
#include <algorithm>
#include <array>
#include <atomic>
#include <cassert>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <future>
#include <iostream>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <random>
#include <set>
#include <sstream>
#include <thread>
#include <vector>

#define private public
#include "atomorph.h"
#undef private

namespace am {
void morph_asynch(point **map, size_t width, size_t height, size_t repeat, unsigned rng_seed, double *gain);
}

namespace {
inline uint64_t word_of(const am::point &p) {
    return uint64_t(p.s.x) | (uint64_t(p.s.y) << 16) | (uint64_t(p.s.x_fract) << 32) |
           (uint64_t(p.s.y_fract) << 40) | (uint64_t(p.s.flags) << 48);
}
inline am::point point_of(uint64_t w) {
    am::point p;
    p.word = 0;
    p.s.x = uint16_t(w & 0xFFFF);
    p.s.y = uint16_t((w >> 16) & 0xFFFF);
    p.s.x_fract = uint8_t((w >> 32) & 0xFF);
    p.s.y_fract = uint8_t((w >> 40) & 0xFF);
    p.s.flags = uint8_t((w >> 48) & 0xFF);
    return p;
}
inline uint32_t pack(const am::color &c) {
    return uint32_t(c.r) | (uint32_t(c.g) << 8) | (uint32_t(c.b) << 16) | (uint32_t(c.a) << 24);
}
inline am::color unpack(uint32_t v) {
    am::color c;
    c.r = v & 255; c.g = (v >> 8) & 255; c.b = (v >> 16) & 255; c.a = v >> 24;
    return c;
}
}  // namespace

extern "C" {

uint32_t ref_rgb_to_hsp(uint32_t c) { return pack(am::rgb_to_hsp(unpack(c))); }
uint32_t ref_hsp_to_rgb(uint32_t c) { return pack(am::hsp_to_rgb(unpack(c))); }

void ref_perlin_perm(unsigned seed, int32_t *p512) {
    PerlinNoise pn(seed);
    for (int i = 0; i < 512; ++i) p512[i] = pn.p[i];
}
double ref_perlin_octave2(unsigned seed, double x, double y, int octaves) {
    PerlinNoise pn(seed);
    return pn.octaveNoise(x, y, octaves);
}
double ref_perlin_noise3(unsigned seed, double x, double y, double z) {
    PerlinNoise pn(seed);
    return pn.noise(x, y, z);
}

void ref_spline_eval(const double *vx, const double *vy, int n, double t, double *ox, double *oy) {
    glnemo::CRSpline s;
    for (int i = 0; i < n; ++i) s.AddSplinePoint(glnemo::Vec3D(vx[i], vy[i], 0.0));
    glnemo::Vec3D v = s.GetInterpolatedSplinePoint(t);
    *ox = v.x;
    *oy = v.y;
}

uint64_t ref_point_distance(uint64_t a, uint64_t b) { return am::point_distance(point_of(a), point_of(b)); }

/* Runs am::morph_asynch on an atom-major buffer pts[x*h + y]. */
double ref_morph_asynch(uint64_t *pts, size_t w, size_t h, size_t repeat, unsigned seed) {
    std::vector<am::point *> rows(w);
    std::vector<am::point> store(w * h);
    for (size_t x = 0; x < w; ++x) {
        rows[x] = &store[x * h];
        for (size_t y = 0; y < h; ++y) store[x * h + y] = point_of(pts[x * h + y]);
    }
    double gain = 0.0;
    am::morph_asynch(rows.data(), w, h, repeat, seed, &gain);
    for (size_t i = 0; i < w * h; ++i) pts[i] = word_of(store[i]);
    return gain;
}

/* One FluidModel::step on particles given as 20 doubles each (layout in
 * oracle/oracle.h). */
void ref_fluid_step(double *P, int count, int gsx, int gsy, int steps_left, double freedom_radius,
                    unsigned jitter_seed) {
    FluidModel fm(gsx, gsy, count);
    Particle *ps = fm.getParticles();
    for (int i = 0; i < gsx; ++i)
        for (int j = 0; j < gsy; ++j) {
            fm.grid[i][j].clear();
        }
    for (int i = 0; i < count; ++i) {
        const double *q = P + 20 * size_t(i);
        Particle *p = &ps[i];
        p->clear();
        p->x = q[0]; p->y = q[1]; p->u = q[2]; p->v = q[3];
        p->gravity_x = q[4]; p->gravity_y = q[5]; p->freedom_r = q[6];
        p->R = q[7]; p->G = q[8]; p->B = q[9]; p->A = q[10];
        p->r = q[11]; p->g = q[12]; p->b = q[13]; p->a = q[14];
        p->strength = q[15];
        p->active = q[16] != 0.0;
        p->mature = q[17] != 0.0;
    }
    if (jitter_seed) srand(jitter_seed);
    fm.step(size_t(steps_left), freedom_radius, 0.0);
    for (int i = 0; i < count; ++i) {
        double *q = P + 20 * size_t(i);
        Particle *p = &ps[i];
        q[0] = p->x; q[1] = p->y; q[2] = p->u; q[3] = p->v;
        q[11] = p->r; q[12] = p->g; q[13] = p->b; q[14] = p->a;
    }
}

/* ---------------------------------------------------------------- morph */
void *ref_morph_new(void) { return new am::morph(); }
void ref_morph_free(void *m) { delete static_cast<am::morph *>(m); }

void ref_morph_set_option(void *mp, const char *name, double v) {
    am::morph *m = static_cast<am::morph *>(mp);
    std::string n(name);
    if (n == "seed") m->set_seed(unsigned(v));
    else if (n == "threads") m->set_threads(size_t(v));
    else if (n == "cycle_length") m->set_cycle_length(size_t(v));
    else if (n == "motion") m->set_motion((unsigned char)v);
    else if (n == "fading") m->set_fading((unsigned char)v);
    else if (n == "density") m->set_density(uint16_t(v));
    else if (n == "fluid") m->set_fluid(unsigned(v));
    else if (n == "feather") m->set_feather(size_t(v));
    else if (n == "keep_background") m->set_keep_background(v != 0.0);
    else if (n == "finite") m->set_finite(v != 0.0);
    else if (n == "show_blobs") m->set_show_blobs(unsigned(v));
    else if (n == "blob_delimiter") m->set_blob_delimiter((unsigned char)v);
    else if (n == "degeneration") m->set_degeneration(size_t(v));
    else if (n == "blob_threshold") m->set_blob_threshold(v);
    else if (n == "blob_max_size") m->set_blob_max_size(size_t(v));
    else if (n == "blob_min_size") m->set_blob_min_size(size_t(v));
    else if (n == "blob_box_grip") m->set_blob_box_grip(uint16_t(v));
    else if (n == "blob_box_samples") m->set_blob_box_samples(size_t(v));
    else if (n == "blob_number") m->set_blob_number(size_t(v));
    else if (n == "blob_rgba_weight") m->set_blob_rgba_weight((unsigned char)v);
    else if (n == "blob_size_weight") m->set_blob_size_weight((unsigned char)v);
    else if (n == "blob_xy_weight") m->set_blob_xy_weight((unsigned char)v);
}

/* demo/main.cpp:96-127 (fill_morph): skips alpha == 0 pixels. */
void ref_morph_add_rgba(void *mp, size_t frame, const uint8_t *rgba, unsigned w, unsigned h) {
    am::morph *m = static_cast<am::morph *>(mp);
    bool empty = true;
    for (size_t i = 0; i < size_t(w) * h; ++i) {
        const uint8_t *q = rgba + 4 * i;
        if (q[3] == 0) continue;
        m->add_pixel(frame, am::create_pixel(uint16_t(i % w), uint16_t(i / w), q[0], q[1], q[2], q[3]));
        empty = false;
    }
    if (empty) m->add_frame(frame);
}
void ref_morph_set_resolution(void *mp, unsigned w, unsigned h) {
    static_cast<am::morph *>(mp)->set_resolution(uint16_t(w), uint16_t(h));
}
/* n worker steps, run synchronously on the calling thread (thread::run does
 * nothing else per step, thread.cpp:173-206), each side synchronized. */
unsigned ref_morph_run(void *mp, size_t n) {
    am::morph *m = static_cast<am::morph *>(mp);
    m->synchronize();
    if (n > 0) {
        for (size_t i = 0; i < n; ++i) m->worker.step();
        m->synchronize();
    }
    return m->get_state();
}
/* Worker steps until its state reaches `state` (or max_steps were taken),
 * then synchronize; returns the state. */
unsigned ref_morph_run_until(void *mp, unsigned state, size_t max_steps) {
    am::morph *m = static_cast<am::morph *>(mp);
    m->synchronize();
    for (size_t i = 0; i < max_steps && m->worker.get_state() < state; ++i) m->worker.step();
    m->synchronize();
    return m->get_state();
}
void ref_morph_next_state(void *mp) {
    am::morph *m = static_cast<am::morph *>(mp);
    m->next_state();
}
unsigned ref_morph_state(void *mp) { return static_cast<am::morph *>(mp)->get_state(); }
double ref_morph_energy(void *mp) { return static_cast<am::morph *>(mp)->get_energy(); }

size_t ref_morph_frame_count(void *mp) { return static_cast<am::morph *>(mp)->frames.size(); }
size_t ref_morph_blob_count(void *mp, size_t f) { return static_cast<am::morph *>(mp)->get_blob_count(f); }
/* Blob b of frame f: returns surface size; fills group, means[6] = x,y,r,g,b,a
 * and (if surf != NULL) the surface positions (y*65536+x, ascending). */
size_t ref_morph_blob(void *mp, size_t f, size_t b, size_t *group, double *means, uint64_t *surf) {
    am::morph *m = static_cast<am::morph *>(mp);
    const am::blob *bl = m->get_blob(f, b);
    if (!bl) return SIZE_MAX;
    if (group) *group = bl->group;
    if (means) {
        means[0] = bl->x; means[1] = bl->y; means[2] = bl->r;
        means[3] = bl->g; means[4] = bl->b; means[5] = bl->a;
    }
    if (surf) {
        size_t i = 0;
        for (size_t p : bl->surface) surf[i++] = p;
    }
    return bl->surface.size();
}
/* The worker's random engine (std::default_random_engine, one state word),
 * e.g. where blob matching starts drawing from it. */
uint64_t ref_morph_rng_state(void *mp) {
    am::morph *m = static_cast<am::morph *>(mp);
    std::ostringstream os;
    os << m->worker.e1;
    return std::stoull(os.str());
}
/* Frame f's running means x, y, r, g, b, a as morph::add_pixel keeps them
 * (morph.cpp:317-334). */
void ref_morph_frame_means(void *mp, size_t f, double *m6) {
    am::morph *m = static_cast<am::morph *>(mp);
    const am::frame &fr = m->frames[f];
    m6[0] = fr.x; m6[1] = fr.y; m6[2] = fr.r; m6[3] = fr.g; m6[4] = fr.b; m6[5] = fr.a;
}
/* Blob matching energies of the worker: the current map's running value,
 * the best seen (infinity before the first accepted swap) and the value
 * get_energy() reports while matching. */
void ref_morph_match_energy(void *mp, double *e3) {
    am::morph *m = static_cast<am::morph *>(mp);
    e3[0] = m->worker.blob_map_e;
    e3[1] = m->worker.best_e;
    e3[2] = m->worker.best_blob_map_e;
}
namespace {
void dump_particles(const Particle *ps, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i) {
        const Particle *p = &ps[i];
        double *q = out + 20 * i;
        q[0] = p->x; q[1] = p->y; q[2] = p->u; q[3] = p->v;
        q[4] = p->gravity_x; q[5] = p->gravity_y; q[6] = p->freedom_r;
        q[7] = p->R; q[8] = p->G; q[9] = p->B; q[10] = p->A;
        q[11] = p->r; q[12] = p->g; q[13] = p->b; q[14] = p->a;
        q[15] = p->strength; q[16] = p->active ? 1.0 : 0.0; q[17] = p->mature ? 1.0 : 0.0;
        q[18] = 0.0; q[19] = 0.0;
    }
}
}  // namespace

/* Number of particles of the morph's fluid model (0: none yet). */
size_t ref_morph_particle_count(void *mp) {
    am::morph *m = static_cast<am::morph *>(mp);
    return m->fluid ? m->fluid->get_particle_count() : 0;
}
/* One fluid frame at global time `time`, split at the boundary between
 * step_fluid and the drawing half of morph::draw_fluid (morph.cpp:1057-1300):
 * from the current state step_fluid runs and its particles are written to
 * P_pre; the particles and the random engine are then put back and
 * get_pixels(time) renders the frame the normal way (step_fluid + drawing)
 * into img; P_post receives the particles after it.  rand() is reseeded
 * with srand_seed before each step_fluid.  The drawn frame was therefore
 * made from P_pre when the second step_fluid reproduced it: returns 1 when
 * the positions after get_pixels equal P_pre's, 0 if not, -1 without a
 * fluid model. */
int ref_fluid_frame(void *mp, double time, unsigned srand_seed, double *P_pre, uint8_t *img, double *P_post) {
    am::morph *m = static_cast<am::morph *>(mp);
    if (!m->fluid) return -1;
    FluidModel *fm = m->fluid;
    const size_t n = fm->get_particle_count();
    Particle *ps = fm->getParticles();
    std::vector<Particle> saved(ps, ps + n);
    const std::default_random_engine e_saved = m->e1;
    const double tn = m->normalize_time(time);
    const size_t fk = m->get_frame_key(tn);
    const double dt = 1.0 / double(m->frames.size());
    const double tl = std::max(0.0, (tn - (fk * dt)) / dt);
    srand(srand_seed);
    m->step_fluid(fk, tl, tn);
    dump_particles(ps, n, P_pre);
    std::copy(saved.begin(), saved.end(), ps);
    m->e1 = e_saved;
    srand(srand_seed);
    std::vector<am::pixel> px;
    m->get_pixels(time, &px);
    const unsigned W = m->get_width(), H = m->get_height();
    std::memset(img, 0, size_t(W) * H * 4);
    for (const am::pixel &q : px) {
        const size_t pos = (size_t(q.y) * W + q.x) * 4;
        if (pos >= size_t(W) * H * 4) continue;
        img[pos + 0] = q.c.r; img[pos + 1] = q.c.g; img[pos + 2] = q.c.b; img[pos + 3] = q.c.a;
    }
    dump_particles(ps, n, P_post);
    for (size_t i = 0; i < n; ++i)
        if (P_post[20 * i] != P_pre[20 * i] || P_post[20 * i + 1] != P_pre[20 * i + 1]) return 0;
    return 1;
}
/* get_background(x, y, time) for every pixel (morph.cpp:1431-1465). */
void ref_morph_background(void *mp, double time, uint8_t *img) {
    am::morph *m = static_cast<am::morph *>(mp);
    const unsigned W = m->get_width(), H = m->get_height();
    for (unsigned y = 0; y < H; ++y)
        for (unsigned x = 0; x < W; ++x) {
            const am::color c = m->get_background(uint16_t(x), uint16_t(y), time);
            uint8_t *o = img + (size_t(y) * W + x) * 4;
            o[0] = c.r; o[1] = c.g; o[2] = c.b; o[3] = c.a;
        }
}
size_t ref_morph_chain_count(void *mp) { return static_cast<am::morph *>(mp)->chains.size(); }
/* Chain with key `key`: returns width, writes height and max_surface; if pts
 * != NULL copies atom-major points (padding byte cleared). */
size_t ref_morph_chain(void *mp, size_t key, size_t *height, size_t *max_surface, uint64_t *pts) {
    am::morph *m = static_cast<am::morph *>(mp);
    auto it = m->chains.find(key);
    if (it == m->chains.end()) return SIZE_MAX;
    const am::chain &c = it->second;
    if (height) *height = c.height;
    if (max_surface) *max_surface = c.max_surface;
    if (pts)
        for (size_t x = 0; x < c.width; ++x)
            for (size_t y = 0; y < c.height; ++y) pts[x * c.height + y] = word_of(c.points[x][y]);
    return c.width;
}
/* Overwrite a chain's points (and rebuild its splines) on the morph side
 * only, to render from given trajectories. */
void ref_morph_put_chain(void *mp, size_t key, const uint64_t *pts) {
    am::morph *m = static_cast<am::morph *>(mp);
    am::chain &c = m->chains[key];
    for (size_t x = 0; x < c.width; ++x) {
        c.splines[x].clearCPoints();
        for (size_t y = 0; y < c.height; ++y) {
            c.points[x][y] = point_of(pts[x * c.height + y]);
            double vx = c.points[x][y].s.x + c.points[x][y].s.x_fract / double(UINT8_MAX + 1);
            double vy = c.points[x][y].s.y + c.points[x][y].s.y_fract / double(UINT8_MAX + 1);
            c.splines[x].AddSplinePoint(glnemo::Vec3D(vx, vy, 0.0));
        }
    }
}
double ref_morph_get_time(void *mp, size_t i, size_t n) { return static_cast<am::morph *>(mp)->get_time(i, n); }
/* morph::get_pixels(t) into a width*height RGBA8 image. */
void ref_morph_render(void *mp, double t, uint8_t *out) {
    am::morph *m = static_cast<am::morph *>(mp);
    std::vector<am::pixel> px;
    m->get_pixels(t, &px);
    unsigned W = m->get_width(), H = m->get_height();
    std::memset(out, 0, size_t(W) * H * 4);
    for (const am::pixel &p : px) {
        size_t pos = (size_t(p.y) * W + p.x) * 4;
        if (pos >= size_t(W) * H * 4) continue;
        out[pos + 0] = p.c.r; out[pos + 1] = p.c.g; out[pos + 2] = p.c.b; out[pos + 3] = p.c.a;
    }
}
void ref_morph_bbox(void *mp, int32_t *b4) {
    am::morph *m = static_cast<am::morph *>(mp);
    b4[0] = m->bbox_x1; b4[1] = m->bbox_y1; b4[2] = m->bbox_x2; b4[3] = m->bbox_y2;
}
/* morph's stored pixel of frame f at (x,y) as get_pixel returns it. */
uint32_t ref_morph_get_pixel(void *mp, size_t f, unsigned x, unsigned y) {
    am::morph *m = static_cast<am::morph *>(mp);
    return pack(m->get_pixel(f, am::xy2pos(uint16_t(x), uint16_t(y))).c);
}

}  /* extern "C" */
