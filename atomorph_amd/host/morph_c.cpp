/*
 * C shim over am::morph for hosts without C++ linkage (Python ctypes in the
 * tests and the benchmark, or any other FFI).  Each function forwards to the
 * am::morph member of the same name.
 */
#include <algorithm>
#include <cstring>
#include <string>

#include "atomorph.h"

extern "C" {

void *am_morph_new(void) { return new (std::nothrow) am::morph(); }
void am_morph_free(void *m) { delete static_cast<am::morph *>(m); }

/* The option setters of morph.h, by name.  Returns 0, or -1 for an unknown
 * name. */
int am_morph_set(void *mp, const char *name, double v) {
    am::morph *m = static_cast<am::morph *>(mp);
    const std::string n(name);
    if (n == "seed") m->set_seed(unsigned(v));
    else if (n == "blob_delimiter") m->set_blob_delimiter((unsigned char)v);
    else if (n == "blob_threshold") m->set_blob_threshold(v);
    else if (n == "blob_max_size") m->set_blob_max_size(size_t(v));
    else if (n == "blob_min_size") m->set_blob_min_size(size_t(v));
    else if (n == "blob_box_grip") m->set_blob_box_grip(uint16_t(v));
    else if (n == "blob_box_samples") m->set_blob_box_samples(size_t(v));
    else if (n == "blob_number") m->set_blob_number(size_t(v));
    else if (n == "blob_rgba_weight") m->set_blob_rgba_weight((unsigned char)v);
    else if (n == "blob_size_weight") m->set_blob_size_weight((unsigned char)v);
    else if (n == "blob_xy_weight") m->set_blob_xy_weight((unsigned char)v);
    else if (n == "degeneration") m->set_degeneration(size_t(v));
    else if (n == "motion") m->set_motion((unsigned char)v);
    else if (n == "fading") m->set_fading((unsigned char)v);
    else if (n == "threads") m->set_threads(size_t(v));
    else if (n == "cycle_length") m->set_cycle_length(size_t(v));
    else if (n == "feather") m->set_feather(size_t(v));
    else if (n == "keep_background") m->set_keep_background(v != 0.0);
    else if (n == "finite") m->set_finite(v != 0.0);
    else if (n == "show_blobs") m->set_show_blobs(unsigned(v));
    else if (n == "fluid") m->set_fluid(unsigned(v));
    else if (n == "density") m->set_density(uint16_t(v));
    else if (n == "match_budget") m->set_match_budget(size_t(v));
    else if (n == "fluid_sparseness") m->set_fluid_sparseness(v);
    else return -1;
    return 0;
}
void am_morph_set_refine(void *m, unsigned mode, unsigned rounds, int tile, unsigned graph_passes) {
    static_cast<am::morph *>(m)->set_refine(mode, rounds, tile, graph_passes);
}
int am_morph_add_frame_rgba(void *m, size_t frame, const uint8_t *rgba, unsigned w, unsigned h, size_t stride) {
    return static_cast<am::morph *>(m)->add_frame_rgba(frame, rgba, uint16_t(w), uint16_t(h), stride) ? 0 : -1;
}
int am_morph_add_frame(void *m, size_t frame) { return static_cast<am::morph *>(m)->add_frame(frame) ? 0 : -1; }
void am_morph_set_resolution(void *m, unsigned w, unsigned h) {
    static_cast<am::morph *>(m)->set_resolution(uint16_t(w), uint16_t(h));
}
/* iterate(n), wait until the n steps are done, synchronize: the blocking
 * form the tests and benchmark drive.  Returns the state, or -1 on error. */
int am_morph_run(void *mp, size_t n) {
    am::morph *m = static_cast<am::morph *>(mp);
    m->suspend();
    if (!m->synchronize()) return -1;
    if (n) {
        m->iterate(n);
        m->wait();
        if (!m->synchronize()) return -1;
    }
    return int(m->get_state());
}
void am_morph_wait(void *m) { static_cast<am::morph *>(m)->wait(); }
void am_morph_restart(void *m) { static_cast<am::morph *>(m)->restart(); }
/* n steps timed with HIP events on the compute stream, then synchronize.
 * Returns the state or -1. */
int am_morph_run_timed(void *mp, size_t n, float *gpu_ms) {
    am::morph *m = static_cast<am::morph *>(mp);
    m->suspend();
    if (!m->synchronize()) return -1;
    if (!m->iterate_timed(n, gpu_ms)) return -1;
    if (!m->synchronize()) return -1;
    return int(m->get_state());
}
int am_morph_render_device(void *m, const double *times, size_t n, void *out_dev, float *gpu_ms) {
    return static_cast<am::morph *>(m)->render_device(times, n, out_dev, gpu_ms) ? 0 : -1;
}
void am_morph_compute(void *m) { static_cast<am::morph *>(m)->compute(); }
void am_morph_compute_seconds(void *m, double s) { static_cast<am::morph *>(m)->compute(s); }
void am_morph_iterate(void *m, size_t n) { static_cast<am::morph *>(m)->iterate(n); }
void am_morph_suspend(void *m) { static_cast<am::morph *>(m)->suspend(); }
int am_morph_suspend_timeout(void *m, double s) { return static_cast<am::morph *>(m)->suspend(s) ? 1 : 0; }
int am_morph_is_busy(void *m) { return static_cast<am::morph *>(m)->is_busy() ? 1 : 0; }
int am_morph_synchronize(void *m) { return static_cast<am::morph *>(m)->synchronize() ? 1 : 0; }
void am_morph_next_state(void *m) { static_cast<am::morph *>(m)->next_state(); }
unsigned am_morph_state(void *m) { return static_cast<am::morph *>(m)->get_state(); }
double am_morph_energy(void *m) { return static_cast<am::morph *>(m)->get_energy(); }
int am_morph_cost(void *m, uint64_t *lo, uint64_t *hi) {
    return static_cast<am::morph *>(m)->transport_cost(lo, hi) ? 0 : -1;
}
void am_morph_counters(void *m, uint64_t *attempts, uint64_t *accepted) {
    static_cast<am::morph *>(m)->get_refine_counters(attempts, accepted);
}
double am_morph_get_time(void *m, size_t i, size_t n) { return static_cast<am::morph *>(m)->get_time(i, n); }
int am_morph_render(void *m, double t, uint8_t *out) {
    return static_cast<am::morph *>(m)->get_frame_rgba(t, out) ? 0 : -1;
}
/* get_pixels(t, &vector) into width*height RGBA8 through the pixel vector
 * path (for testing that path against the bulk one). */
int am_morph_render_pixels(void *mp, double t, uint8_t *out, unsigned w, unsigned h) {
    std::vector<am::pixel> px;
    static_cast<am::morph *>(mp)->get_pixels(t, &px);
    std::memset(out, 0, size_t(w) * h * 4);
    for (const am::pixel &p : px) {
        if (p.x >= w || p.y >= h) continue;
        uint8_t *q = out + (size_t(p.y) * w + p.x) * 4;
        q[0] = p.c.r; q[1] = p.c.g; q[2] = p.c.b; q[3] = p.c.a;
    }
    return int(px.size());
}
size_t am_morph_blob_count(void *m, size_t f) { return static_cast<am::morph *>(m)->get_blob_count(f); }
/* Blob b of frame f: surface size, group, and (if surf) its positions. */
long long am_morph_blob(void *mp, size_t f, size_t b, size_t *group, double *means, uint64_t *surf) {
    const am::blob *bl = static_cast<am::morph *>(mp)->get_blob(f, b);
    if (!bl) return -1;
    if (group) *group = bl->group;
    if (means) {
        means[0] = bl->x; means[1] = bl->y; means[2] = bl->r;
        means[3] = bl->g; means[4] = bl->b; means[5] = bl->a;
    }
    if (surf) {
        size_t i = 0;
        for (size_t p : bl->surface) surf[i++] = p;
    }
    return (long long)bl->surface.size();
}
/* Chains: returns total atoms; writes height; if widths / pts non-NULL fills
 * them (pts key-frame-major [h][atoms]). */
size_t am_morph_chains(void *mp, size_t *height, int64_t *widths, uint64_t *pts) {
    am::morph *m = static_cast<am::morph *>(mp);
    std::vector<uint64_t> p;
    std::vector<size_t> w;
    size_t hh = 0;
    const size_t n = m->get_chain_points(pts ? &p : nullptr, &hh, &w);
    if (height) *height = hh;
    if (widths)
        for (size_t i = 0; i < w.size(); ++i) widths[i] = int64_t(w[i]);
    if (pts && !p.empty()) std::memcpy(pts, p.data(), p.size() * 8);
    return n;
}
size_t am_morph_chain_count(void *mp) {
    std::vector<size_t> w;
    size_t hh = 0;
    static_cast<am::morph *>(mp)->get_chain_points(nullptr, &hh, &w);
    return w.size();
}
int am_morph_put_chains(void *mp, const uint64_t *pts, size_t n) {
    std::vector<uint64_t> p(pts, pts + n);
    return static_cast<am::morph *>(mp)->put_chain_points(p) ? 0 : -1;
}
// report: max_report rows of (tile, rounds, passes, ms, gain_at_budget);
// *rounds > 0 on input: only that rounds value is raced
int am_morph_autotune(void *mp, double budget_ms, int *tile, unsigned *rounds, double *report, size_t max_report,
                      size_t *nreport) {
    std::vector<am::tune_entry> rep;
    if (!static_cast<am::morph *>(mp)->autotune(budget_ms, tile, rounds, &rep)) return -1;
    for (size_t i = 0; i < rep.size() && i < max_report; ++i) {
        double *r = report + 5 * i;
        r[0] = rep[i].tile;
        r[1] = rep[i].rounds;
        r[2] = rep[i].passes;
        r[3] = rep[i].ms;
        r[4] = rep[i].gain_at_budget;
    }
    if (nreport) *nreport = rep.size();
    return 0;
}
int am_morph_save_chains(void *m, const char *path) { return static_cast<am::morph *>(m)->save_chains(path) ? 0 : -1; }
int am_morph_load_chains(void *m, const char *path) { return static_cast<am::morph *>(m)->load_chains(path) ? 0 : -1; }
void am_morph_set_learned(void *m, const float *params) { static_cast<am::morph *>(m)->set_learned(params); }
int am_morph_warm_start(void *m, float *gpu_ms) { return static_cast<am::morph *>(m)->warm_start(gpu_ms) ? 0 : -1; }
size_t am_morph_learn_samples(void *mp, float *x, float *t) {
    am::morph *m = static_cast<am::morph *>(mp);
    if (!x && !t) return m->learn_samples(nullptr, nullptr);
    std::vector<float> vx, vt;
    const size_t n = m->learn_samples(x ? &vx : nullptr, t ? &vt : nullptr);
    if (x && vx.size() == n * 8) std::copy(vx.begin(), vx.end(), x);
    if (t && vt.size() == n * 2) std::copy(vt.begin(), vt.end(), t);
    return n;
}
size_t am_morph_fluid_particles(void *mp, double *out) {
    am::morph *m = static_cast<am::morph *>(mp);
    std::vector<double> v;
    const size_t n = m->get_fluid_particles(out ? &v : nullptr);
    if (out && v.size() == n * 20) std::copy(v.begin(), v.end(), out);
    return n;
}
const char *am_morph_error(void *m) { return static_cast<am::morph *>(m)->last_error(); }

/* Host-only members (no device work), for tests of the host logic. */
size_t am_morph_frame_key(void *m, double t) { return static_cast<am::morph *>(m)->get_frame_key(t); }
double am_morph_normalize_time(void *m, double t) { return static_cast<am::morph *>(m)->normalize_time(t); }
uint32_t am_morph_get_pixel(void *m, size_t f, size_t pos) {
    return am::color_word(static_cast<am::morph *>(m)->get_pixel(f, pos).c);
}
int am_morph_add_pixel(void *m, size_t f, unsigned x, unsigned y, uint32_t rgba) {
    return static_cast<am::morph *>(m)->add_pixel(f, am::create_pixel(uint16_t(x), uint16_t(y), am::word_color(rgba)))
               ? 0
               : -1;
}
uint32_t am_morph_interpolate_color(void *m, uint32_t c1, uint32_t c2, double lag, double slope, double w) {
    return am::color_word(static_cast<am::morph *>(m)->interpolate(am::word_color(c1), am::word_color(c2), lag, slope, w));
}
uint64_t am_morph_interpolate_point(void *m, uint64_t p1, uint64_t p2, double w) {
    am::point a, b;
    a.word = p1;
    b.word = p2;
    am::point r = static_cast<am::morph *>(m)->interpolate(a, b, w);
    return r.word & 0x0000FFFFFFFFFFFFull;
}
uint32_t am_morph_background(void *m, unsigned x, unsigned y, double t) {
    return am::color_word(static_cast<am::morph *>(m)->get_background(uint16_t(x), uint16_t(y), t));
}
size_t am_morph_frame_count(void *m) { return static_cast<am::morph *>(m)->get_frame_count(); }
size_t am_morph_pixel_count(void *m, size_t f) { return static_cast<am::morph *>(m)->get_pixel_count(f); }

}  // extern "C"
