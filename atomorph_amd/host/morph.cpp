/*
 * am::morph on one MI355X.  The public behaviour follows the reference's
 * morph.cpp / thread.cpp (citations inline); the heavy work is enqueued on
 * HIP streams through the C ABI of libatomorph_hip.so.
 */
#include "atomorph.h"
#include "blobify.h"

#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>

#include "../../include/atomorph_hip.h"

namespace am {

// ------------------------------------------------------ free functions
color create_color(unsigned char r, unsigned char g, unsigned char b, unsigned char a) {
    color c;
    c.r = r; c.g = g; c.b = b; c.a = a;
    return c;
}
static inline uint8_t to_u8(double v) { return uint8_t(uint32_t(int32_t(v)) & 0xFFu); }
static inline uint16_t to_u16(double v) { return uint16_t(uint32_t(int32_t(v)) & 0xFFFFu); }
color create_color(double r, double g, double b, double a) {  // color.cpp:22-29
    color c;
    c.r = to_u8(std::round(r * 255.0));
    c.g = to_u8(std::round(g * 255.0));
    c.b = to_u8(std::round(b * 255.0));
    c.a = to_u8(std::round(a * 255.0));
    return c;
}
color rgb_to_hsp(color c) { return word_color(amh_rgb_to_hsp(color_word(c))); }
color hsp_to_rgb(color c) { return word_color(amh_hsp_to_rgb(color_word(c))); }
void RGBtoHSP(double R, double G, double B, double *H, double *S, double *P) { amh_rgb2hsp(R, G, B, H, S, P); }
void HSPtoRGB(double H, double S, double P, double *R, double *G, double *B) { amh_hsp2rgb(H, S, P, R, G, B); }

pixel create_pixel(uint16_t x, uint16_t y, unsigned char r, unsigned char g, unsigned char b, unsigned char a) {
    pixel px;
    px.x = x; px.y = y;
    px.c = create_color(r, g, b, a);
    return px;
}
pixel create_pixel(uint16_t x, uint16_t y, color c) { return create_pixel(x, y, c.r, c.g, c.b, c.a); }

const char *get_version() { return "1.0"; }
size_t get_warning() {
    size_t flags = 0;
    if (sizeof(void *) < 8) flags |= WARN_POINTER_SIZE;
    if (sizeof(pixel) > 8) flags |= WARN_PIXEL_SIZE;
    if (sizeof(point) > 8) flags |= WARN_POINT_SIZE;
    return flags;
}
bool uses_opencv() { return false; }

void clear_chain(chain *c) {
    if (c->points) {
        for (size_t i = 0; i < c->width; ++i) delete[] c->points[i];
        delete[] c->points;
        c->points = nullptr;
    }
    if (c->places) {
        for (size_t i = 0; i < c->width; ++i) delete[] c->places[i];
        delete[] c->places;
        c->places = nullptr;
    }
    c->width = c->height = 0;
    c->energy = 0.0;
    c->max_surface = 0;
}
bool renew_chain(chain *c, size_t width, size_t height) {
    clear_chain(c);
    c->points = new (std::nothrow) point *[width];
    c->places = new (std::nothrow) size_t *[width];
    if (!c->points || !c->places) {
        delete[] c->points;
        delete[] c->places;
        c->points = nullptr;
        c->places = nullptr;
        return false;
    }
    for (size_t k = 0; k < width; ++k) {
        c->points[k] = new (std::nothrow) point[height];
        c->places[k] = new (std::nothrow) size_t[height];
    }
    c->width = width;
    c->height = height;
    return true;
}

// ------------------------------------------------------------- engine
namespace {

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) amh_free(p);
        p = nullptr;
        n = 0;
    }
    bool ensure(size_t bytes) {
        if (bytes <= n && p) return true;
        release();
        if (amh_malloc(&p, bytes ? bytes : 8) != AMH_OK) return false;
        n = bytes ? bytes : 8;
        return true;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct Options {  // the setters, as the worker sees them (morph.h:97-119)
    unsigned char fading = PERLIN, motion = SPLINE, blob_delimiter = HSP;
    double blob_threshold = 1.0;
    size_t blob_number = 1, blob_max_size = SIZE_MAX, blob_min_size = 1, blob_box_samples = 10;
    uint16_t blob_box_grip = UINT16_MAX;
    unsigned char blob_rgba_weight = 1, blob_size_weight = 1, blob_xy_weight = 1;
    size_t degeneration = 0;
    uint16_t density = 1;
    size_t threads = 0, cycle_length = 1000, feather = 0;
    bool keep_background = false, blend_blobs = false, finite = false;
    unsigned fluidsteps = 0, show_blobs = TEXTURE;
    unsigned seed = 0;
    // batched: rounds 0 or tile < 0 = autotuned for the chains
    unsigned refine_mode = REFINE_PLAIN, refine_rounds = 0, graph_passes = 0;
    int refine_tile = -1;
    size_t match_budget = 0;
    double fluid_sparseness = 1.0;  // grid cells per pixel (the reference's sparseness)
    bool learned = false;           // learned warm start after the chains are built
};

struct HostFrame {  // caller-side pixels of one key frame, in insertion order
    std::vector<std::pair<uint32_t, uint32_t>> px;  // (y*65536+x, RGBA as given)
};

struct BlobInfo {  // one blob of a key frame after detection or padding
    bool empty = false;
    int id = -1;       // blob number from amh_blob_stats
    size_t size = 0;
    double x = 0, y = 0, r = 0, g = 0, b = 0, a = 0;
    int first = -1;    // smallest pixel index (canvas)
    size_t group0 = 0; // group before matching: y*65536 + x of the pixel the
                       // greedy merge grew the blob from (thread.cpp:243),
                       // of its first pixel for device-detected blobs
    pixel px{};        // the blob as one pixel (blob_distance), set by match_init
};

inline double mono_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct morph::engine {
    // ---------------------------------------------------- caller side
    std::map<size_t, HostFrame> frames;
    Options opt;                    // caller's setters
    uint16_t width = 0, height = 0;
    uint16_t bbox_x1 = UINT16_MAX, bbox_y1 = UINT16_MAX, bbox_x2 = 0, bbox_y2 = 0;
    size_t identifier = 0;
    bool skip_state = false;
    unsigned state = STATE_BLOB_DETECTION;
    double energy = 0.0;
    std::string err;
    std::mt19937 e1{0};             // morph's own engine (morph.cpp:9-12)
    int32_t lag_perm[512], slope_perm[512];
    // synchronized results
    std::vector<std::vector<blob *>> out_blobs;  // [frame index] materialised lazily
    bool blobs_dirty = true;
    uint64_t counters[3] = {0, 0, 0};

    // ---------------------------------------------------- worker side
    Options wopt;
    size_t wident = SIZE_MAX;
    unsigned wstate = STATE_BLOB_DETECTION;
    bool wskip = false;
    size_t counter = 0;
    // The worker's random engine, as the reference's (thread.h:120): greedy
    // blob detection and blob matching draw from it in the reference's
    // order, so matching continues where detection left it.
    std::default_random_engine rng;
    void *cstream = nullptr, *rstream = nullptr;
    // canvas and key frames
    int32_t CW = 0, CH = 0, h = 0;
    std::vector<size_t> keys;                     // frame keys in order
    std::vector<std::vector<uint32_t>> canvas;    // host copy of each key frame
    std::vector<double> fmean;                    // [h][6] frame means (stored colour space)
    DevBuf d_rgba, d_stored, d_shown, d_labels, d_bop, d_plist, d_first, d_stats, d_bws, d_gws;
    std::vector<int32_t> nblobs;                  // per key frame
    std::vector<std::vector<amh_blob_stat>> stats;
    // matching
    std::vector<std::vector<BlobInfo>> fb;        // [y][index in frame order]
    std::vector<int32_t> bmap, best_map;          // [c*h + y] -> index in fb[y]
    size_t bw = 0;                                // blob map width
    double map_e = 0, best_e = std::numeric_limits<double>::infinity();
    double best_report = 0;                       // get_energy() while matching (best_blob_map_e)
    bool matched = false;
    size_t match_steps = 0;
    uint32_t bbox_d = 0;
    double w_rgba = 1.0 / 3, w_size = 1.0 / 3, w_xy = 1.0 / 3;
    // chains
    std::vector<int64_t> off;
    std::vector<int32_t> cw;
    int64_t wtot = 0;
    DevBuf d_pts, d_off, d_w, d_items, d_queue, d_rstats, d_cost, d_cblob, d_exy;
    int32_t nitems = 0, items_tile = -1;
    uint64_t visits = 0, passes = 0;
    void *graph = nullptr;
    // atom-path frames replayed from a captured graph, and what it was
    // captured for (the render arguments and the workspace)
    void *rgraph = nullptr;
    std::vector<unsigned char> rgraph_key;
    int graph_tile = -1;
    unsigned graph_np = 0, graph_rounds = 0;
    // autotuned batched schedule for the current chains (-1: not yet)
    int tuned_tile = -1, tuned_rounds = 0, tuned_fixed = 0;
    DevBuf d_tune;
    // render snapshot (made at synchronize)
    DevBuf d_snap, d_rank, d_brgba, d_lag, d_slope, d_rws, d_out;
    DevBuf d_fkeys;  // the render snapshot's key points atom-major, for the fluid update
    std::vector<int32_t> rank;                    // [h][nchains]
    std::vector<uint32_t> brgba;                  // [h][bw]
    bool have_snap = false;
    // fluid path: the model persists across frames; its per-segment inputs
    // are rebuilt after every snapshot
    void *fluid = nullptr;
    double fluid_sp = 0.0;
    bool fluid_inputs = false;
    bool fluid_stale = false;        // restarted: particles back to a new model's state on next use
    int32_t fluid_dims[5] = {0, 0, 0, 0, 0};  // gsx, gsy, particles, width, height of the handle
    DevBuf d_frank, d_fsurf, d_fbcol;
    // learned warm start: MLP parameters (caller side), the handle and its
    // workspace (worker side)
    std::vector<float> mlp_params;
    size_t mlp_version = 0, mlp_loaded = SIZE_MAX;
    void *mlp = nullptr;
    DevBuf d_wws, d_lx, d_lt;
    bool prepare_mlp();
    bool warm_start();
    bool prepare_fluid();
    bool render_fluid(double t, uint32_t *out_dev);
    void fill_render(amh_render &r) const;

    // ---------------------------------------------------- worker thread
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    bool running = false, paused = true, stop = false, pause_req = false;
    size_t iters = 0;
    double secs = 0, t0 = 0;

    std::vector<int64_t> d_first_offsets;         // per frame offset into d_first

    engine();
    ~engine();
    void set_perms(unsigned seed);
    uint64_t mix_seed() const { return (uint64_t(wopt.seed) << 32) ^ 0x2545F4914F6CDD1Dull; }
    bool fail(const char *what);
    bool ensure_streams();
    void reset_worker();
    bool upload_frames();
    bool step();
    bool detect();
    bool match_init();
    bool match_step();
    double blob_distance(const BlobInfo &b1, const BlobInfo &b2) const;
    double map_energy(const std::vector<int32_t> &m) const;
    void fix_volatiles(size_t c);
    bool init_chains();
    bool morph_step();
    bool autotune(double budget_ms, int rounds_fixed, std::vector<amh_tune_entry> *report);
    bool exact_cost(uint64_t *lo, uint64_t *hi);
    bool snapshot();
    void run();
    void start_or_resume(size_t n, double s);
};

morph::engine::engine() { set_perms(0); }

// PerlinNoise(seed) for the lag map and PerlinNoise(seed + 1) for the slope
// map (morph.cpp:436-444, perlin.cpp:11-17): std::shuffle with mt19937.
void morph::engine::set_perms(unsigned seed) {
    auto make = [](unsigned s, int32_t *p) {
        if (s == 0) s = std::mt19937::default_seed;
        int tmp[256];
        std::iota(tmp, tmp + 256, 0);
        std::shuffle(tmp, tmp + 256, std::mt19937(s));
        for (int i = 0; i < 256; ++i) p[i] = p[256 + i] = tmp[i];
    };
    make(seed, lag_perm);
    make(seed + 1, slope_perm);
}
morph::engine::~engine() {
    {
        std::unique_lock<std::mutex> lk(mu);
        stop = true;
        paused = false;
    }
    cv.notify_all();
    if (th.joinable()) th.join();
    if (graph) amh_graph_destroy(graph);
    if (rgraph) amh_render_graph_destroy(rgraph);
    if (fluid) amh_fluid_destroy(fluid);
    if (mlp) amh_mlp_destroy(mlp);
    for (auto &v : out_blobs)
        for (blob *b : v) delete b;
    // device buffers free themselves; streams last
    if (cstream) amh_stream_destroy(cstream);
    if (rstream) amh_stream_destroy(rstream);
}

bool morph::engine::fail(const char *what) {
    err = std::string(what) + ": " + amh_last_error();
    return false;
}

bool morph::engine::ensure_streams() {
    if (!cstream && amh_stream_create(&cstream) != AMH_OK) return fail("stream");
    if (!rstream && amh_stream_create(&rstream) != AMH_OK) return fail("stream");
    return true;
}

// Worker state back to blob detection (thread::clear, thread.cpp:35-84).
void morph::engine::reset_worker() {
    wstate = STATE_BLOB_DETECTION;
    counter = 1;
    matched = false;
    match_steps = 0;
    fb.clear();
    bmap.clear();
    best_map.clear();
    best_e = std::numeric_limits<double>::infinity();
    best_report = 0;
    map_e = 0;
    wtot = 0;
    off.clear();
    cw.clear();
    nitems = 0;
    items_tile = -1;
    tuned_tile = -1;
    tuned_rounds = 0;
    visits = passes = 0;
    if (graph) {
        amh_graph_destroy(graph);
        graph = nullptr;
    }
    if (rgraph) {
        amh_render_graph_destroy(rgraph);
        rgraph = nullptr;
        rgraph_key.clear();
    }
    have_snap = false;
    fluid_inputs = false;
    // a restart deletes the fluid model (morph.cpp:131-132); the handle is
    // kept and reset on its next use when the sizes allow, with its
    // captured frame graphs
    fluid_stale = fluid != nullptr;
    rng.seed(wopt.seed);
}

void morph::engine::fill_render(amh_render &r) const {
    std::memset(&r, 0, sizeof(r));
    r.width = CW; r.height = CH; r.nkeys = h; r.nblobs = int32_t(bw);
    r.bbox_x1 = bbox_x1; r.bbox_y1 = bbox_y1; r.bbox_x2 = bbox_x2; r.bbox_y2 = bbox_y2;
    r.motion = opt.motion; r.fading = opt.fading; r.show_blobs = int32_t(opt.show_blobs);
    r.density = opt.density; r.feather = int32_t(opt.feather);
    r.keep_background = opt.keep_background ? 1 : 0; r.blend_blobs = opt.blend_blobs ? 1 : 0;
    r.nchains = int32_t(cw.size()); r.wtot = wtot;
    r.keys_dev = d_shown.as<uint32_t>(); r.pts_dev = d_snap.as<uint64_t>();
    r.chain_off_dev = d_off.as<int64_t>(); r.chain_w_dev = d_w.as<int32_t>();
    r.rank_dev = d_rank.as<int32_t>(); r.blob_rgba_dev = d_brgba.as<uint32_t>();
    r.lag_perm_dev = d_lag.as<int32_t>(); r.slope_perm_dev = d_slope.as<int32_t>();
}

// Per key segment: which atoms carry particles (rank: both ends hold fluid,
// then source only, then destination only, hash order inside each class),
// the blob sizes |A|, |B| and the AVERAGE / DISTINCT colours
// (update_particle, morph.cpp:760-777).
bool morph::engine::prepare_fluid() {
    const size_t nch = cw.size();
    if (!have_snap || nch == 0 || wtot == 0) return false;
    const double sp = opt.fluid_sparseness > 0.0 ? opt.fluid_sparseness : 1.0;
    const int32_t gsx = int32_t(CW * sp) + 20, gsy = int32_t(CH * sp) + 20;
    const int32_t dims[5] = {gsx, gsy, int32_t(wtot), CW, CH};
    const bool same = fluid && std::equal(dims, dims + 5, fluid_dims);
    if (!same || fluid_sp != sp) {
        if (fluid) amh_fluid_destroy(fluid);
        fluid = nullptr;
        if (amh_fluid_create(&fluid, gsx, gsy, int32_t(wtot), CW, CH, h, rstream)) return fail("fluid create");
        if (amh_fluid_reset(fluid, sp, rstream)) return fail("fluid reset");
        std::copy(dims, dims + 5, fluid_dims);
        fluid_sp = sp;
    } else if (fluid_stale) {  // a new model's particles on the same handle
        if (amh_fluid_reset(fluid, sp, rstream)) return fail("fluid reset");
    }
    fluid_stale = false;
    if (fluid_inputs) return true;
    // per key segment and chain: blob size and its AVERAGE / DISTINCT colour
    std::vector<int32_t> surf(size_t(h) * nch);
    std::vector<double> bcol(size_t(h) * nch * 4, 0.0);
    for (int y = 0; y < h; ++y) {
        for (size_t c = 0; c < nch; ++c) {
            const BlobInfo &bi = fb[y][best_map[c * h + y]];
            surf[size_t(y) * nch + c] = int32_t(bi.size);
            double *q = &bcol[(size_t(y) * nch + c) * 4];
            if (opt.show_blobs == DISTINCT) {
                std::mt19937 gen(static_cast<std::mt19937::result_type>(c));
                std::uniform_real_distribution<double> u(0.0, 1.0);
                q[0] = u(gen); q[1] = u(gen); q[2] = u(gen); q[3] = 1.0;
            } else {
                q[0] = bi.r; q[1] = bi.g; q[2] = bi.b; q[3] = bi.a;
            }
        }
    }
    if (!d_frank.ensure(4 * size_t(wtot) * h) || !d_fsurf.ensure(4 * surf.size()) ||
        !d_fbcol.ensure(8 * bcol.size()))
        return fail("device allocation");
    // activation order of the atoms, on the device from the render snapshot
    if (amh_fluid_ranks(d_snap.as<uint64_t>(), h, wtot, d_off.as<int64_t>(), d_w.as<int32_t>(), int32_t(nch),
                        mix_seed(), d_frank.as<int32_t>(), rstream))
        return fail("fluid ranks");
    if (!d_fkeys.ensure(8 * size_t(wtot) * h) ||
        amh_fluid_atom_keys(d_snap.as<uint64_t>(), h, wtot, d_fkeys.as<uint64_t>(), rstream))
        return fail("fluid atom keys");
    if (amh_memcpy_h2d(d_fsurf.p, surf.data(), 4 * surf.size(), rstream) ||
        amh_memcpy_h2d(d_fbcol.p, bcol.data(), 8 * bcol.size(), rstream) || amh_stream_sync(rstream))
        return fail("fluid inputs");
    fluid_inputs = true;
    return true;
}

bool morph::engine::render_fluid(double t, uint32_t *out_dev) {
    if (!prepare_fluid()) return false;
    amh_render r;
    fill_render(r);
    amh_fluid_frame fa;
    std::memset(&fa, 0, sizeof(fa));
    fa.r = &r;
    fa.rank_dev = d_frank.as<int32_t>();
    fa.surface_dev = d_fsurf.as<int32_t>();
    fa.blob_col_dev = opt.show_blobs == TEXTURE ? nullptr : d_fbcol.as<double>();
    fa.sparseness = fluid_sp;
    fa.fluidsteps = int32_t(opt.fluidsteps);
    fa.atom_keys_dev = d_fkeys.as<uint64_t>();
    if (amh_fluid_render(fluid, &fa, t, out_dev, rstream)) return fail("fluid render");
    return true;
}

// Key frames onto the device canvas (morph::synchronize -> worker.set_frame,
// morph.cpp:122-136, thread.cpp:92-111).
bool morph::engine::upload_frames() {
    keys.clear();
    for (auto &kv : frames) keys.push_back(kv.first);
    h = int32_t(keys.size());
    CW = std::max<int32_t>(width, bbox_x1 <= bbox_x2 ? bbox_x2 + 1 : 1);
    CH = std::max<int32_t>(height, bbox_y1 <= bbox_y2 ? bbox_y2 + 1 : 1);
    const size_t npx = size_t(CW) * CH;
    canvas.assign(h, std::vector<uint32_t>(npx, 0u));
    fmean.assign(size_t(h) * 6, 0.0);
    for (int y = 0; y < h; ++y) {
        auto &cv_ = canvas[y];
        for (auto &e : frames[keys[y]].px) {
            const uint32_t x = e.first & 0xFFFFu, yy = e.first >> 16;
            if (x < uint32_t(CW) && yy < uint32_t(CH)) cv_[size_t(yy) * CW + x] = e.second;
        }
    }
    if (h == 0) return true;
    if (!ensure_streams()) return false;
    if (!d_rgba.ensure(npx * 4 * h) || !d_stored.ensure(npx * 4 * h) || !d_shown.ensure(npx * 4 * h) ||
        !d_labels.ensure(npx * 4) || !d_bop.ensure(npx * 4 * h) || !d_plist.ensure(npx * 4 * h))
        return fail("device allocation");
    for (int y = 0; y < h; ++y)
        if (amh_memcpy_h2d(d_rgba.as<uint32_t>() + size_t(y) * npx, canvas[y].data(), npx * 4, cstream))
            return fail("upload");
    return amh_stream_sync(cstream) == AMH_OK || fail("upload sync");
}

// BLOB_DETECTION + UNIFICATION in one step per frame set: ingest, segment,
// per-blob statistics and the row-major pixel lists of every blob.  With
// the default blob options the blobs are the 4-connected components of the
// foreground (device segmentation); otherwise the greedy merge and dust
// unification run on the host over the stored colours (blobify.cpp).
bool morph::engine::detect() {
    const size_t npx = size_t(CW) * CH;
    nblobs.assign(h, 0);
    stats.assign(h, {});
    if (!d_bws.ensure(amh_blob_workspace_bytes(CW, CH))) return fail("device allocation");
    std::vector<int64_t> first_off(h + 1, 0);
    am_blobify::options bo;
    bo.threshold = wopt.blob_threshold;
    bo.max_size = wopt.blob_max_size;
    bo.min_size = wopt.blob_min_size;
    bo.number = wopt.blob_number;
    bo.box_samples = wopt.blob_box_samples;
    bo.box_grip = wopt.blob_box_grip;
    const bool greedy = am_blobify::needed(bo);
    std::vector<am_blobify::frame_result> gres;
    std::vector<std::vector<uint32_t>> gstored;
    if (greedy) {
        gstored.assign(h, std::vector<uint32_t>(npx));
        std::vector<const uint32_t *> ptrs(h);
        for (int y = 0; y < h; ++y) {
            uint32_t *stored = d_stored.as<uint32_t>() + size_t(y) * npx;
            if (amh_ingest_frame(d_rgba.as<uint32_t>() + size_t(y) * npx, CW, CH, wopt.blob_delimiter == HSP ? 1 : 0,
                                 stored, d_shown.as<uint32_t>() + size_t(y) * npx, cstream) ||
                amh_memcpy_d2h(gstored[y].data(), stored, npx * 4, cstream))
                return fail("ingest");
            ptrs[y] = gstored[y].data();
        }
        if (amh_stream_sync(cstream)) return fail("ingest sync");
        if (!am_blobify::segment(ptrs, CW, CH, bo, rng, gres))
            return fail("greedy blob detection: canvas too large for 31-bit pixel indices");
        for (int y = 0; y < h; ++y) {
            nblobs[y] = int32_t(gres[y].blobs.size());
            if (amh_memcpy_h2d(d_bop.as<int32_t>() + size_t(y) * npx, gres[y].bop.data(), npx * 4, cstream))
                return fail("blob upload");
            first_off[y + 1] = first_off[y] + nblobs[y] + 1;
        }
        if (amh_stream_sync(cstream)) return fail("blob upload sync");
    }
    for (int y = 0; y < h && !greedy; ++y) {
        uint32_t *rgba = d_rgba.as<uint32_t>() + size_t(y) * npx;
        uint32_t *stored = d_stored.as<uint32_t>() + size_t(y) * npx;
        uint32_t *shown = d_shown.as<uint32_t>() + size_t(y) * npx;
        if (amh_ingest_frame(rgba, CW, CH, wopt.blob_delimiter == HSP ? 1 : 0, stored, shown, cstream))
            return fail("ingest");
        if (amh_segment(rgba, CW, CH, d_labels.as<int32_t>(), cstream)) return fail("segment");
        int32_t nb = 0;
        if (amh_blob_count(d_labels.as<int32_t>(), CW, CH, d_bws.p, &nb, cstream)) return fail("blob count");
        nblobs[y] = nb;
        if (!d_stats.ensure(sizeof(amh_blob_stat) * std::max(nb, 1))) return fail("device allocation");
        int32_t *bop = d_bop.as<int32_t>() + size_t(y) * npx;
        if (amh_blob_stats(d_labels.as<int32_t>(), stored, CW, CH, d_bws.p, nb, bop, d_stats.as<amh_blob_stat>(),
                           cstream))
            return fail("blob stats");
        stats[y].resize(nb);
        if (nb && amh_memcpy_d2h(stats[y].data(), d_stats.p, sizeof(amh_blob_stat) * nb, cstream))
            return fail("stats download");
        if (amh_stream_sync(cstream)) return fail("detect sync");
        first_off[y + 1] = first_off[y] + nb + 1;
    }
    // pixel lists, grouped by blob, per frame
    if (!d_first.ensure(sizeof(int64_t) * first_off[h])) return fail("device allocation");
    for (int y = 0; y < h; ++y) {
        const size_t gb = amh_group_workspace_bytes(int32_t(npx), nblobs[y]);
        if (!d_gws.ensure(gb)) return fail("device allocation");
        if (amh_group_pixels(d_bop.as<int32_t>() + size_t(y) * npx, int32_t(npx), nblobs[y], d_gws.p, gb,
                             d_first.as<int64_t>() + first_off[y], d_plist.as<int32_t>() + size_t(y) * npx,
                             cstream))
            return fail("group pixels");
        if (amh_stream_sync(cstream)) return fail("group sync");
    }
    // frame means (morph::add_pixel's running means, as exact averages over
    // the stored colours) and the blob table of every frame
    fb.assign(h, {});
    for (int y = 0; y < h; ++y) {
        double sx = 0, sy = 0, sc[4] = {0, 0, 0, 0};
        size_t n = 0;
        if (greedy) {  // the same sums over the stored pixels
            const auto &bop = gres[y].bop;
            std::vector<int> first(nblobs[y], -1);
            int64_t ssx = 0, ssy = 0, ssc[4] = {0, 0, 0, 0};
            for (size_t i = 0; i < npx; ++i) {
                if (bop[i] < 0) continue;
                if (first[bop[i]] < 0) first[bop[i]] = int(i);
                ssx += int64_t(i % CW);
                ssy += int64_t(i / CW);
                const color c = word_color(gstored[y][i]);
                ssc[0] += c.r; ssc[1] += c.g; ssc[2] += c.b; ssc[3] += c.a;
                ++n;
            }
            sx = double(ssx); sy = double(ssy);
            for (int k = 0; k < 4; ++k) sc[k] = double(ssc[k]);
            for (int b = 0; b < nblobs[y]; ++b) {
                const am_blobify::blob_info &g = gres[y].blobs[b];
                BlobInfo bi;
                bi.id = b;
                bi.size = g.size;
                bi.x = g.x; bi.y = g.y;
                bi.r = g.r; bi.g = g.g; bi.b = g.b; bi.a = g.a;
                bi.first = first[b];
                bi.group0 = size_t(g.group);
                fb[y].push_back(bi);
            }
        }
        for (const auto &s : stats[y]) {
            sx += double(s.sum_x); sy += double(s.sum_y);
            for (int k = 0; k < 4; ++k) sc[k] += double(s.sum_c[k]);
            n += size_t(s.count);
        }
        double *m = &fmean[size_t(y) * 6];
        if (n) {
            m[0] = sx / n; m[1] = sy / n;
            for (int k = 0; k < 4; ++k) m[2 + k] = sc[k] / (255.0 * n);
        } else {  // morph::refresh_frames (morph.cpp:271-286)
            m[0] = bbox_x1 <= bbox_x2 ? (bbox_x1 + bbox_x2) / 2.0 : 0.0;
            m[1] = bbox_y1 <= bbox_y2 ? (bbox_y1 + bbox_y2) / 2.0 : 0.0;
        }
        for (int b = 0; b < nblobs[y] && !greedy; ++b) {
            const amh_blob_stat &s = stats[y][b];
            BlobInfo bi;
            bi.id = b;
            bi.size = size_t(s.count);
            bi.x = double(s.sum_x) / s.count;
            bi.y = double(s.sum_y) / s.count;
            bi.r = double(s.sum_c[0]) / (255.0 * s.count);
            bi.g = double(s.sum_c[1]) / (255.0 * s.count);
            bi.b = double(s.sum_c[2]) / (255.0 * s.count);
            bi.a = double(s.sum_c[3]) / (255.0 * s.count);
            bi.first = int(s.first);
            bi.group0 = xy2pos(uint16_t(s.first % CW), uint16_t(s.first / CW));
            fb[y].push_back(bi);
        }
    }
    d_first_offsets = first_off;
    return true;
}

// thread::blob_distance (thread.cpp:1151-1174).
double morph::engine::blob_distance(const BlobInfo &b1, const BlobInfo &b2) const {
    const size_t sz1 = b1.size, sz2 = b2.size, szs = sz1 + sz2;
    double pix = 0.0, col = 0.0, siz = 0.0;
    if (szs > 0) siz = std::fabs(double(sz1) - double(sz2)) / double(szs);
    if (sz1 > 0 && sz2 > 0) {  // the pixels are made once per blob by match_init
        pix = std::sqrt(double(pixel_distance(b1.px, b2.px)) / bbox_d);
        col = color_distance(b1.px.c, b2.px.c);
    }
    return w_xy * pix + w_rgba * col + w_size * siz;
}

// thread::get_energy(blob***) (thread.cpp:1087-1107).
double morph::engine::map_energy(const std::vector<int32_t> &m) const {
    double e = 0.0;
    for (size_t c = 0; c < bw; ++c) {
        const BlobInfo *prev = &fb[h - 1][m[c * h + h - 1]];
        for (int y = 0; y < h; ++y) {
            const BlobInfo *cur = &fb[y][m[c * h + y]];
            e += blob_distance(*prev, *cur);
            prev = cur;
        }
    }
    return e;
}

// First matching step (thread.cpp:599-665): pad every frame with empty blobs
// at the frame mean, shuffle each frame's blobs, column c = group c.
bool morph::engine::match_init() {
    bw = 0;
    for (int y = 0; y < h; ++y) bw = std::max(bw, fb[y].size());
    for (int y = 0; y < h; ++y) {
        const double *m = &fmean[size_t(y) * 6];
        while (fb[y].size() < bw) {
            BlobInfo e;
            e.empty = true;
            e.x = m[0]; e.y = m[1]; e.r = m[2]; e.g = m[3]; e.b = m[4]; e.a = m[5];
            fb[y].push_back(e);
        }
        std::shuffle(fb[y].begin(), fb[y].end(), rng);
        for (BlobInfo &bi : fb[y])  // create_pixel as thread::blob_distance does per call
            bi.px = create_pixel(to_u16(bi.x), to_u16(bi.y), create_color(bi.r, bi.g, bi.b, bi.a));
    }
    bmap.assign(bw * h, 0);
    for (size_t c = 0; c < bw; ++c)
        for (int y = 0; y < h; ++y) bmap[c * h + y] = int32_t(c);
    best_map = bmap;
    const int dx = bbox_x2 - bbox_x1, dy = bbox_y2 - bbox_y1;
    bbox_d = bbox_x1 <= bbox_x2 ? uint32_t(dx * dx + dy * dy) : 0u;
    const double sum = double(wopt.blob_rgba_weight) * wopt.blob_rgba_weight +
                       double(wopt.blob_size_weight) * wopt.blob_size_weight +
                       double(wopt.blob_xy_weight) * wopt.blob_xy_weight;
    if (sum > 0.0) {  // thread::set_blob_weights (thread.cpp:1141-1149)
        w_rgba = double(wopt.blob_rgba_weight * wopt.blob_rgba_weight) / sum;
        w_size = double(wopt.blob_size_weight * wopt.blob_size_weight) / sum;
        w_xy = 1.0 - (w_rgba + w_size);
    }
    // best_e stays infinite until the first accepted swap (thread.cpp:51),
    // so that swap always becomes the best map, even at zero gain.
    map_e = bw > 0 && h > 0 ? map_energy(bmap) : 0.0;
    best_e = std::numeric_limits<double>::infinity();
    best_report = 0.0;
    matched = true;
    return false;
}

// One random blob swap at a random key frame (thread.cpp:667-738).
// Returns true when matching is finished.
bool morph::engine::match_step() {
    if (!matched) return match_init();
    if (h == 0 || bw <= 1) return true;
    ++match_steps;
    std::uniform_int_distribution<size_t> ux(0, bw - 1), uy(0, size_t(h) - 1);
    const size_t y = uy(rng), yn = (y + 1) % h, yp = y > 0 ? y - 1 : h - 1;
    const size_t x1 = ux(rng);
    size_t x2;
    do { x2 = ux(rng); } while (x1 == x2);
    const BlobInfo &a_p = fb[yp][bmap[x1 * h + yp]], &b_p = fb[yp][bmap[x2 * h + yp]];
    const BlobInfo &a_c = fb[y][bmap[x1 * h + y]], &b_c = fb[y][bmap[x2 * h + y]];
    const BlobInfo &a_n = fb[yn][bmap[x1 * h + yn]], &b_n = fb[yn][bmap[x2 * h + yn]];
    if (a_c.empty && b_c.empty) return wopt.match_budget && match_steps >= wopt.match_budget;
    // summed per blob first, as thread.cpp:690-696 does
    const double c1 = (blob_distance(a_p, a_c) + blob_distance(a_c, a_n)) +
                      (blob_distance(b_p, b_c) + blob_distance(b_c, b_n));
    const double c2 = (blob_distance(b_p, a_c) + blob_distance(a_c, b_n)) +
                      (blob_distance(a_p, b_c) + blob_distance(b_c, a_n));
    if (c1 >= c2 || (wopt.degeneration && (counter % wopt.degeneration) == 0)) {
        std::swap(bmap[x1 * h + y], bmap[x2 * h + y]);
        map_e -= c1 - c2;
        if (map_e < 0.0) map_e = 0.0;
        if (map_e < best_e) {  // the groups follow the best map seen
            best_e = map_e;
            best_report = best_e;
            best_map = bmap;
            counter = 0;
        }
    }
    if (best_e == 0.0) return true;
    return wopt.match_budget && match_steps >= wopt.match_budget;
}

// thread::fix_volatiles (thread.cpp:1187-1233): empty blobs of a chain take
// positions interpolated between the neighbouring real blobs, cyclically.
void morph::engine::fix_volatiles(size_t c) {
    const size_t sz = size_t(h);
    if (sz <= 1) return;
    auto at = [&](size_t y) -> BlobInfo & { return fb[y][best_map[c * h + y]]; };
    size_t i = 0;
    bool started = false;
    std::vector<BlobInfo *> vol;
    BlobInfo *first = nullptr, *prev = nullptr;
    for (;;) {
        i = (i + 1) % sz;
        BlobInfo *bl = &at(i);
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
        if (!vol.empty()) {
            const size_t vsz = vol.size();
            for (size_t v = 0; v < vsz; ++v) {
                const double t = (v + 1.0) / double(vsz + 1.0);
                vol[v]->x = t * bl->x + (1.0 - t) * prev->x;
                vol[v]->y = t * bl->y + (1.0 - t) * prev->y;
            }
            vol.clear();
        }
        prev = bl;
        if (prev == first) break;
    }
}

// thread::init_morph (thread.cpp:740-891) on the device.
bool morph::engine::init_chains() {
    const size_t nch = bw;
    for (size_t c = 0; c < nch; ++c) fix_volatiles(c);
    cw.assign(nch, 0);
    off.assign(nch, 0);
    std::vector<int32_t> cblob(nch * h);
    std::vector<uint32_t> exy(nch * h);
    int64_t tot = 0;
    int32_t max_w = 0;
    for (size_t c = 0; c < nch; ++c) {
        size_t mx = 0;
        for (int y = 0; y < h; ++y) mx = std::max(mx, fb[y][best_map[c * h + y]].size);
        const int64_t w = int64_t(mx) * std::max<uint16_t>(wopt.density, 1);
        cw[c] = int32_t(w);
        off[c] = tot;
        tot += w;
        max_w = std::max(max_w, cw[c]);
        for (int y = 0; y < h; ++y) {
            const BlobInfo &bi = fb[y][best_map[c * h + y]];
            cblob[c * h + y] = bi.empty ? -1 : bi.id;
            exy[c * h + y] = uint32_t(to_u16(std::round(bi.x))) | (uint32_t(to_u16(std::round(bi.y))) << 16);
        }
    }
    wtot = tot;
    if (h == 0 || nch == 0 || max_w <= 1 || wtot == 0) return false;
    if (!d_pts.ensure(sizeof(uint64_t) * size_t(wtot) * h) || !d_off.ensure(8 * nch) || !d_w.ensure(4 * nch) ||
        !d_cblob.ensure(4 * nch * h) || !d_exy.ensure(4 * nch * h) || !d_queue.ensure(16) ||
        !d_rstats.ensure(sizeof(amh_refine_stats)) || !d_cost.ensure(16 * nch) || !d_snap.ensure(8 * size_t(wtot) * h))
        return fail("device allocation");
    if (amh_memcpy_h2d(d_off.p, off.data(), 8 * nch, cstream) || amh_memcpy_h2d(d_w.p, cw.data(), 4 * nch, cstream) ||
        amh_memcpy_h2d(d_cblob.p, cblob.data(), 4 * nch * h, cstream) ||
        amh_memcpy_h2d(d_exy.p, exy.data(), 4 * nch * h, cstream) || amh_memset(d_queue.p, 0, 16, cstream) ||
        amh_memset(d_rstats.p, 0, sizeof(amh_refine_stats), cstream))
        return fail("chain upload");
    const size_t npx = size_t(CW) * CH;
    std::vector<const int32_t *> plist(h);
    std::vector<const int64_t *> first(h);
    for (int y = 0; y < h; ++y) {
        plist[y] = d_plist.as<int32_t>() + size_t(y) * npx;
        first[y] = d_first.as<int64_t>() + d_first_offsets[y];
    }
    amh_chain_build a;
    std::memset(&a, 0, sizeof(a));
    a.nkeys = h;
    a.nchains = int32_t(nch);
    a.wtot = wtot;
    a.frame_width = CW;
    a.chain_off_dev = d_off.as<int64_t>();
    a.chain_w_dev = d_w.as<int32_t>();
    a.chain_blob_dev = d_cblob.as<int32_t>();
    a.empty_xy_dev = d_exy.as<uint32_t>();
    a.pixel_list_dev = plist.data();
    a.blob_first_dev = first.data();
    a.seed = mix_seed();
    if (amh_build_chains(&a, d_pts.as<uint64_t>(), cstream)) return fail("build chains");
    if (amh_stream_sync(cstream)) return fail("build sync");
    // composite order of each key frame: blob index of every chain
    rank.assign(size_t(h) * nch, -1);
    for (size_t c = 0; c < nch; ++c)
        for (int y = 0; y < h; ++y) rank[size_t(y) * nch + c] = best_map[c * h + y];
    return true;
}

// Autotuned batched schedules, shared by every engine of the process and
// keyed by the chain shapes (key frames, then every chain width): a job on
// the same shapes reuses the tile and rounds measured for the first one.
namespace {
std::mutex g_tune_mu;
std::map<std::vector<int64_t>, std::pair<int, int>> g_tune;
constexpr double kTuneBudgetMs = 3.0;
}  // namespace

bool morph::engine::autotune(double budget_ms, int rounds_fixed, std::vector<amh_tune_entry> *report) {
    const size_t nch = cw.size();
    if (nch == 0 || h < 2 || wtot == 0) {
        err = "autotune: no chains";
        return false;
    }
    if (!d_tune.ensure(sizeof(uint64_t) * size_t(wtot) * h)) return fail("device allocation");
    std::vector<amh_tune_entry> rep(64);
    int32_t t = -1, r = 0, n = 0;
    const bool ok = amh_refine_autotune(d_pts.as<uint64_t>(), d_tune.as<uint64_t>(), h, wtot, d_off.as<int64_t>(),
                                        d_w.as<int32_t>(), cw.data(), int32_t(nch), budget_ms, mix_seed(),
                                        int32_t(rounds_fixed), &t, &r, rep.data(), int32_t(rep.size()), &n,
                                        cstream) == AMH_OK;
    d_tune.release();
    if (!ok) return fail("autotune");
    rep.resize(std::min<size_t>(size_t(std::max(n, 0)), rep.size()));
    if (report) *report = rep;
    tuned_tile = t;
    tuned_rounds = r;
    std::vector<int64_t> key{int64_t(h), int64_t(rounds_fixed)};
    key.insert(key.end(), cw.begin(), cw.end());
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = {t, r};
    return true;
}

// kNN-guided refinement: its kernel variants (the first four tiles), the
// default one (4096 slots, 512 threads) and rounds per pass.
constexpr int kKnnTiles = 4, kKnnTile = 3, kKnnRounds = 16;

// ATOM_MORPHING: one refinement step.  Plain: one visit of one chain
// (thread::morph, thread.cpp:1043-1068: chain counter % W, one random key
// frame, cycle_length attempts as conflict-free rounds).  Batched: one pass
// over every chain, or one hipGraph replay of graph_passes passes.  kNN: one
// pass over every chain with nearest-neighbour partners (thread.cpp:893-986).
bool morph::engine::morph_step() {
    const size_t nch = cw.size();
    if (nch == 0 || h < 2) return false;
    if (wopt.refine_mode == REFINE_BATCHED || wopt.refine_mode == REFINE_KNN) {
        const bool knn = wopt.refine_mode == REFINE_KNN;
        int tile = wopt.refine_tile, rounds = int(wopt.refine_rounds);
        if (knn) {  // kernel variants of up to 4096 slots; not autotuned
            if (tile < 0 || tile >= kKnnTiles) tile = kKnnTile;
            if (rounds <= 0) rounds = kKnnRounds;
        } else if (tile < 0 || rounds <= 0) {  // autotuned: once per chain shape (and fixed rounds)
            const int fixed = rounds > 0 ? rounds : 0;
            if (tuned_tile < 0 || tuned_fixed != fixed) {
                tuned_tile = -1;
                tuned_fixed = fixed;
                std::vector<int64_t> key{int64_t(h), int64_t(fixed)};
                key.insert(key.end(), cw.begin(), cw.end());
                {
                    std::lock_guard<std::mutex> lk(g_tune_mu);
                    auto it = g_tune.find(key);
                    if (it != g_tune.end()) {
                        tuned_tile = it->second.first;
                        tuned_rounds = it->second.second;
                    }
                }
                if (tuned_tile < 0 && !autotune(kTuneBudgetMs, fixed, nullptr)) return true;
            }
            if (tile < 0) tile = tuned_tile;
            if (rounds <= 0) rounds = tuned_rounds;
        }
        if (items_tile != tile) {
            int32_t n = 0;
            if (amh_refine_items(cw.data(), int32_t(nch), tile, nullptr, 0, &n)) return fail("refine items");
            std::vector<int32_t> items(2 * size_t(std::max(n, 1)));
            if (amh_refine_items(cw.data(), int32_t(nch), tile, items.data(), n, &n)) return fail("refine items");
            if (!d_items.ensure(items.size() * 4)) return fail("device allocation");
            if (amh_memcpy_h2d(d_items.p, items.data(), items.size() * 4, cstream)) return fail("items upload");
            nitems = n;
            items_tile = tile;
        }
        if (nitems == 0) return false;
        if (knn) {
            if (amh_refine_knn(d_pts.as<uint64_t>(), h, wtot, d_off.as<int64_t>(), d_w.as<int32_t>(),
                               d_items.as<int32_t>(), nitems, tile, passes, rounds, mix_seed(),
                               d_queue.as<uint32_t>(), d_rstats.as<amh_refine_stats>(), cstream))
                return fail("refine knn");
            ++passes;
        } else if (wopt.graph_passes > 0) {
            if (!graph || graph_tile != tile || graph_np != wopt.graph_passes || graph_rounds != unsigned(rounds)) {
                if (graph) amh_graph_destroy(graph);
                graph = nullptr;
                if (amh_refine_graph_create(&graph, d_pts.as<uint64_t>(), h, wtot, d_off.as<int64_t>(),
                                            d_w.as<int32_t>(), d_items.as<int32_t>(), nitems, tile,
                                            int32_t(wopt.graph_passes), rounds, mix_seed(), d_queue.as<uint32_t>(),
                                            d_rstats.as<amh_refine_stats>(), cstream))
                    return fail("refine graph");
                graph_tile = tile;
                graph_np = wopt.graph_passes;
                graph_rounds = unsigned(rounds);
            }
            if (amh_graph_launch(graph, cstream)) return fail("graph launch");
            passes += wopt.graph_passes;
        } else {
            if (amh_refine_batched(d_pts.as<uint64_t>(), h, wtot, d_off.as<int64_t>(), d_w.as<int32_t>(),
                                   d_items.as<int32_t>(), nitems, tile, passes, rounds, mix_seed(),
                                   d_queue.as<uint32_t>(), d_rstats.as<amh_refine_stats>(), cstream))
                return fail("refine batched");
            ++passes;
        }
        return false;
    }
    const size_t c = counter % nch;
    if (cw[c] <= 1) return false;
    const int32_t y = int32_t(rng() % uint64_t(h));
    // rounds per visit: fixed when set, else cycle_length attempts
    const uint64_t pairs = uint64_t(cw[c] / 2);
    const uint64_t rounds = wopt.refine_rounds > 0 ? uint64_t(wopt.refine_rounds)
                                                   : std::max<uint64_t>(1, (wopt.cycle_length + pairs - 1) / pairs);
    if (amh_refine_chain(d_pts.as<uint64_t>(), h, wtot, off[c], cw[c], y, rounds, mix_seed(),
                         (uint64_t(c) << 32) ^ visits, d_rstats.as<amh_refine_stats>(), cstream))
        return fail("refine chain");
    ++visits;
    return false;
}

// thread::step (thread.cpp:139-171).
bool morph::engine::step() {
    switch (wstate) {
        case STATE_BLOB_DETECTION:
            if (!detect()) wstate = STATE_DONE;
            else wstate = STATE_BLOB_UNIFICATION;
            counter = 0;
            break;
        case STATE_BLOB_UNIFICATION:  // dust unification: a no-op at blob_min_size 1
            wstate = STATE_BLOB_MATCHING;
            counter = 0;
            break;
        case STATE_BLOB_MATCHING:
            if (wskip || match_step()) {
                if (!matched) match_init();
                wstate = init_chains() ? STATE_ATOM_MORPHING : STATE_DONE;
                if (wstate == STATE_ATOM_MORPHING && wopt.learned && !warm_start()) wstate = STATE_DONE;
                counter = 0;
            }
            break;
        case STATE_ATOM_MORPHING:
            if (wskip || morph_step()) {
                wstate = STATE_DONE;
                counter = 0;
            }
            break;
        default:
            break;
    }
    wskip = false;
    counter++;
    return true;
}

bool morph::engine::exact_cost(uint64_t *lo, uint64_t *hi) {
    *lo = *hi = 0;
    const size_t nch = cw.size();
    if (!nch || wtot == 0 || !have_snap) return true;
    if (amh_chain_cost(d_snap.as<uint64_t>(), h, wtot, d_off.as<int64_t>(), d_w.as<int32_t>(), int32_t(nch),
                       d_cost.as<uint64_t>(), rstream))
        return fail("chain cost");
    std::vector<uint64_t> c(2 * nch);
    if (amh_memcpy_d2h(c.data(), d_cost.p, 16 * nch, rstream) || amh_stream_sync(rstream)) return fail("cost sync");
    unsigned __int128 s = 0;
    for (size_t i = 0; i < nch; ++i) s += (unsigned __int128)c[2 * i] | ((unsigned __int128)c[2 * i + 1] << 64);
    *lo = uint64_t(s);
    *hi = uint64_t(s >> 64);
    return true;
}

// The MLP handle with the caller's current parameters, and the warm start
// workspace for the current chains (allocations and uploads, done before
// any timed work).
bool morph::engine::prepare_mlp() {
    if (mlp_params.size() != AMH_MLP_PARAMS) return fail("warm start: no MLP parameters (set_learned)");
    if (!mlp && amh_mlp_create(&mlp, 65536, cstream)) return fail("mlp create");
    if (mlp_loaded != mlp_version) {
        if (amh_mlp_set_params(mlp, mlp_params.data(), cstream)) return fail("mlp params");
        mlp_loaded = mlp_version;
    }
    if (!d_wws.ensure(amh_warm_workspace_bytes(wtot, int32_t(cw.size()), h))) return fail("warm start workspace");
    return true;
}

// Learned warm start of the freshly built chains (in place on d_pts).
bool morph::engine::warm_start() {
    const size_t nch = cw.size();
    if (!nch || wtot == 0 || h < 2) return true;
    if (!prepare_mlp()) return false;
    const size_t bytes = amh_warm_workspace_bytes(wtot, int32_t(nch), h);
    amh_render r;
    std::memset(&r, 0, sizeof(r));
    r.width = CW; r.height = CH; r.nkeys = h; r.nchains = int32_t(nch); r.wtot = wtot;
    r.keys_dev = d_shown.as<uint32_t>(); r.pts_dev = d_pts.as<uint64_t>();
    r.chain_off_dev = d_off.as<int64_t>(); r.chain_w_dev = d_w.as<int32_t>();
    if (amh_warm_start(mlp, &r, d_pts.as<uint64_t>(), d_wws.p, bytes, cstream)) return fail("warm start");
    return true;
}

// Copies the refined chains (and what rendering needs) for the caller side;
// the reference deep-copies chains at synchronize (morph.cpp:174-214).
bool morph::engine::snapshot() {
    const size_t nch = cw.size();
    if (!nch || wtot == 0) return true;
    if (amh_memcpy_d2d(d_snap.p, d_pts.p, 8 * size_t(wtot) * h, cstream)) return fail("snapshot");
    brgba.assign(size_t(h) * bw, 0u);
    // group of blob b in frame y = the chain whose column holds it
    std::vector<size_t> group(bw);
    for (int y = 0; y < h; ++y) {
        std::fill(group.begin(), group.end(), size_t(0));
        for (size_t c = 0; c < nch; ++c) {
            const int32_t b = rank[size_t(y) * nch + c];
            if (b >= 0 && size_t(b) < bw) group[size_t(b)] = c;
        }
        for (size_t b = 0; b < bw; ++b) {
            const BlobInfo &bi = fb[y][b];
            const size_t grp = group[b];
            color cc;
            if (wopt.show_blobs == DISTINCT) {  // morph.cpp:1323-1328
                std::mt19937 gen(static_cast<std::mt19937::result_type>(grp));
                std::uniform_int_distribution<unsigned char> ub(0, 255);
                cc.r = ub(gen); cc.g = ub(gen); cc.b = ub(gen); cc.a = 255;
            } else {  // blob2pixel (morph.cpp:1423-1429)
                cc = create_color(bi.r, bi.g, bi.b, bi.a);
                if (wopt.blob_delimiter == HSP) cc = hsp_to_rgb(cc);
            }
            brgba[size_t(y) * bw + b] = color_word(cc);
        }
    }
    const size_t nb = std::max<size_t>(bw, 1);
    if (!d_rank.ensure(4 * rank.size()) || !d_brgba.ensure(4 * size_t(h) * nb) || !d_lag.ensure(2048) ||
        !d_slope.ensure(2048))
        return fail("device allocation");
    if (amh_memcpy_h2d(d_rank.p, rank.data(), 4 * rank.size(), cstream) ||
        amh_memcpy_h2d(d_brgba.p, brgba.data(), 4 * brgba.size(), cstream) ||
        amh_memcpy_h2d(d_lag.p, lag_perm, 2048, cstream) || amh_memcpy_h2d(d_slope.p, slope_perm, 2048, cstream))
        return fail("snapshot upload");
    const size_t wsb = amh_render_workspace_bytes(CW, CH, wtot);
    if (d_rws.n < wsb) {
        if (!d_rws.ensure(wsb)) return fail("device allocation");
        if (amh_render_workspace_init(d_rws.p, wsb, CW, CH, wtot, cstream)) return fail("render workspace");
    }
    if (!d_out.ensure(4 * size_t(CW) * CH)) return fail("device allocation");
    if (amh_stream_sync(cstream)) return fail("snapshot sync");
    have_snap = true;
    fluid_inputs = false;  // chains moved: the fluid's per-segment inputs are rebuilt on use
    return true;
}

// The worker thread (thread::run, thread.cpp:173-206), with a condition
// variable instead of spinning.
void morph::engine::run() {
    std::unique_lock<std::mutex> lk(mu);
    size_t since_sync = 0;
    while (!stop) {
        if (paused) {
            cv.wait(lk, [&] { return stop || !paused; });
            t0 = mono_seconds();
            continue;
        }
        lk.unlock();
        step();
        if (++since_sync >= 256) {  // bound the depth of the stream
            amh_stream_sync(cstream);
            since_sync = 0;
        }
        lk.lock();
        bool pause = pause_req;
        if (iters > 0 && --iters == 0) pause = true;
        if (secs > 0.0 && mono_seconds() - t0 > secs) {
            secs = 0.0;
            pause = true;
        }
        if (pause) {
            lk.unlock();
            if (cstream) amh_stream_sync(cstream);
            since_sync = 0;
            lk.lock();
            paused = true;
            pause_req = false;
            cv.notify_all();
        }
    }
    running = false;
    paused = true;
    cv.notify_all();
}

void morph::engine::start_or_resume(size_t n, double s) {
    std::unique_lock<std::mutex> lk(mu);
    if (running && !paused) return;
    iters = n;
    secs = s;
    t0 = mono_seconds();
    pause_req = false;
    paused = false;
    if (!running) {
        running = true;
        th = std::thread(&engine::run, this);
    }
    cv.notify_all();
}

// ----------------------------------------------------------- am::morph
morph::morph() : E(new engine) {}
morph::~morph() { delete E; }

void morph::clear() {  // morph.cpp:18-49
    suspend();
    E->frames.clear();
    E->identifier = 0;
    E->energy = 0.0;
    E->bbox_x1 = E->bbox_y1 = UINT16_MAX;
    E->bbox_x2 = E->bbox_y2 = 0;
    E->wident = SIZE_MAX;
    E->reset_worker();
    E->state = STATE_BLOB_DETECTION;
    for (auto &v : E->out_blobs)
        for (blob *b : v) delete b;
    E->out_blobs.clear();
}

bool morph::add_frame(size_t f) {  // morph.cpp:288-296
    if (E->frames.count(f)) return false;
    E->frames[f];
    ++E->identifier;
    return true;
}

bool morph::add_pixel(size_t f, pixel px) {  // morph.cpp:298-337
    HostFrame &fr = E->frames[f];
    fr.px.emplace_back(uint32_t(px.y) << 16 | px.x, color_word(px.c));
    ++E->identifier;
    E->bbox_x1 = std::min(E->bbox_x1, px.x);
    E->bbox_y1 = std::min(E->bbox_y1, px.y);
    E->bbox_x2 = std::max(E->bbox_x2, px.x);
    E->bbox_y2 = std::max(E->bbox_y2, px.y);
    return true;
}

bool morph::add_frame_rgba(size_t f, const uint8_t *rgba, uint16_t w, uint16_t h, size_t stride) {
    if (!rgba || w == 0 || h == 0) return false;
    if (stride == 0) stride = size_t(w) * 4;
    HostFrame &fr = E->frames[f];
    ++E->identifier;
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t *row = rgba + y * stride;
        for (uint32_t x = 0; x < w; ++x) {
            const uint8_t *q = row + 4 * x;
            if (q[3] == 0) continue;
            fr.px.emplace_back(y << 16 | x, uint32_t(q[0]) | (uint32_t(q[1]) << 8) | (uint32_t(q[2]) << 16) |
                                                 (uint32_t(q[3]) << 24));
            E->bbox_x1 = std::min<uint16_t>(E->bbox_x1, uint16_t(x));
            E->bbox_y1 = std::min<uint16_t>(E->bbox_y1, uint16_t(y));
            E->bbox_x2 = std::max<uint16_t>(E->bbox_x2, uint16_t(x));
            E->bbox_y2 = std::max<uint16_t>(E->bbox_y2, uint16_t(y));
        }
    }
    return true;
}

size_t morph::get_pixel_count(size_t f) {
    auto it = E->frames.find(f);
    if (it == E->frames.end()) return 0;
    std::set<uint32_t> pos;
    for (auto &e : it->second.px) pos.insert(e.first);
    return pos.size();
}

size_t morph::get_frame_key(double t) {  // morph.cpp:404-423
    double integ;
    t = std::modf(t, &integ);
    if (t < 0.0) t += 1.0;
    const size_t f = size_t(t * E->frames.size());
    size_t i = 0;
    for (auto &kv : E->frames)
        if (i++ == f) return kv.first;
    return SIZE_MAX;
}

size_t morph::get_blob_count(size_t f) {
    size_t idx = 0;
    for (size_t k : E->keys) {
        if (k == f) break;
        ++idx;
    }
    if (idx >= E->keys.size() || idx >= E->fb.size()) return 0;
    if (E->state >= STATE_BLOB_MATCHING && E->matched) return E->fb[idx].size();
    return size_t(E->nblobs.size() > idx ? E->nblobs[idx] : 0);
}

size_t morph::get_blob_count() {
    size_t n = 0;
    for (size_t k : E->keys) n += get_blob_count(k);
    return n;
}

pixel morph::get_average_pixel(size_t f) {  // morph.cpp:339-355
    pixel px = create_pixel(0, 0, 0, 0, 0, 0);
    size_t idx = 0;
    for (size_t k : E->keys) {
        if (k == f) break;
        ++idx;
    }
    if (idx >= E->keys.size() || E->fmean.size() < (idx + 1) * 6) return px;
    const double *m = &E->fmean[idx * 6];
    px.x = to_u16(std::round(m[0]));
    px.y = to_u16(std::round(m[1]));
    px.c = create_color(m[2], m[3], m[4], m[5]);
    if (E->opt.blob_delimiter == HSP) px.c = hsp_to_rgb(px.c);
    return px;
}

pixel morph::get_average_pixel(size_t f, size_t b) {  // morph.cpp:357-376
    pixel px = create_pixel(0, 0, 0, 0, 0, 0);
    const blob *bl = get_blob(f, b);
    if (!bl) return px;
    px.x = to_u16(std::round(bl->x));
    px.y = to_u16(std::round(bl->y));
    px.c = create_color(bl->r, bl->g, bl->b, bl->a);
    if (E->opt.blob_delimiter == HSP) px.c = hsp_to_rgb(px.c);
    return px;
}

pixel morph::get_pixel(size_t f, size_t pos) {  // morph.cpp:378-392
    pixel px = create_pixel(0, 0, 0, 0, 0, 0);
    auto it = E->frames.find(f);
    if (it == E->frames.end()) return px;
    const uint32_t key = uint32_t((pos / 65536) << 16 | (pos % 65536));
    const auto &v = it->second.px;
    for (size_t i = v.size(); i-- > 0;) {
        if (v[i].first != key) continue;
        px.x = uint16_t(pos % 65536);
        px.y = uint16_t(pos / 65536);
        px.c = word_color(v[i].second);
        if (E->opt.blob_delimiter == HSP) px.c = hsp_to_rgb(rgb_to_hsp(px.c));
        return px;
    }
    return px;
}

// Perlin fBm on the host for get_background (perlin.cpp:19-69, z = 0).
static double host_octave(const int32_t *p, double x, double y) {
    auto fade = [](double t) { return t * t * t * (t * (t * 6 - 15) + 10); };
    auto lerp = [](double t, double a, double b) { return a + t * (b - a); };
    auto grad = [](int hash, double gx, double gy) {
        const int hh = hash & 15;
        const double u = hh < 8 ? gx : gy, v = hh < 4 ? gy : (hh == 12 || hh == 14) ? gx : 0.0;
        return ((hh & 1) == 0 ? u : -u) + ((hh & 2) == 0 ? v : -v);
    };
    double result = 0.0, amp = 1.0;
    for (int i = 0; i < 8; ++i) {
        const int X = static_cast<int>(std::floor(x)) & 255, Y = static_cast<int>(std::floor(y)) & 255;
        double fx = x - std::floor(x), fy = y - std::floor(y);
        const double u = fade(fx), v = fade(fy);
        const int A = p[X] + Y, B = p[X + 1] + Y;
        const double n = lerp(v, lerp(u, grad(p[p[A]], fx, fy), grad(p[p[B]], fx - 1, fy)),
                              lerp(u, grad(p[p[A + 1]], fx, fy - 1), grad(p[p[B + 1]], fx - 1, fy - 1)));
        result += n * amp;
        x *= 2.0;
        y *= 2.0;
        amp *= 0.5;
    }
    return result;
}

color morph::get_background(uint16_t x, uint16_t y, double time) {  // morph.cpp:1431-1465
    if (E->frames.empty()) return create_color(0.0, 0.0, 0.0, 0.0);
    time = normalize_time(time);
    const size_t fk = get_frame_key(time);
    auto it = E->frames.find(fk);
    if (it == E->frames.end()) return create_color(0.0, 0.0, 0.0, 0.0);
    ++it;
    if (it == E->frames.end()) it = E->frames.begin();
    const size_t fn = it->first;
    const double dt = 1.0 / double(E->frames.size());
    const double t = std::max(0.0, (time - (fk * dt)) / dt);
    const color c1 = get_pixel(fk, xy2pos(x, y)).c, c2 = get_pixel(fn, xy2pos(x, y)).c;
    if (E->opt.fading == PERLIN) {
        const double bw = E->bbox_x2 - E->bbox_x1 + 1.0, bh = E->bbox_y2 - E->bbox_y1 + 1.0;
        const double px = ((x - E->bbox_x1) / double(bw)) * 8.0, py = ((y - E->bbox_y1) / double(bh)) * 8.0;
        const double lag = host_octave(E->lag_perm, px, py) * 0.5 + 0.5;
        const double slope = host_octave(E->slope_perm, px, py) * 0.5 + 0.5;
        return interpolate(c1, c2, lag, slope, 1.0 - t);
    }
    if (E->opt.fading == COSINE) return interpolate(c1, c2, 0.5, 0.5, 1.0 - t);
    return interpolate(c1, c2, 1.0 - t);
}

color morph::interpolate(color c1, color c2, double lag, double slope, double str) {  // morph.cpp:1467-1479
    if (E->opt.fading == COSINE || E->opt.fading == PERLIN) {
        const double pi = 3.14159265358;
        const double s = (slope + 0.1) / 1.1;
        const double l = (1.0 - s) * lag;
        if (str <= l) str = 0.0;
        else if (str >= (l + s)) str = 1.0;
        else str = ((-std::cos((str - l) * (pi / s)) + 1.0) / 2.0);
    }
    return interpolate(c1, c2, str);
}

color morph::interpolate(color c1, color c2, double w) {  // morph.cpp:1481-1488
    color c;
    c.r = to_u8(std::round(w * double(c1.r) + (1.0 - w) * double(c2.r)));
    c.g = to_u8(std::round(w * double(c1.g) + (1.0 - w) * double(c2.g)));
    c.b = to_u8(std::round(w * double(c1.b) + (1.0 - w) * double(c2.b)));
    c.a = to_u8(std::round(w * double(c1.a) + (1.0 - w) * double(c2.a)));
    return c;
}

pixel morph::interpolate(pixel p1, pixel p2, double w) {  // morph.cpp:1490-1499
    pixel px;
    px.x = to_u16(std::round(w * double(p1.x) + (1.0 - w) * double(p2.x)));
    px.y = to_u16(std::round(w * double(p1.y) + (1.0 - w) * double(p2.y)));
    px.c = interpolate(p1.c, p2.c, w);
    return px;
}

point morph::interpolate(point pt1, point pt2, double w) {  // morph.cpp:1501-1515
    point pt;
    pt.word = 0;
    const double x1 = (256.0 * pt1.s.x + pt1.s.x_fract), y1 = (256.0 * pt1.s.y + pt1.s.y_fract);
    const double x2 = (256.0 * pt2.s.x + pt2.s.x_fract), y2 = (256.0 * pt2.s.y + pt2.s.y_fract);
    const double x = (w * x1 + (1.0 - w) * x2), y = (w * y1 + (1.0 - w) * y2);
    pt.s.x = to_u16(x / 256.0);
    pt.s.x_fract = to_u8(x - (pt.s.x * 256));
    pt.s.y = to_u16(y / 256.0);
    pt.s.y_fract = to_u8(y - (pt.s.y * 256));
    return pt;
}

void morph::set_seed(unsigned seed) {  // morph.cpp:436-444
    E->e1.seed(seed);
    E->set_perms(seed);
    E->opt.seed = seed;
}

double morph::normalize_time(double t) {  // morph.cpp:446-450
    double integ, time = std::modf(t, &integ);
    if (time < 0.0) time += 1.0;
    return time;
}

double morph::get_time(size_t fi, size_t total) {  // morph.cpp:1522-1538
    if (total == 0) return 0.0;
    if (E->opt.finite) {
        if (total == 1) {
            if (get_frame_count() == 0) return 0.0;
            return (1.0 - (1.0 / double(get_frame_count()))) / 2.0;
        }
        double t = double(fi) / double(total - 1);
        t *= 1.0 - (1.0 / double(get_frame_count()));
        return t;
    }
    return double(fi) / double(total);
}

uint16_t morph::get_width() { return E->width; }
uint16_t morph::get_height() { return E->height; }
size_t morph::get_frame_count() { return E->frames.size(); }
unsigned morph::get_state() { return E->state; }
void morph::next_state() { E->skip_state = true; }
double morph::get_energy() { return E->energy; }

pixel morph::blob2pixel(const blob *bl) {  // morph.cpp:1423-1429
    pixel px = create_pixel(to_u16(bl->x), to_u16(bl->y), create_color(bl->r, bl->g, bl->b, bl->a));
    if (E->opt.blob_delimiter == HSP) px.c = hsp_to_rgb(px.c);
    return px;
}

void morph::compute() { E->start_or_resume(0, 0.0); }
void morph::iterate(size_t n) {
    if (n) E->start_or_resume(n, 0.0);
}
void morph::compute(double s) { E->start_or_resume(0, s); }

void morph::suspend() {
    std::unique_lock<std::mutex> lk(E->mu);
    if (!E->running || E->paused) return;
    E->pause_req = true;
    E->cv.wait(lk, [&] { return E->paused || !E->running; });
}

bool morph::suspend(double timeout) {
    std::unique_lock<std::mutex> lk(E->mu);
    if (!E->running || E->paused) return true;
    E->pause_req = true;
    return E->cv.wait_for(lk, std::chrono::duration<double>(timeout), [&] { return E->paused || !E->running; });
}

void morph::wait() {
    std::unique_lock<std::mutex> lk(E->mu);
    E->cv.wait(lk, [&] { return E->paused || !E->running; });
}

void morph::restart() { ++E->identifier; }

bool morph::iterate_timed(size_t steps, float *gpu_ms) {
    engine &e = *E;
    if (is_busy() || !e.ensure_streams()) return false;
    void *timer = nullptr;
    if (amh_timer_begin(&timer, e.cstream)) return e.fail("timer");
    iterate(steps);
    wait();
    float ms = 0.0f;
    const int r = amh_timer_end(timer, e.cstream, &ms);
    amh_timer_free(timer);
    if (r) return e.fail("timer");
    if (gpu_ms) *gpu_ms = ms;
    return true;
}

bool morph::render_device(const double *times, size_t n, void *out_dev, float *gpu_ms) {
    engine &e = *E;
    if (!e.have_snap || !out_dev || !times) {
        e.err = "render_device: no synchronized chains or NULL buffer";
        return false;
    }
    amh_render r;
    e.fill_render(r);
    const bool fluid = e.opt.fluidsteps > 0;
    if (fluid && !e.prepare_fluid()) return false;
    if (!fluid && n > 1) {
        // several atom-path frames: one captured frame replayed for each
        // (kept while the render arguments and the workspace stay the same)
        std::vector<unsigned char> key(sizeof(r) + sizeof(void *) + sizeof(size_t));
        std::memcpy(key.data(), &r, sizeof(r));  // fill_render zeroes the padding
        std::memcpy(key.data() + sizeof(r), &e.d_rws.p, sizeof(void *));
        std::memcpy(key.data() + sizeof(r) + sizeof(void *), &e.d_rws.n, sizeof(size_t));
        if (e.rgraph && key != e.rgraph_key) {
            amh_render_graph_destroy(e.rgraph);
            e.rgraph = nullptr;
        }
        if (!e.rgraph) {
            if (amh_render_graph_create(&e.rgraph, &r, e.d_rws.p, e.d_rws.n, e.rstream) == AMH_OK)
                e.rgraph_key = key;
            else
                e.rgraph = nullptr;  // frames run eagerly
        }
    }
    void *timer = nullptr;
    if (amh_timer_begin(&timer, e.rstream)) return e.fail("timer");
    const size_t npx = size_t(e.CW) * e.CH;
    for (size_t i = 0; i < n; ++i) {
        uint32_t *dst = static_cast<uint32_t *>(out_dev) + i * npx;
        bool ok;
        if (fluid) ok = e.render_fluid(times[i], dst);
        else if (n > 1 && e.rgraph) ok = amh_render_graph_frame(e.rgraph, times[i], dst, e.rstream) == AMH_OK;
        else ok = amh_render_frame(&r, times[i], e.d_rws.p, e.d_rws.n, dst, e.rstream) == AMH_OK;
        if (!ok) {
            amh_timer_free(timer);
            return fluid ? false : e.fail("render");
        }
    }
    float ms = 0.0f;
    const int rc = amh_timer_end(timer, e.rstream, &ms);
    amh_timer_free(timer);
    if (rc) return e.fail("timer");
    if (gpu_ms) *gpu_ms = ms;
    return true;
}

bool morph::is_busy() const {
    std::unique_lock<std::mutex> lk(E->mu);
    return E->running && !E->paused;
}

bool morph::synchronize() {  // morph.cpp:102-269
    if (is_busy()) return false;
    engine &e = *E;
    if (e.opt.seed != e.wopt.seed) e.rng.seed(e.opt.seed);
    e.wopt = e.opt;
    if (e.identifier != e.wident) {
        e.reset_worker();
        if (!e.upload_frames()) return false;
        e.wident = e.identifier;
        for (auto &v : e.out_blobs)
            for (blob *b : v) delete b;
        e.out_blobs.clear();
    }
    e.state = e.wstate;
    e.energy = -1.0;
    if (e.wstate == STATE_BLOB_MATCHING) e.energy = e.best_report;  // best_blob_map_e (thread.cpp:1178)
    if (e.wstate >= STATE_ATOM_MORPHING && e.wtot > 0) {
        if (!e.snapshot()) return false;
        uint64_t lo, hi;
        if (e.wstate == STATE_ATOM_MORPHING && e.exact_cost(&lo, &hi))
            e.energy = double(hi) * 18446744073709551616.0 + double(lo);
        amh_refine_stats st;
        if (amh_memcpy_d2h(&st, e.d_rstats.p, sizeof(st), e.cstream) == AMH_OK && amh_stream_sync(e.cstream) == AMH_OK) {
            e.counters[0] = st.attempts;
            e.counters[1] = st.accepted;
            e.counters[2] = uint64_t(st.gain);
        }
    }
    e.blobs_dirty = true;
    if (e.skip_state) {
        e.wskip = true;
        e.skip_state = false;
    }
    return true;
}

const blob *morph::get_blob(size_t f, size_t b) {
    engine &e = *E;
    size_t y = 0;
    for (size_t k : e.keys) {
        if (k == f) break;
        ++y;
    }
    if (y >= e.keys.size() || y >= e.fb.size() || b >= e.fb[y].size()) return nullptr;
    if (e.blobs_dirty || e.out_blobs.size() != e.fb.size()) {
        for (auto &v : e.out_blobs)
            for (blob *bb : v) delete bb;
        e.out_blobs.assign(e.fb.size(), {});
        e.blobs_dirty = false;
    }
    auto &v = e.out_blobs[y];
    if (v.empty()) {
        // surfaces from the blob-of-pixel map of this frame
        const size_t npx = size_t(e.CW) * e.CH;
        std::vector<int32_t> bop(npx);
        if (amh_memcpy_d2h(bop.data(), e.d_bop.as<int32_t>() + y * npx, npx * 4, e.rstream) ||
            amh_stream_sync(e.rstream))
            return nullptr;
        std::vector<std::vector<size_t>> surf(e.nblobs[y]);
        for (size_t i = 0; i < npx; ++i)
            if (bop[i] >= 0) surf[bop[i]].push_back(xy2pos(uint16_t(i % e.CW), uint16_t(i / e.CW)));
        const size_t nch = e.cw.size();
        for (size_t i = 0; i < e.fb[y].size(); ++i) {
            const BlobInfo &bi = e.fb[y][i];
            blob *bl = new blob;
            bl->index = i;
            bl->x = bi.x; bl->y = bi.y; bl->r = bi.r; bl->g = bi.g; bl->b = bi.b; bl->a = bi.a;
            bl->unified = true;
            // the reference's blob->group: the start pixel before matching,
            // the column of the best blob map while matching (thread.cpp:
            // 655, 713-730), the chain afterwards
            bl->group = bi.group0;
            if (e.matched && e.best_map.size() == e.bw * size_t(e.h))
                for (size_t c = 0; c < e.bw; ++c)
                    if (e.best_map[c * e.h + y] == int32_t(i)) bl->group = c;
            for (size_t c = 0; c < nch; ++c)
                if (e.rank.size() == size_t(e.h) * nch && e.rank[y * nch + c] == int32_t(i)) bl->group = c;
            if (!bi.empty && bi.id >= 0) bl->surface.insert(surf[bi.id].begin(), surf[bi.id].end());
            v.push_back(bl);
        }
    }
    return b < v.size() ? v[b] : nullptr;
}

bool morph::get_frame_rgba(double t, uint8_t *out) {
    engine &e = *E;
    const int W = e.width ? e.width : e.CW, H = e.height ? e.height : e.CH;
    if (!out || W <= 0 || H <= 0) return false;
    std::vector<uint32_t> img;
    const size_t npx = size_t(e.CW) * e.CH;
    if (e.have_snap) {
        amh_render r;
        e.fill_render(r);
        if (e.opt.fading == PERLIN && (e.opt.seed != e.wopt.seed)) {  // perms follow the caller's seed
            amh_memcpy_h2d(e.d_lag.p, e.lag_perm, 2048, e.rstream);
            amh_memcpy_h2d(e.d_slope.p, e.slope_perm, 2048, e.rstream);
        }
        if (e.opt.fluidsteps > 0) {  // draw_fluid (morph.cpp:1418)
            if (!e.render_fluid(t, e.d_out.as<uint32_t>())) return false;
        } else if (amh_render_frame(&r, t, e.d_rws.p, e.d_rws.n, e.d_out.as<uint32_t>(), e.rstream)) {
            return e.fail("render");
        }
        img.resize(npx);
        if (amh_memcpy_d2h(img.data(), e.d_out.p, npx * 4, e.rstream) || amh_stream_sync(e.rstream))
            return e.fail("render download");
    } else {
        // no chains yet: the key frame's pixels over the background
        // (get_pixels(blob) with chains.empty(), morph.cpp:468-474)
        img.assign(npx, 0u);
        const double tn = normalize_time(t);
        const size_t fk = e.keys.empty() ? SIZE_MAX : size_t(tn * e.keys.size());
        for (int y = 0; y < e.CH; ++y)
            for (int x = 0; x < e.CW; ++x) {
                uint32_t c = 0;
                if (e.opt.keep_background) c = color_word(get_background(uint16_t(x), uint16_t(y), t));
                if (e.state >= STATE_BLOB_UNIFICATION && fk < e.canvas.size()) {
                    const uint32_t k = e.canvas[fk][size_t(y) * e.CW + x];
                    if (k >> 24) {
                        color kc = word_color(k);
                        if (e.opt.blob_delimiter == HSP) kc = hsp_to_rgb(rgb_to_hsp(kc));
                        if (e.opt.keep_background) {
                            const color bg = word_color(c);
                            const double a = kc.a / 255.0, bga = bg.a / 255.0;
                            kc = create_color(a * (kc.r / 255.0) + (1.0 - a) * (bg.r / 255.0),
                                              a * (kc.g / 255.0) + (1.0 - a) * (bg.g / 255.0),
                                              a * (kc.b / 255.0) + (1.0 - a) * (bg.b / 255.0),
                                              bga + (1.0 - bga) * a);
                        }
                        c = color_word(kc);
                    }
                }
                img[size_t(y) * e.CW + x] = c;
            }
    }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const uint32_t c = (x < e.CW && y < e.CH) ? img[size_t(y) * e.CW + x] : 0u;
            std::memcpy(out + (size_t(y) * W + x) * 4, &c, 4);
        }
    return true;
}

void morph::get_pixels(double t, std::vector<pixel> *to) {  // morph.cpp:1405-1421
    to->clear();
    const int W = E->width ? E->width : E->CW, H = E->height ? E->height : E->CH;
    if (W <= 0 || H <= 0) return;
    std::vector<uint8_t> img(size_t(W) * H * 4);
    if (!get_frame_rgba(t, img.data())) std::fill(img.begin(), img.end(), 0);
    to->reserve(size_t(W) * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const uint8_t *q = &img[(size_t(y) * W + x) * 4];
            to->push_back(create_pixel(uint16_t(x), uint16_t(y), q[0], q[1], q[2], q[3]));
        }
}

const blob *morph::get_pixels(size_t b, double t, std::vector<pixel> *to) {  // morph.cpp:452-678
    const size_t fk = get_frame_key(normalize_time(t));
    const blob *bl = get_blob(fk, b);
    if (!bl) return nullptr;
    engine &e = *E;
    if (!e.have_snap) {
        for (size_t pos : bl->surface) to->push_back(get_pixel(fk, pos));
        return bl;
    }
    // render this blob alone: its resolved pixels with alpha > 0
    std::vector<int32_t> solo(e.rank.size(), -1);
    const size_t nch = e.cw.size();
    size_t y = 0;
    for (size_t k : e.keys) {
        if (k == fk) break;
        ++y;
    }
    for (size_t c = 0; c < nch; ++c)
        if (e.rank[y * nch + c] == int32_t(b)) solo[y * nch + c] = int32_t(b);
    amh_memcpy_h2d(e.d_rank.p, solo.data(), 4 * solo.size(), e.rstream);
    const Options saved = e.opt;
    e.opt.keep_background = false;
    e.opt.show_blobs = TEXTURE;
    e.opt.blend_blobs = false;
    std::vector<uint8_t> img(size_t(e.width ? e.width : e.CW) * (e.height ? e.height : e.CH) * 4);
    const bool ok = get_frame_rgba(t, img.data());
    e.opt = saved;
    amh_memcpy_h2d(e.d_rank.p, e.rank.data(), 4 * e.rank.size(), e.rstream);
    amh_stream_sync(e.rstream);
    if (!ok) return bl;
    const int W = e.width ? e.width : e.CW;
    for (size_t i = 0; i < img.size() / 4; ++i)
        if (img[4 * i + 3]) to->push_back(create_pixel(uint16_t(i % W), uint16_t(i / W), img[4 * i], img[4 * i + 1],
                                                       img[4 * i + 2], img[4 * i + 3]));
    return bl;
}

void morph::set_blob_delimiter(unsigned char d) { E->opt.blob_delimiter = d; }
void morph::set_blob_threshold(double t) { E->opt.blob_threshold = t; }
void morph::set_blob_max_size(size_t s) { E->opt.blob_max_size = s; }
void morph::set_blob_min_size(size_t s) { E->opt.blob_min_size = s; }
void morph::set_blob_box_grip(uint16_t g) { E->opt.blob_box_grip = g; }
void morph::set_blob_box_samples(size_t s) { E->opt.blob_box_samples = s; }
void morph::set_blob_number(size_t n) { E->opt.blob_number = n; }
void morph::set_blob_rgba_weight(unsigned char w) { E->opt.blob_rgba_weight = w; }
void morph::set_blob_size_weight(unsigned char w) { E->opt.blob_size_weight = w; }
void morph::set_blob_xy_weight(unsigned char w) { E->opt.blob_xy_weight = w; }
void morph::set_degeneration(size_t d) { E->opt.degeneration = d; }
void morph::set_motion(unsigned char m) { E->opt.motion = m; }
void morph::set_fading(unsigned char f) { E->opt.fading = f; }
void morph::set_threads(size_t t) { E->opt.threads = t; }
void morph::set_cycle_length(size_t c) { E->opt.cycle_length = c; }
void morph::set_feather(size_t f) { E->opt.feather = f; }
void morph::set_keep_background(bool k) { E->opt.keep_background = k; }
void morph::set_finite(bool f) { E->opt.finite = f; }
void morph::set_show_blobs(unsigned b) { E->opt.show_blobs = b; }

void morph::set_fluid(unsigned f) {  // morph.cpp:1553-1560
    if (E->opt.fluidsteps != f) {
        E->opt.fluidsteps = f;
        if (E->opt.fluidsteps > 0 && E->opt.density > 1) set_density(1);
    }
}
void morph::set_density(uint16_t d) {  // morph.cpp:1562-1570
    if (E->opt.density != d) {
        E->opt.density = d;
        ++E->identifier;
        if (E->opt.fluidsteps != 0 && E->opt.density > 1) set_fluid(0);
    }
}
void morph::set_resolution(uint16_t w, uint16_t h) {
    E->width = w;
    E->height = h;
}

void morph::set_refine(unsigned mode, unsigned rounds, int tile, unsigned graph_passes) {
    E->opt.refine_mode = mode;
    E->opt.refine_rounds = rounds;
    E->opt.refine_tile = tile;
    E->opt.graph_passes = graph_passes;
}
void morph::set_match_budget(size_t steps) { E->opt.match_budget = steps; }
void morph::set_learned(const float *params) {
    if (params) {
        E->mlp_params.assign(params, params + AMH_MLP_PARAMS);
        E->mlp_version += 1;
    }
    E->opt.learned = params != nullptr;
}

bool morph::warm_start(float *gpu_ms) {
    engine &e = *E;
    if (is_busy() || e.wstate != STATE_ATOM_MORPHING) { e.err = "warm_start: not paused in ATOM_MORPHING"; return false; }
    if (!e.prepare_mlp() || amh_stream_sync(e.cstream)) return false;
    void *timer = nullptr;
    if (gpu_ms && amh_timer_begin(&timer, e.cstream)) return false;
    bool ok = e.warm_start();
    if (gpu_ms) ok = amh_timer_end(timer, e.cstream, gpu_ms) == 0 && ok;
    amh_timer_free(timer);
    return ok && amh_stream_sync(e.cstream) == 0 && e.snapshot();
}

bool morph::autotune(double budget_ms, int *tile, unsigned *rounds, std::vector<tune_entry> *report) {
    engine &e = *E;
    if (is_busy() || e.wstate != STATE_ATOM_MORPHING) { e.err = "autotune: not paused in ATOM_MORPHING"; return false; }
    std::vector<amh_tune_entry> rep;
    const int fixed = rounds && *rounds > 0 ? int(*rounds) : 0;
    if (!e.autotune(budget_ms, fixed, &rep)) return false;
    e.tuned_fixed = fixed;
    if (tile) *tile = e.tuned_tile;
    if (rounds) *rounds = unsigned(e.tuned_rounds);
    if (report) {
        report->clear();
        for (const amh_tune_entry &r : rep)
            report->push_back({r.tile, unsigned(r.rounds), r.passes, r.ms, r.gain_at_budget});
    }
    return true;
}

size_t morph::learn_samples(std::vector<float> *x, std::vector<float> *t) {
    engine &e = *E;
    if (!e.have_snap || e.wtot == 0 || e.h < 2) return 0;
    const size_t n = size_t(e.wtot) * size_t(e.h);
    if (!x && !t) return n;
    const size_t bytes = amh_warm_workspace_bytes(e.wtot, int32_t(e.cw.size()), e.h);
    if (!e.d_wws.ensure(bytes) || !e.d_lx.ensure(size_t(e.wtot) * 32) || !e.d_lt.ensure(size_t(e.wtot) * 8)) return 0;
    amh_render r;
    e.fill_render(r);
    if (x) x->resize(n * 8);
    if (t) t->resize(n * 2);
    for (int y = 0; y < e.h; ++y) {  // row y -> row y + 1 (cyclically)
        if (amh_learn_features(&r, y, e.d_lx.as<float>(), e.d_lt.as<float>(), e.d_wws.p, bytes, e.rstream)) return 0;
        const size_t o = size_t(y) * size_t(e.wtot);
        if ((x && amh_memcpy_d2h(x->data() + o * 8, e.d_lx.p, size_t(e.wtot) * 32, e.rstream)) ||
            (t && amh_memcpy_d2h(t->data() + o * 2, e.d_lt.p, size_t(e.wtot) * 8, e.rstream)) ||
            amh_stream_sync(e.rstream))
            return 0;
    }
    return n;
}
void morph::set_fluid_sparseness(double s) {
    if (s > 0.0) E->opt.fluid_sparseness = s;
}
void morph::get_refine_counters(uint64_t *attempts, uint64_t *accepted) {
    if (attempts) *attempts = E->counters[0];
    if (accepted) *accepted = E->counters[1];
}
bool morph::transport_cost(uint64_t *lo, uint64_t *hi) { return E->exact_cost(lo, hi); }

size_t morph::get_chain_points(std::vector<uint64_t> *pts, size_t *height, std::vector<size_t> *widths) {
    engine &e = *E;
    if (height) *height = size_t(e.h);
    if (widths) widths->assign(e.cw.begin(), e.cw.end());
    if (pts && e.have_snap) {
        pts->resize(size_t(e.wtot) * e.h);
        if (amh_memcpy_d2h(pts->data(), e.d_snap.p, 8 * pts->size(), e.rstream) || amh_stream_sync(e.rstream))
            pts->clear();
    }
    return size_t(e.wtot);
}

bool morph::put_chain_points(const std::vector<uint64_t> &pts) {
    engine &e = *E;
    if (is_busy() || pts.size() != size_t(e.wtot) * e.h || e.wtot == 0) return false;
    if (amh_memcpy_h2d(e.d_pts.p, pts.data(), 8 * pts.size(), e.cstream) || amh_stream_sync(e.cstream)) return false;
    return e.snapshot();
}

// Chain checkpoint file (no counterpart in the reference, which keeps no
// state on disk), little-endian:
//   "AMCHAIN1"  u32 version = 1, u32 key frames h, u32 chains, u32 0,
//   u64 atoms, u32 width[chains], u64 point[h][atoms] (key-frame-major),
//   u64 FNV-1a 64 of every byte before it.
namespace {
constexpr char kChainMagic[8] = {'A', 'M', 'C', 'H', 'A', 'I', 'N', '1'};
uint64_t fnv1a(const uint8_t *p, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
    for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 0x100000001b3ull;
    return h;
}
template <class T> void put_le(std::vector<uint8_t> &v, T x) {
    for (size_t k = 0; k < sizeof(T); ++k) v.push_back(uint8_t(uint64_t(x) >> (8 * k)));
}
template <class T> T get_le(const uint8_t *p) {
    uint64_t x = 0;
    for (size_t k = 0; k < sizeof(T); ++k) x |= uint64_t(p[k]) << (8 * k);
    return T(x);
}
}  // namespace

bool morph::save_chains(const char *path) {
    engine &e = *E;
    std::vector<uint64_t> pts;
    size_t h = 0;
    std::vector<size_t> widths;
    if (!path || !e.have_snap || get_chain_points(&pts, &h, &widths) == 0 || pts.empty()) {
        e.err = "save_chains: no synchronized chains";
        return false;
    }
    std::vector<uint8_t> out(kChainMagic, kChainMagic + 8);
    put_le<uint32_t>(out, 1);
    put_le<uint32_t>(out, uint32_t(h));
    put_le<uint32_t>(out, uint32_t(widths.size()));
    put_le<uint32_t>(out, 0);
    put_le<uint64_t>(out, uint64_t(e.wtot));
    for (size_t w : widths) put_le<uint32_t>(out, uint32_t(w));
    const size_t head = out.size();
    out.resize(head + pts.size() * 8);
    for (size_t i = 0; i < pts.size(); ++i)
        for (int k = 0; k < 8; ++k) out[head + i * 8 + k] = uint8_t(pts[i] >> (8 * k));
    put_le<uint64_t>(out, fnv1a(out.data(), out.size()));
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        e.err = std::string("save_chains: cannot create ") + path;
        return false;
    }
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (std::fclose(f) != 0 || !ok) {
        e.err = std::string("save_chains: write failed: ") + path;
        return false;
    }
    return true;
}

bool morph::load_chains(const char *path) {
    engine &e = *E;
    if (!path || is_busy() || e.wstate != STATE_ATOM_MORPHING || e.wtot == 0) {
        e.err = "load_chains: not paused in ATOM_MORPHING";
        return false;
    }
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        e.err = std::string("load_chains: cannot open ") + path;
        return false;
    }
    std::vector<uint8_t> in;
    uint8_t buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) in.insert(in.end(), buf, buf + n);
    std::fclose(f);
    auto bad = [&](const char *why) {
        e.err = std::string("load_chains: ") + why;
        return false;
    };
    if (in.size() < 40 || std::memcmp(in.data(), kChainMagic, 8) != 0) return bad("not a chain file");
    if (fnv1a(in.data(), in.size() - 8) != get_le<uint64_t>(in.data() + in.size() - 8)) return bad("checksum mismatch");
    const uint32_t version = get_le<uint32_t>(in.data() + 8), h = get_le<uint32_t>(in.data() + 12);
    const uint32_t nch = get_le<uint32_t>(in.data() + 16);
    const uint64_t atoms = get_le<uint64_t>(in.data() + 24);
    if (version != 1) return bad("unknown version");
    if (h != uint32_t(e.h) || nch != e.cw.size() || atoms != uint64_t(e.wtot))
        return bad("the file's chains do not have the engine's shape");
    const size_t head = 32 + size_t(nch) * 4;
    if (in.size() != head + size_t(h) * atoms * 8 + 8) return bad("truncated file");
    for (uint32_t c = 0; c < nch; ++c)
        if (get_le<uint32_t>(in.data() + 32 + 4 * c) != uint32_t(e.cw[c]))
            return bad("the file's chain widths differ from the engine's");
    std::vector<uint64_t> pts(size_t(h) * atoms);
    for (size_t i = 0; i < pts.size(); ++i) pts[i] = get_le<uint64_t>(in.data() + head + 8 * i);
    if (!put_chain_points(pts)) return bad("upload failed");
    return true;
}

size_t morph::get_fluid_particles(std::vector<double> *out) {
    engine &e = *E;
    if (!e.fluid || e.wtot == 0) return 0;
    if (out) {
        out->resize(size_t(e.wtot) * 20);
        if (amh_stream_sync(e.rstream) || amh_fluid_get_particles(e.fluid, out->data(), e.rstream)) {
            out->clear();
            return 0;
        }
    }
    return size_t(e.wtot);
}

const char *morph::last_error() { return E->err.c_str(); }

}  // namespace am
