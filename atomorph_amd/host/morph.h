/*
 * am::morph -- the morph engine, with the public interface of the
 * reference's morph.h (morph.h:9-76, defaults morph.h:97-119).
 *
 * The worker runs the same state machine as the reference's am::thread
 * (thread.cpp:139-206): blob detection, unification, matching, atom
 * morphing, done.  Detection, chain building, refinement and rendering are
 * HIP kernels on one MI355X reached through include/atomorph_hip.h; the
 * small blob matching runs on the host.
 *
 * Semantics kept from the reference: setters are plain writes that reach the
 * worker at synchronize() while it is paused; compute()/iterate()/compute(s)
 * are non-blocking; suspend() blocks; synchronize() returns false while
 * busy; get_pixels() renders the state of the last synchronize().
 * Differences: get_energy() in ATOM_MORPHING is the exact transport cost of
 * the chains (the reference's running value is not the cost); with the
 * default blob options the blobs are the 4-connected components of present
 * pixels, found on the device, listed in label order with exact means (the
 * reference's merge gives the same blobs in a shuffled order with running
 * means); any other blob option runs the reference's greedy merge and dust
 * unification on the host with its random draws; detection and unification
 * take one worker step; frames live on a canvas of max(resolution, pixel
 * bounding box).
 */
#ifndef AM_MORPH_H
#define AM_MORPH_H

#include <cstddef>
#include <cstdint>
#include <vector>

#include "atomorph.h"

namespace am {

// One candidate measured by morph::autotune.
struct tune_entry {
    int      tile;            // batched kernel variant
    unsigned rounds;          // rounds per pass
    int      passes;          // passes run to reach the budget
    double   ms;              // their device time
    double   gain_at_budget;  // exact cost decrease at the budget
};

class morph {
  public:
    morph();
    ~morph();
    morph(const morph &) = delete;
    morph &operator=(const morph &) = delete;

    void        clear              ();
    bool        add_pixel          (size_t frame, pixel px);
    bool        add_frame          (size_t frame);
    size_t      get_pixel_count    (size_t frame);
    size_t      get_frame_key      (double t);
    size_t      get_blob_count     (size_t frame);
    size_t      get_blob_count     ();
    pixel       get_average_pixel  (size_t frame);
    pixel       get_average_pixel  (size_t frame, size_t blob);
    pixel       get_pixel          (size_t frame, size_t position);
    color       get_background     (uint16_t x, uint16_t y, double t);
    point       interpolate        (point pt1, point pt2, double pt1_weight);
    pixel       interpolate        (pixel px1, pixel px2, double px1_weight);
    color       interpolate        (color c1, color c2, double c1_weight);
    color       interpolate        (color c1, color c2, double lag, double slope, double c1_weight);
    const blob* get_pixels         (size_t blob, double t, std::vector<pixel> *to);
    void        get_pixels         (double t, std::vector<pixel> *to);
    void        set_seed           (unsigned seed);
    double      get_time           (size_t current_frame, size_t total_frames);
    double      normalize_time     (double t);

    uint16_t    get_width       ();
    uint16_t    get_height      ();
    size_t      get_frame_count ();
    unsigned    get_state       ();
    void        next_state      ();
    const blob* get_blob        (size_t f, size_t b);
    double      get_energy      ();
    pixel       blob2pixel      (const blob *bl);
    void        compute         ();                  // If not busy yet, sets worker thread to run indefinitely.
    void        iterate         (size_t iterations); // If not busy yet, runs the defined number of iterations.
    void        compute         (double seconds);    // If not busy yet, runs for the defined number of seconds.
    void        suspend         ();                  // Blocks until the worker is paused.
    bool        suspend         (double timeout);    // Blocks for timeout seconds at most; false on timeout.
    bool        is_busy         () const;

    bool        synchronize     (); // Synchronizes with the worker; false while it is busy.

    void        set_blob_delimiter  (unsigned char d);
    void        set_blob_threshold  (double        t);
    void        set_blob_max_size   (size_t        s);
    void        set_blob_min_size   (size_t        s);
    void        set_blob_box_grip   (uint16_t      g);
    void        set_blob_box_samples(size_t        s);
    void        set_blob_number     (size_t        n);
    void        set_blob_rgba_weight(unsigned char w);
    void        set_blob_size_weight(unsigned char w);
    void        set_blob_xy_weight  (unsigned char w);
    void        set_degeneration    (size_t        d);
    void        set_motion          (unsigned char m);
    void        set_fading          (unsigned char f);
    void        set_threads         (size_t        t);
    void        set_cycle_length    (size_t        c);
    void        set_feather         (size_t        f);
    void        set_keep_background (bool          k);
    void        set_finite          (bool          f);
    void        set_show_blobs      (unsigned      b);

    // These change identifier, causing the morph to restart:
    void        set_fluid           (unsigned      f);
    void        set_density         (uint16_t      d);

    void        set_resolution      (uint16_t w, uint16_t h);

    // ---- extensions of this build -------------------------------------
    // Blocks until the worker pauses by itself (end of an iterate(n) or
    // compute(seconds) budget); unlike suspend() it does not ask it to stop.
    void        wait                ();
    // Restart from blob detection at the next synchronize() (what adding a
    // pixel does in the reference), keeping frames and options.
    void        restart             ();
    // iterate(steps) and wait, with HIP events around the work on the
    // engine's compute stream: *gpu_ms = device time of those steps.
    bool        iterate_timed       (size_t steps, float *gpu_ms);
    // Renders n frames at times[i] into out_dev (n * width * height RGBA8,
    // device memory, left in HBM); *gpu_ms = device time on the render
    // stream.
    bool        render_device       (const double *times, size_t n, void *out_dev, float *gpu_ms);
    // Bulk ingest of a whole RGBA8 frame (alpha 0 = no pixel), like the
    // CLI's fill_morph (demo/main.cpp:96-127) in one call.
    bool        add_frame_rgba      (size_t frame, const uint8_t *rgba, uint16_t w, uint16_t h, size_t stride);
    // Renders time t into width*height RGBA8 (what get_pixels(t) produces).
    bool        get_frame_rgba      (double t, uint8_t *out);
    // Exact transport cost of the synchronized chains (128-bit).
    bool        transport_cost      (uint64_t *lo, uint64_t *hi);
    // Refinement schedule: REFINE_PLAIN (default; `rounds` rounds per chain
    // visit, 0 = cycle_length attempts like the reference) or REFINE_BATCHED
    // (`rounds` rounds per pass, kernel variant `tile`; rounds 0 or tile < 0
    // = autotuned on the first pass over new chains, see autotune) with
    // `graph_passes` passes per hipGraph replay (0 = no graph), or REFINE_KNN
    // (batched passes whose swap partners come from a nearest-neighbour
    // search of the next key frame's row, as the reference's OpenCV build
    // draws them; `tile` 0-3, default 3; `rounds` per pass, default 16; no
    // graph).
    void        set_refine          (unsigned mode, unsigned rounds, int tile, unsigned graph_passes);
    // Autotunes the batched descent on the current chains now (paused in
    // ATOM_MORPHING): every kernel variant x rounds per pass races from the
    // chains' present state for budget_ms of device time on a scratch copy
    // (the chains are not changed); the largest exact cost decrease wins.
    // With *rounds > 0 on input only that rounds value is raced (the tile is
    // tuned for it).  The result is used by batched passes with automatic
    // tile or rounds and remembered for chains of the same shapes.  report
    // may be NULL.
    bool        autotune            (double budget_ms, int *tile, unsigned *rounds,
                                     std::vector<tune_entry> *report);
    // Matching ends by itself after this many steps (0 = only at
    // best energy 0 or next_state(), as in the reference).
    void        set_match_budget    (size_t steps);
    // Fluid grid cells per pixel (the reference fixes it at 1: a
    // (width+20) x (height+20) grid); 0.25 gives a quarter-resolution grid.
    void        set_fluid_sparseness(double s);
    // Swap attempts and accepted swaps so far (synchronized).
    void        get_refine_counters (uint64_t *attempts, uint64_t *accepted);
    // Atom trajectories of the synchronized chains: total atoms, key frames,
    // and (if pts != NULL) the points key-frame-major [h][atoms].
    size_t      get_chain_points    (std::vector<uint64_t> *pts, size_t *height,
                                     std::vector<size_t> *widths);
    // Replace the chains' refined points (same widths), e.g. a warm start.
    bool        put_chain_points    (const std::vector<uint64_t> &pts);
    // Chain checkpoint (no counterpart in the reference): save_chains writes
    // the synchronized chains to a file; load_chains (paused in
    // ATOM_MORPHING, on chains of the same shapes: same key frames, seed
    // and matching) puts them back, so a refinement can resume.  The file
    // carries the shapes and a checksum; a mismatch is an error.
    bool        save_chains         (const char *path);
    bool        load_chains         (const char *path);
    // Learned warm start (no counterpart in the reference): with params
    // (AMH_MLP_PARAMS floats, see atomorph_hip.h) the chains are reordered by
    // the MLP's sort-and-match right after they are built; NULL turns it off.
    void        set_learned         (const float *params);
    // Applies the warm start to the current chains now (paused in
    // ATOM_MORPHING; synchronizes).  gpu_ms: device time, may be NULL.
    bool        warm_start          (float *gpu_ms);
    // Training samples from the synchronized chains: for every key frame y
    // and atom, 8 features and the 2 normalised coordinates the atom holds
    // in key frame y+1 (cyclically).  Returns the count (h * atoms).
    size_t      learn_samples       (std::vector<float> *x, std::vector<float> *t);
    // Fluid particles of the last rendered fluid frame, 20 doubles each (x y
    // u v gravity_x gravity_y freedom_r R G B A r g b a strength active
    // mature - -, grid coordinates); returns the count (0: no fluid model).
    size_t      get_fluid_particles (std::vector<double> *out);
    const char *last_error          ();

    struct engine;

  private:
    engine *E;
};

}  // namespace am

#endif
