/*
 * atomorph_hip.h -- C ABI of the MI355X (gfx950) morph engine.
 *
 * The drop-in boundary of the atomorph morph-engine hot path.  The C++ class
 * am::morph (atomorph_amd/host/morph.h, same interface as the reference's
 * morph.h:9-76) calls these entry points; any other host language binds them
 * directly (ctypes, cgo, JNI: plain pointers and sizes only).
 *
 * Conventions
 *  - Every entry point returns AMH_OK (0) or a negative AMH_E_* code and never
 *    throws.  amh_last_error() returns the calling thread's last message.
 *  - Buffers named *_dev are device (HBM) pointers; all others are host.
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).  Work
 *    is enqueued asynchronously unless the comment says "synchronous".
 *  - A point is the reference's 8-byte am::point word (atomorph.h:275-284):
 *      bits 0-15 x, 16-31 y, 32-39 x_fract, 40-47 y_fract, 48-55 flags.
 *  - Chain storage on the device is key-frame-major: the point of atom a in
 *    key frame y is pts_dev[y * wtot + a], where atom a = off[c] + x is slot x
 *    of chain c and wtot = sum of all chain widths.  (The reference keeps
 *    points[x][y] per chain: atomorph.h:286-298, atomorph.cpp:1085-1157.)
 *  - Colours are packed RGBA8 words: r | g<<8 | b<<16 | a<<24.
 */
#ifndef ATOMORPH_HIP_H
#define ATOMORPH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMH_OK 0
#define AMH_E_ARG (-1)   /* invalid argument or size                       */
#define AMH_E_HIP (-2)   /* HIP runtime error (message in amh_last_error)  */
#define AMH_E_NOMEM (-3) /* device allocation failed                       */

/* ------------------------------------------------------------ plumbing -- */
/* Library version and the calling thread's last error message. */
const char *amh_version(void);
const char *amh_last_error(void);
/* Device 0 properties (synchronous). */
int amh_device_info(int32_t *compute_units, uint64_t *hbm_bytes, char *name, size_t name_len);
int amh_malloc(void **ptr_dev, size_t bytes);
int amh_free(void *ptr_dev);
int amh_memcpy_h2d(void *dst_dev, const void *src, size_t bytes, void *stream);
int amh_memcpy_d2h(void *dst, const void *src_dev, size_t bytes, void *stream);
int amh_memcpy_d2d(void *dst_dev, const void *src_dev, size_t bytes, void *stream);
int amh_memset(void *dst_dev, int value, size_t bytes, void *stream);
int amh_stream_create(void **stream);
int amh_stream_destroy(void *stream);
int amh_stream_sync(void *stream);
/* Event timer on `stream`: amh_timer_begin records a start event into
 * *timer (allocating it on first use), amh_timer_end records the stop event,
 * waits for it and returns the elapsed milliseconds. */
int amh_timer_begin(void **timer, void *stream);
int amh_timer_end(void *timer, void *stream, float *ms);
int amh_timer_free(void *timer);
/* Per-kernel device timing.  While enabled, the entry points bracket their
 * main kernels (k_refine_pass, refine_graph, k_prepare, k_splat,
 * k_composite, k_p2g, k_forces, k_g2p, k_advect, k_fluid_update,
 * k_fluid_splat, k_blob_stats) with HIP events on the launch stream.
 * amh_profile_read waits for them and returns the number of launches and
 * the summed device milliseconds recorded under `kernel`. */
int amh_profile_enable(int on);
int amh_profile_read(const char *kernel, uint64_t *launches, double *ms);
int amh_profile_reset(void);

/* -------------------------------------------------------------- ingest -- */
/* One key frame: per pixel, `stored` = what morph::add_pixel keeps
 * (rgb_to_hsp when hsp != 0; morph.cpp:298-301, color.cpp:31-44) and
 * `shown` = what morph::get_pixel returns (hsp_to_rgb of it; morph.cpp:
 * 378-392, color.cpp:46-59).  Pixels with alpha 0 are absent (the CLI's
 * fill_morph skips them, demo/main.cpp:96-127) and become 0 in both. */
int amh_ingest_frame(const uint32_t *rgba_dev, int32_t width, int32_t height, int32_t hsp,
                     uint32_t *stored_dev, uint32_t *shown_dev, void *stream);

/* -------------------------------------------------------- segmentation -- */
/* Blob detection at the library default blob_threshold 1.0
 * (thread::blobify_frame, thread.cpp:225-412): every 4-neighbour pair of
 * present pixels is mergeable, so blobs are the 4-connected components of
 * alpha > 0.  labels_dev[i] = smallest row-major index of i's component,
 * -1 for background.  Needs no workspace. */
int amh_segment(const uint32_t *rgba_dev, int32_t width, int32_t height, int32_t *labels_dev,
                void *stream);

/* Per-blob statistics, blobs numbered 0..nblobs-1 in increasing order of
 * their smallest pixel index.  Replaces the running means the reference
 * keeps per blob (thread.cpp:235-243, 346-354): sums are exact integers. */
typedef struct amh_blob_stat {
    int64_t count;                 /* surface size                         */
    int64_t sum_x, sum_y;          /* pixel coordinates                    */
    int64_t sum_c[4];              /* stored colour channels (HSP or RGB)  */
    int32_t x1, y1, x2, y2;        /* bounding box                         */
    int64_t first;                 /* smallest row-major pixel index       */
} amh_blob_stat;
/* Workspace for amh_blob_count / amh_blob_stats, in bytes. */
size_t amh_blob_workspace_bytes(int32_t width, int32_t height);
/* Number of blobs in a labelled frame (synchronous: returns it in *nblobs). */
int amh_blob_count(const int32_t *labels_dev, int32_t width, int32_t height, void *ws_dev,
                   int32_t *nblobs, void *stream);
/* blob_of_pixel_dev[i] = blob number of pixel i or -1; stats_dev[nblobs].
 * Must follow amh_blob_count on the same ws_dev. */
int amh_blob_stats(const int32_t *labels_dev, const uint32_t *stored_dev, int32_t width, int32_t height,
                   void *ws_dev, int32_t nblobs, int32_t *blob_of_pixel_dev, amh_blob_stat *stats_dev,
                   void *stream);
/* Stable grouping of pixel indices by blob: pixels of blob b, in row-major
 * order (the std::set order of blob::surface), land at
 * pixel_list_dev[blob_first_dev[b] ...], blob_first_dev = exclusive scan of
 * the counts (nblobs + 1 entries, written here). */
size_t amh_group_workspace_bytes(int32_t npixels, int32_t nblobs);
int amh_group_pixels(const int32_t *blob_of_pixel_dev, int32_t npixels, int32_t nblobs, void *ws_dev,
                     size_t ws_bytes, int64_t *blob_first_dev, int32_t *pixel_list_dev, void *stream);

/* ---------------------------------------------------------- chain build -- */
/* thread::init_morph (thread.cpp:740-891): row y of chain c holds the pixels
 * of the chain's blob in key frame y in row-major order (flags HAS_PIXEL |
 * HAS_FLUID, zero fractions), then duplicates of a shuffled copy with random
 * fractions and HAS_FLUID cleared; an empty blob contributes one HAS_FLUID
 * point at its rounded centroid first. */
typedef struct amh_chain_build {
    int32_t nkeys;                  /* h                                     */
    int32_t nchains;
    int64_t wtot;                   /* sum of widths                         */
    int32_t frame_width;            /* width of the key-frame images         */
    int32_t pad_;
    const int64_t *chain_off_dev;   /* [nchains] first atom of each chain    */
    const int32_t *chain_w_dev;     /* [nchains] width                       */
    /* [nchains * nkeys]: blob of chain c in key frame y (-1 = empty blob),
     * and for empty blobs the centroid pixel (x | y<<16) to use. */
    const int32_t *chain_blob_dev;
    const uint32_t *empty_xy_dev;
    /* per key frame y: pixel lists grouped by blob (amh_group_pixels) */
    const int32_t *const *pixel_list_dev;   /* host array of nkeys device ptrs */
    const int64_t *const *blob_first_dev;   /* host array of nkeys device ptrs */
    uint64_t seed;
} amh_chain_build;
int amh_build_chains(const amh_chain_build *args, uint64_t *pts_dev, void *stream);

/* ------------------------------------------------ transport cost + refine */
/* Exact cost per chain, sum_x sum_y d(p[x][y], p[x][y+1 mod h]) with
 * point_distance (atomorph.h:334-339; thread::get_energy(chain*),
 * thread.cpp:1109-1125).  cost_dev[2c] / [2c+1] = low / high 64 bits. */
int amh_chain_cost(const uint64_t *pts_dev, int32_t nkeys, int64_t wtot, const int64_t *chain_off_dev,
                   const int32_t *chain_w_dev, int32_t nchains, uint64_t *cost_dev, void *stream);

/* Counters kept on the device by the refine entry points (accumulated). */
typedef struct amh_refine_stats {
    uint64_t attempts;   /* pair swaps evaluated                        */
    uint64_t accepted;   /* swaps made (c1 >= c2, ties accepted)        */
    int64_t gain;        /* exact cost decrease                         */
    uint64_t reserved;
} amh_refine_stats;

/* Plain descent: one visit of thread::morph / morph_asynch
 * (thread.cpp:989-1068) to chain c, row y, as `rounds` conflict-free rounds.
 * Each round is one kernel over a random perfect matching of the chain's
 * slots; every pair whose swap does not raise the cost is swapped.  Pairs of
 * one round are disjoint and read rows y-1 / y+1 only, so the round is
 * exactly a sequence of independent reference attempts.  stream_id selects
 * the random matchings (see pass_stream in the kernel source). */
int amh_refine_chain(uint64_t *pts_dev, int32_t nkeys, int64_t wtot, int64_t chain_off, int32_t chain_w,
                     int32_t y, uint64_t rounds, uint64_t seed, uint64_t stream_id,
                     amh_refine_stats *stats_dev, void *stream);

/* Batched descent: one pass over all chains in a single persistent launch.
 * Work items (chain, slot group) are dequeued from an atomic queue in the
 * order given by items_dev (largest first); each loads its chain's row
 * y = pass_row(seed, c, pass) into LDS, runs `rounds` rounds there, and
 * writes the row back.  Chains wider than the tile capacity are split into
 * groups by a random slot permutation that changes every pass.
 * items_dev[2*i], [2*i+1] = (chain, group); build it with
 * amh_refine_items.  `tile` selects the kernel variant (amh_refine_tiles).
 * queue_dev: 4 uint32 words, zeroed once before first use; the kernel
 * leaves words 0-1 at zero and counts completed passes in word 2.  With
 * pass == UINT64_MAX the pass number is read from word 2 instead. */
int amh_refine_tiles(int32_t *count, int32_t *capacity, int32_t *threads, int32_t max_count);
int amh_refine_items(const int32_t *chain_w, int32_t nchains, int32_t tile, int32_t *items,
                     int32_t max_items, int32_t *nitems);
int amh_refine_batched(uint64_t *pts_dev, int32_t nkeys, int64_t wtot, const int64_t *chain_off_dev,
                       const int32_t *chain_w_dev, const int32_t *items_dev, int32_t nitems,
                       int32_t tile, uint64_t pass, int32_t rounds, uint64_t seed,
                       uint32_t *queue_dev, amh_refine_stats *stats_dev, void *stream);
/* kNN-guided batched descent: replaces the OpenCV build's thread::morph
 * (thread.cpp:893-986), whose second slot is drawn from the two nearest
 * row-(y+1) points of slot x1's point in row y, the two nearest of x1's own
 * row-(y+1) point (FLANN search), or one random slot.  Same items, queue,
 * pass and row schedule as amh_refine_batched; per round every slot of a
 * work item proposes a partner that way (an index of row y+1 in LDS: slots
 * sorted by the Morton key of their pixel, nearest taken among the four
 * entries around a point's place), and the proposals holding both of their
 * slots' claims run.  `tile` < 4 (capacities up to 4096; build the items
 * with the same tile). */
int amh_refine_knn(uint64_t *pts_dev, int32_t nkeys, int64_t wtot, const int64_t *chain_off_dev,
                   const int32_t *chain_w_dev, const int32_t *items_dev, int32_t nitems, int32_t tile,
                   uint64_t pass, int32_t rounds, uint64_t seed, uint32_t *queue_dev,
                   amh_refine_stats *stats_dev, void *stream);
/* hipGraph of `npasses` batched launches whose pass numbers come from the
 * queue's counter, so every amh_graph_launch runs the next npasses passes
 * with one submission.  Capture needs a stream made by amh_stream_create. */
int amh_refine_graph_create(void **graph, uint64_t *pts_dev, int32_t nkeys, int64_t wtot,
                            const int64_t *chain_off_dev, const int32_t *chain_w_dev, const int32_t *items_dev,
                            int32_t nitems, int32_t tile, int32_t npasses, int32_t rounds, uint64_t seed,
                            uint32_t *queue_dev, amh_refine_stats *stats_dev, void *stream);
int amh_graph_launch(void *graph, void *stream);
/* Waits for the device's pending work (replays may still be queued), then
 * frees the graph. */
int amh_graph_destroy(void *graph);

/* Autotuner of the batched descent (no counterpart in the reference: its
 * descent has no tunable schedule).  Races every kernel variant
 * (amh_refine_tiles) x rounds per pass in {4, 8, 16, 32, 64} (only
 * rounds_fixed when it is > 0) from the same start -- a copy of pts_dev in
 * work_dev (same size, distinct) -- for budget_ms of device time each (<= 0:
 * 3 ms), and returns in *tile_out / *rounds_out the candidate with the
 * largest exact cost decrease at the budget.  pts_dev is not modified.  chain_w is the host copy of
 * chain_w_dev.  report (optional, max_report entries) receives every
 * candidate; *nreport their number.  Synchronous; allocates a small
 * scratch buffer. */
typedef struct amh_tune_entry {
    int32_t tile, rounds;     /* candidate                                   */
    int32_t passes;           /* passes run until the budget was reached     */
    int32_t reserved;
    double ms;                /* device time of those passes                 */
    double gain_at_budget;    /* exact cost decrease at budget_ms (linear    */
                              /* within the last pass)                       */
    int64_t gain;             /* exact cost decrease after all the passes    */
} amh_tune_entry;
int amh_refine_autotune(const uint64_t *pts_dev, uint64_t *work_dev, int32_t nkeys, int64_t wtot,
                        const int64_t *chain_off_dev, const int32_t *chain_w_dev, const int32_t *chain_w,
                        int32_t nchains, double budget_ms, uint64_t seed, int32_t rounds_fixed, int32_t *tile_out,
                        int32_t *rounds_out, amh_tune_entry *report, int32_t max_report, int32_t *nreport,
                        void *stream);

/* ------------------------------------------------------ host colour ----- */
/* The kernels' colour arithmetic on the host (color.cpp:31-59, 71-178). */
uint32_t amh_rgb_to_hsp(uint32_t rgba);
uint32_t amh_hsp_to_rgb(uint32_t hspa);
void amh_rgb2hsp(double r, double g, double b, double *h, double *s, double *p);
void amh_hsp2rgb(double h, double s, double p, double *r, double *g, double *b);

/* ----------------------------------------------------- frame rendering -- */
/* morph::get_pixels(t, image) on the atom path (morph.cpp:1405-1421 with
 * draw_atoms 1302-1403, get_pixels(blob) 452-678, get_background 1431-1465,
 * interpolate 1467-1515).  Per atom: trajectory (NONE / LINEAR / closed
 * Catmull-Rom SPLINE), colour fade (NONE/LINEAR, COSINE, PERLIN lag/slope),
 * bilinear splat with the reference's /255 weights and guards; per blob and
 * pixel the weighted mean, alpha * min(1, n/density), optional inward
 * feather; blobs composited with "over" in blob order; optional background
 * cross-dissolve.  Sums are exact integers (the bilinear weights are
 * integers over 255^2), so the result does not depend on atom order; it
 * agrees with the reference's FP64 sums within 1 LSB per channel. */
typedef struct amh_render {
    int32_t width, height;          /* set_resolution                        */
    int32_t nkeys;                  /* key frames                            */
    int32_t nblobs;                 /* blobs per key frame (blob map width)  */
    int32_t bbox_x1, bbox_y1, bbox_x2, bbox_y2;
    int32_t motion;                 /* 2 NONE, 3 LINEAR, 4 SPLINE            */
    int32_t fading;                 /* 2 NONE, 3 LINEAR, 5 COSINE, 6 PERLIN  */
    int32_t show_blobs;             /* 0 TEXTURE, 1 AVERAGE, 2 DISTINCT      */
    int32_t density;
    int32_t feather;
    int32_t keep_background;
    int32_t blend_blobs;
    int32_t nchains;
    int64_t wtot;
    const uint32_t *keys_dev;       /* [nkeys][height*width] get_pixel colour */
    const uint64_t *pts_dev;        /* [nkeys][wtot] chain points             */
    const int64_t *chain_off_dev;   /* [nchains]                              */
    const int32_t *chain_w_dev;     /* [nchains]                              */
    const int32_t *rank_dev;        /* [nkeys][nchains] blob index of chain c
                                       in key frame y (composite order), -1:
                                       not drawn                              */
    const uint32_t *blob_rgba_dev;  /* [nkeys][nblobs] AVERAGE / DISTINCT
                                       colours (NULL for TEXTURE)             */
    const int32_t *lag_perm_dev;    /* PerlinNoise(seed).p[512]               */
    const int32_t *slope_perm_dev;  /* PerlinNoise(seed + 1).p[512]           */
} amh_render;
size_t amh_render_workspace_bytes(int32_t width, int32_t height, int64_t wtot);
/* Prepares a workspace once; every amh_render_frame leaves it clean again. */
int amh_render_workspace_init(void *ws_dev, size_t ws_bytes, int32_t width, int32_t height, int64_t wtot,
                              void *stream);
/* Renders global time t (any real; normalised like morph::normalize_time)
 * into out_dev[height*width] packed RGBA8. */
int amh_render_frame(const amh_render *r, double t, void *ws_dev, size_t ws_bytes, uint32_t *out_dev,
                     void *stream);
/* amh_render_frame without a caller workspace and without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The workspace is a fixed device array of
 * amh_render_noalloc_max_bytes() (96 MiB) per GPU, held until the process
 * ends; frames whose amh_render_workspace_bytes() exceed it are refused with
 * AMH_E_ARG.  Each call clears the array with a fill kernel first.  All
 * calls on one GPU share the array: calls that overlap on one GPU (two
 * streams, two threads) race. */
size_t amh_render_noalloc_max_bytes(void);
int amh_render_frame_noalloc(const amh_render *r, double t, uint32_t *out_dev, void *stream);
/* amh_render_frame replayed from a hipGraph (the reference renders frame by
 * frame, morph.cpp:1302-1429; the graph is this engine's own).  Create
 * captures one frame's launches for r and the workspace (prepared by
 * amh_render_workspace_init) on a stream made by amh_stream_create; each
 * amh_render_graph_frame then writes the frame's time and out_dev into a
 * small device block and replays the graph, with the pixels of
 * amh_render_frame(r, t, ...).  r's sizes, options and device arrays (not
 * their contents) and the workspace must stay as at creation; frames with
 * nothing to draw run eagerly. */
int amh_render_graph_create(void **graph, const amh_render *r, void *ws_dev, size_t ws_bytes, void *stream);
int amh_render_graph_frame(void *graph, double t, uint32_t *out_dev, void *stream);
/* Waits for the device's pending work (replays may still be queued), then
 * frees the graph. */
int amh_render_graph_destroy(void *graph);

/* --------------------------------------------------------- fluid path -- */
/* The fluid renderer (fluidsteps > 0): morph::step_fluid, FluidModel::step
 * and morph::draw_fluid (morph.cpp:807-1300, fluidmodel.cpp:107-580).  The
 * model holds `count` particles (= atoms of the chains; morph.cpp:235-240
 * sizes it by the chains' max_surface, which equals the atoms at density 1)
 * on a gsx x gsy node grid (the reference uses width*s+20 x height*s+20 with
 * s = 1), plus width x height render buffers.  State persists from frame to
 * frame: rendering advances the simulation, as in the reference.  Sizes:
 * 8 <= gsx, gsy < 65535 and 1 <= width, height <= 65535 (AMH_E_ARG beyond). */
int amh_fluid_create(void **fluid, int32_t gsx, int32_t gsy, int32_t count, int32_t width, int32_t height,
                     int32_t nkeys, void *stream);
/* Waits for the device's pending work (frames may still be queued), then
 * frees the model and its graphs. */
int amh_fluid_destroy(void *fluid);
/* Particles back to the state morph::synchronize gives a new model
 * (morph.cpp:243-262): centred, inactive, green. */
int amh_fluid_reset(void *fluid, double sparseness, void *stream);
/* Particle state as 20 doubles per particle, particle i = atom i: x y u v
 * gravity_x gravity_y freedom_r R G B A r g b a strength active mature, 2
 * unused (synchronous).  The model keeps its particles in the order of its
 * last cell sort; these two convert to and from atom order. */
int amh_fluid_set_particles(void *fluid, const double *host, void *stream);
int amh_fluid_get_particles(void *fluid, double *host, void *stream);
/* One FluidModel::step (fluidmodel.cpp:165-580); the wall bounce jitter
 * (rand() in the reference) comes from `seed`, and seed 0 bounces without
 * jitter. */
int amh_fluid_step(void *fluid, int32_t steps_left, double freedom_radius, uint64_t seed, void *stream);
/* Activation order of the fluid particles (no counterpart in the reference,
 * whose std::map bookkeeping births and kills particles, morph.cpp:915-1025):
 * rank_dev[y * wtot + off[c] + x] = place of slot x of chain c among the
 * chain's slots in key segment y -> y+1, ordered by class (HAS_FLUID in
 * rows y and y+1: both, only y, only y+1, neither), then by a pseudo-random
 * permutation of the slots keyed by (seed, c, y).  Every chain's ranks in a
 * segment are a permutation of [0, width). */
int amh_fluid_ranks(const uint64_t *pts_dev, int32_t nkeys, int64_t wtot, const int64_t *chain_off_dev,
                    const int32_t *chain_w_dev, int32_t nchains, uint64_t seed, int32_t *rank_dev, void *stream);
typedef struct amh_fluid_frame {
    const amh_render *r;          /* chains, key frames, options, Perlin     */
    const int32_t *rank_dev;      /* [nkeys][wtot]: activation order of each
                                     atom within its chain, per key segment  */
    const int32_t *surface_dev;   /* [nkeys][nchains]: blob sizes            */
    const double *blob_col_dev;   /* [nkeys][nchains][4] AVERAGE / DISTINCT
                                     colours in [0,1] (NULL for TEXTURE)     */
    double sparseness;            /* grid cells per pixel (reference: 1)     */
    int32_t fluidsteps;           /* FluidModel::step calls per frame        */
    int32_t pad_;
    const uint64_t *atom_keys_dev; /* [wtot][nkeys]: r->pts_dev atom-major
                                     (amh_fluid_atom_keys), one record per
                                     atom for the update; NULL: read r->pts_dev */
} amh_fluid_frame;
/* The key points atom-major: out_dev[a * nkeys + y] = pts_dev[y * wtot + a]
 * (no counterpart in the reference; the fluid update reads an atom's key
 * points in one record instead of one row per key). */
int amh_fluid_atom_keys(const uint64_t *pts_dev, int32_t nkeys, int64_t wtot, uint64_t *out_dev, void *stream);
/* One frame at global time t: particle update, fluidsteps model steps,
 * splat, resolve, outward feather, background; out_dev[width*height].  On a
 * stream made by amh_stream_create the frame's launches (about 5 per model
 * step) are captured once into a hipGraph and replayed for every later
 * frame with the same arguments; the values that change per frame (time,
 * freedom radius, jitter seed, out_dev) reach the kernels through a small
 * parameter block written before each replay.  NULL stream: launched one by
 * one. */
int amh_fluid_render(void *fluid, const amh_fluid_frame *frame, double t, uint32_t *out_dev, void *stream);
/* Graph capture of amh_fluid_render on (1, the default) or off (0).  A handle
 * starts with it off when the environment holds AMH_FLUID_GRAPHS=0. */
int amh_fluid_use_graphs(void *fluid, int32_t on);
/* The drawing half of morph::draw_fluid (morph.cpp:1082-1298) alone, on the
 * model's particles as they are (after amh_fluid_set_particles, say): each
 * active particle's colour relaxes towards its target (kept in the model)
 * and is splatted bilinearly; pixels take the weighted means, the feather
 * grows the drawn region outwards, the background is blended.  Uses only
 * frame->r (size, key frames, bbox, feather, keep_background, fading and
 * Perlin tables) and frame->sparseness. */
int amh_fluid_draw(void *fluid, const amh_fluid_frame *frame, double t, uint32_t *out_dev, void *stream);

/* ---- learned proposal: MLP on the matrix cores -------------------------- */
/* No counterpart in the reference (its swap descent starts from row-major
 * order, thread.cpp:740-891).  8 -> 128 -> 128 -> 128 -> 128 -> 2 with GELU;
 * AMH_MLP_PARAMS FP32 parameters packed as W1[8][128] b1 W2[128][128] b2 W3
 * b3 W4 b4 W5[128][2] b5 (W row-major [in][out]).  Products use bfloat16
 * operands (weights and activations) with FP32 accumulation. */
#define AMH_MLP_PARAMS 50946
/* Handle for batches of up to batch_max rows (parameters zero). */
int amh_mlp_create(void **mlp, int32_t batch_max, void *stream);
int amh_mlp_destroy(void *mlp);
/* Host parameters in and out (synchronous); set resets the Adam state. */
int amh_mlp_set_params(void *mlp, const float *host_params, void *stream);
int amh_mlp_get_params(void *mlp, float *host_params, void *stream);
/* Inference: out_dev[n][2] for x_dev[n][8] (any n, run in batches). */
int amh_mlp_forward(void *mlp, const float *x_dev, int64_t n, float *out_dev, void *stream);
/* Gradient of mean_s sum_o (out - t)^2 over n <= batch_max rows into the
 * handle (and grad_dev[AMH_MLP_PARAMS] when not NULL); adds the loss to
 * *loss_dev when not NULL. */
int amh_mlp_gradients(void *mlp, const float *x_dev, const float *t_dev, int32_t n, float *grad_dev,
                      float *loss_dev, void *stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with the last gradients. */
int amh_mlp_adam_step(void *mlp, float lr, void *stream);
/* c_dev[M][N] = a_dev[M][K] . b_dev[K][N] through the MLP's matrix-core
 * kernel (FP32 row-major in and out, bfloat16 operands). */
int amh_mlp_gemm_test(const float *a_dev, const float *b_dev, float *c_dev, int32_t M, int32_t N, int32_t K,
                      void *stream);

/* ---- learned warm start: features and sort-and-match --------------------- */
/* No counterpart in the reference.  Workspace for chains of wtot atoms in
 * nchains chains over nkeys key frames. */
size_t amh_warm_workspace_bytes(int64_t wtot, int32_t nchains, int32_t nkeys);
/* MLP inputs of the atoms of key frame y (x_dev[wtot][8]: x/W, y/H, RGBA/255
 * of the key-frame pixel, position relative to the chain row's centroid in
 * units of its RMS radius + 1 px) and, if t_dev != NULL, training targets
 * (t_dev[wtot][2]: the same slots' positions in key frame y+1, normalised by
 * that row's centroid and radius).  Uses r's keys_dev, pts_dev, chain
 * offsets and widths. */
int amh_learn_features(const amh_render *r, int32_t y, float *x_dev, float *t_dev, void *ws_dev, size_t ws_bytes,
                       void *stream);
/* Reorders rows 1 .. nkeys-1 of pts_dev (== r->pts_dev) in place: row y+1 of
 * every chain is matched to the MLP's predictions from row y by rank in
 * raster order (LSD radix sort of (chain, y, x) keys at 1/16 px), so each
 * row stays a permutation of its points.  nchains <= 65535. */
int amh_warm_start(void *mlp, const amh_render *r, uint64_t *pts_dev, void *ws_dev, size_t ws_bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif
