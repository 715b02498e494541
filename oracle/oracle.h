/*
 * CPU oracle for the atomorph morph-engine hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (atomorph_amd/) links,
 * loads or calls this library; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / CPU timing
 * baseline.  Each function restates one piece of the reference algorithm and
 * cites the reference file:line it follows (paths relative to the atomorph
 * source tree).
 *
 * Point word layout (same bytes as am::point, atomorph.h:275-284, with the
 * padding byte forced to zero):
 *   bits  0-15 x, 16-31 y, 32-39 x_fract, 40-47 y_fract, 48-55 flags.
 *
 * Chain layout used by every oracle entry point: atom-major, like the
 * reference's points[x][y]:  pts[(x) * h + y]  for atom x, key frame y.
 *
 * Pinning (what checks each function, see tests/):
 *   or_point_distance, or_chain_cost  - closed form + reference harness values
 *   or_morph_asynch                   - bit-exact vs the compiled reference
 *   or_refine_rounds                  - invariants, Hungarian lower bound,
 *                                       brute force on w <= 8, statistical
 *                                       parity with or_morph_asynch
 *   or_rgb_to_hsp / or_hsp_to_rgb     - exhaustive-sample vs compiled reference
 *   or_perlin_*                       - vs compiled reference
 *   or_spline_eval                    - vs compiled reference + closed form
 *                                       (passes through control points)
 *   or_segment                        - vs reference blob membership fixtures
 *                                       + brute-force flood fill
 *   or_render_atoms                   - vs reference rendered-frame fixtures
 *   or_fluid_step                     - vs compiled reference FluidModel
 *   or_draw_fluid                     - vs the compiled reference's fluid
 *                                       frames drawn from dumped particles
 *   or_match                          - vs the compiled reference's blob
 *                                       groups, list order, energies and
 *                                       fix_volatiles positions (match.npz)
 *   or_match_brute                    - exhaustive enumeration (the check
 *                                       of the descent's optimum)
 *   or_lap_min_cost                   - vs scipy linear_sum_assignment and
 *                                       brute force permutations (n <= 8)
 *   or_mlp_forward, or_mlp_grad       - central finite differences of the
 *                                       loss; closed-form one-unit cases
 */
#ifndef ATOMORPH_ORACLE_H
#define ATOMORPH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- points and transport cost (atomorph.h:334-339, thread.cpp:1109-1125) */
uint64_t or_point_distance(uint64_t p1, uint64_t p2);
/* Exact cost sum_x sum_y d(p[x][y], p[x][(y+1)%h]); returns the low 64 bits,
 * writes the high 64 bits of the 128-bit sum to *hi (may be NULL). */
uint64_t or_chain_cost(const uint64_t *pts, size_t w, size_t h, uint64_t *hi);

/* ---- reference swap descent, single thread (thread.cpp:989-1041) -------- */
/* Replays morph_asynch with std::mt19937(rng_seed) exactly (libstdc++
 * distributions).  *gain accumulates c1-c2 of accepted swaps, like the
 * reference. Returns the number of accepted swaps. */
uint64_t or_morph_asynch(uint64_t *pts, size_t w, size_t h, size_t repeat,
                         unsigned rng_seed, double *gain);
/* Reference driver thread::morph (thread.cpp:1043-1068) for one chain visit
 * with the engine RNG (std::default_random_engine = minstd_rand0) state
 * *e1_state: draws the morph_asynch seed the way the reference does. */
uint64_t or_morph_visit(uint64_t *pts, size_t w, size_t h, size_t cycle_length,
                        uint64_t *e1_state, double *gain);

/* ---- the GPU's conflict-free round schedule, restated on the CPU -------- */
/* Same hash RNG, Feistel permutation and pairings as the HIP kernels
 * (atomorph_amd/csrc/refine.hip), so GPU output is bit-identical.
 * One round = the fold matching {i, c - i mod w} of row y's slots for a
 * random c; every pair whose swap does not raise the cost (c1 >= c2, ties
 * accepted, the reference's four distances in int64) is swapped.
 * Returns accepted swaps; *gain_out (exact) = cost decrease,
 * *attempts_out = pairs evaluated. */
uint64_t or_refine_rounds(uint64_t *pts, size_t w, size_t h, size_t y,
                          uint64_t rounds, uint64_t seed, uint64_t stream,
                          int64_t *gain_out, uint64_t *attempts_out);
/* One batched pass (amh_refine_batched) replayed on the CPU: every chain c
 * refines row pass_row(seed, c, pass) with `rounds` rounds per slot group of
 * at most `cap` slots.  pts holds the chains atom-major and concatenated
 * (chain c at pts + off[c]*h).  Returns accepted swaps; exact gain. */
uint64_t or_refine_pass(uint64_t *pts, const int64_t *off, const int32_t *widths, int32_t nchains,
                        int32_t h, int32_t cap, uint64_t pass, int32_t rounds, uint64_t seed,
                        int64_t *gain_out, uint64_t *attempts_out);
/* One kNN-guided pass (amh_refine_knn) replayed on the CPU: the schedule of
 * or_refine_pass with partners proposed by the OpenCV build's nearest-
 * neighbour rule (thread.cpp:893-986) and overlapping proposals resolved by
 * the smallest claim.  Returns accepted swaps; exact gain; attempts = the
 * proposals that ran.  Parity unpinned for the choice of neighbours: the
 * reference draws them from OpenCV's FLANN search, which is not available
 * here; the acceptance rule, the exact gains, the row invariants and the
 * two-slot reduction to one reference attempt are pinned by tests. */
uint64_t or_refine_knn_pass(uint64_t *pts, const int64_t *off, const int32_t *widths, int32_t nchains,
                            int32_t h, int32_t cap, uint64_t pass, int32_t rounds, uint64_t seed,
                            int64_t *gain_out, uint64_t *attempts_out);
/* Feistel permutation of [0, n) used by the kernels (exposed for tests). */
uint32_t or_perm_index(uint32_t i, uint32_t n, uint64_t key);
uint64_t or_hash64(uint64_t x);

/* ---- colour (color.cpp:31-59, 71-178) ----------------------------------- */
uint32_t or_rgb_to_hsp(uint32_t rgba);   /* packed r | g<<8 | b<<16 | a<<24 */
uint32_t or_hsp_to_rgb(uint32_t hspa);
/* morph::add_pixel + morph::get_pixel for n pixels: stored = rgb_to_hsp(c)
 * (or c when hsp == 0), shown = hsp_to_rgb(stored) (or stored); alpha 0 ->
 * both 0. */
void or_ingest(const uint32_t *rgba, size_t n, int hsp, uint32_t *stored, uint32_t *shown);

/* ---- Perlin (perlin.cpp:11-89) ------------------------------------------ */
/* p[512] permutation exactly as PerlinNoise(seed) builds it. */
void   or_perlin_perm(unsigned seed, int32_t *p512);
double or_perlin_noise3(const int32_t *p512, double x, double y, double z);
double or_perlin_octave2(const int32_t *p512, double x, double y, int octaves);

/* ---- Catmull-Rom (spline.cpp:29-57) ------------------------------------- */
void or_spline_eval(const double *vx, const double *vy, int n, double t,
                    double *ox, double *oy);

/* ---- segmentation at blob_threshold 1.0 (thread.cpp:225-412) ------------ */
/* 4-connected components of pixels with alpha > 0.  labels[i] = smallest
 * row-major index in the component, -1 for background.  Returns the number
 * of components. */
int or_segment(const uint8_t *rgba, int width, int height, int32_t *labels);

/* ---- rendering, atom path (morph.cpp:452-678, 1302-1421, 1431-1488) ----- */
typedef struct or_render_cfg {
    int32_t width, height;        /* morph::set_resolution               */
    int32_t nkeys;                /* key frames h                        */
    int32_t bbox_x1, bbox_y1, bbox_x2, bbox_y2;
    int32_t motion;               /* am::NONE(2) / LINEAR(3) / SPLINE(4) */
    int32_t fading;               /* NONE/LINEAR/COSINE(5)/PERLIN(6)     */
    int32_t show_blobs;           /* TEXTURE 0 / AVERAGE 1 / DISTINCT 2  */
    int32_t density;
    int32_t feather;
    int32_t keep_background;
    int32_t blend_blobs;
    int32_t pad_;
    const int32_t *lag_perm;      /* p[512] of PerlinNoise(seed)         */
    const int32_t *slope_perm;    /* p[512] of PerlinNoise(seed+1)       */
} or_render_cfg;

/* keys: nkeys images of width*height RGBA8 holding what morph::get_pixel
 *       returns (RGB after the HSP round trip when the delimiter is HSP);
 *       alpha == 0 means "no pixel".
 * chains: nchains chains; chain c has width widths[c] and atom-major points
 *       at pts + offsets[c]*nkeys.
 * order: nkeys rows of nblobs entries: order[y*nblobs + b] = group (chain
 *       index) of blob b of key frame y, -1 = no blob.
 * blob_rgba: nkeys*nblobs packed colours used by AVERAGE / DISTINCT
 *       (blob2pixel or the mt19937 colour), may be NULL for TEXTURE.
 * out: width*height RGBA8. */
void or_render_atoms(const or_render_cfg *cfg, const uint8_t *keys,
                     const uint64_t *pts, const int64_t *offsets,
                     const int64_t *widths, int32_t nchains,
                     const int32_t *order, int32_t nblobs,
                     const uint32_t *blob_rgba, double t, uint8_t *out);

/* color.cpp:22-29: create_color(double r, g, b, a) -> packed RGBA8. */
uint32_t or_create_color(double r, double g, double b, double a);
/* Colour of a blob in show_blobs DISTINCT mode: three bytes of
 * std::mt19937(group) through uniform_int_distribution<unsigned char>
 * (morph.cpp:1323-1328), alpha 255. */
uint32_t or_distinct_color(size_t group);

/* morph::get_time (morph.cpp:1522-1538) */
double or_get_time(size_t frame_index, size_t total_frames, int finite,
                   size_t key_count);

/* ---- fluid (fluidmodel.cpp:165-580), FP64 ------------------------------- */
/* Particle state, 20 doubles each, in this order:
 *  0 x 1 y 2 u 3 v 4 gravity_x 5 gravity_y 6 freedom_r 7 R 8 G 9 B 10 A
 *  11 r 12 g 13 b 14 a 15 strength 16 active 17 mature 18-19 unused
 * Wall jitter uses rand() after srand(jitter_seed) when jitter_seed != 0,
 * otherwise zero jitter. */
void or_fluid_step(double *particles, int32_t count, int32_t gsx, int32_t gsy,
                   int32_t steps_left, double freedom_radius,
                   unsigned jitter_seed);

/* The drawing half of morph::draw_fluid (morph.cpp:1082-1298), after
 * step_fluid: particles (layout above; colours relaxed and mature set in
 * place) at local time t of the key segment on a width x height image,
 * fluid grid sparseness, bbox x1 y1 x2 y2, outward feather layers,
 * keep_background (bg: width*height RGBA8 get_background colours, used only
 * then).  out: width*height RGBA8. */
void or_draw_fluid(double *particles, int32_t count, double t, int32_t width, int32_t height, double sparseness,
                   const int32_t *bbox, int32_t feather, int32_t keep_background, const uint8_t *bg, uint8_t *out);

/* ---- blob matching (thread.cpp:598-738, 1087-1107, 1141-1174, 1187-1233) */
/* thread::match for `steps` worker steps in BLOB_MATCHING (the first pads
 * every frame with blobs at the frame means and shuffles it, the others try
 * one random swap each), driven like thread::step drives it (its counter
 * sets the degeneration period), from std::default_random_engine state
 * rng_state.  Input per key frame y: nblobs[y] blobs in list order with
 * sizes[] and means[][6] = x, y, r, g, b, a (concatenated over frames);
 * fmeans[y][6] the frame's running means (morph::add_pixel); bbox x1 y1 x2
 * y2; weights rgba, size, xy (0..255).  Output, bw = max nblobs:
 * order[y][bw] = caller's index of each listed blob (-1: padding blob),
 * group[y][bw] = its group (the column it held at the last improvement),
 * energy = {current map energy, best energy (infinity before any accepted
 * swap), what get_energy() reports while matching}; with fix != 0,
 * fixed[y][bw][2] = each blob's x, y after init_morph's fix_volatiles;
 * *rng_out = the engine state afterwards.  Returns bw (-1: bad input). */
int32_t or_match(int32_t h, const int32_t *nblobs, const int64_t *sizes, const double *means,
                 const double *fmeans, const int32_t *bbox, const uint8_t *weights, uint64_t degeneration,
                 uint64_t rng_state, uint64_t steps, int fix, int32_t *order, int32_t *group, double *energy,
                 double *fixed, uint64_t *rng_out);
/* get_energy(blob***) of padded lists sizes[h][bw], means[h][bw][6] under
 * group[h][bw] (the column of each list entry). */
double or_match_energy(int32_t h, int32_t bw, const int64_t *sizes, const double *means, const int32_t *bbox,
                       const uint8_t *weights, const int32_t *group);
/* Exhaustive minimum of that energy over every assignment of key frames
 * 1..h-1 to the columns (h <= 3, bw <= 8); best_group receives one
 * minimiser.  Returns the minimum (-1: sizes out of range). */
double or_match_brute(int32_t h, int32_t bw, const int64_t *sizes, const double *means, const int32_t *bbox,
                      const uint8_t *weights, int32_t *best_group);

/* ---- exact assignment (for the OT lower bound) -------------------------- */
/* Minimum-cost perfect matching of an n x n int64 cost matrix (row-major),
 * Hungarian / shortest augmenting path in int64.  assign[i] = column of row
 * i.  Returns the minimum cost. */
int64_t or_lap_min_cost(const int64_t *cost, int32_t n, int32_t *assign);

/* ---- learned-proposal MLP (no counterpart in the reference) ------------- */
/* 8 -> 128 -> 128 -> 128 -> 128 -> 2, exact GELU after each hidden layer.
 * params (OR_MLP_PARAMS floats): W1[8][128] b1[128] W2[128][128] b2[128]
 * W3 b3 W4 b4 (as W2 b2) W5[128][2] b5[2], each W row-major [in][out].
 * x: n x 8 inputs, out: n x 2.  bf16_operands != 0 rounds both operands of
 * every matrix product to bfloat16 first, as the GPU kernels do; all
 * arithmetic is FP64.  (mlp_oracle.cpp) */
#define OR_MLP_PARAMS 50946
void or_mlp_forward(const float *params, const float *x, int32_t n, int bf16_operands, double *out);
/* Loss mean_s sum_o (out - t)^2 (returned) and its gradient with respect to
 * every parameter (OR_MLP_PARAMS doubles, packed like params).  With
 * bf16_operands the backward products round their operands too (dz and the
 * activations for dW, dz and W for the data gradient); bias gradients are
 * unrounded column sums. */
double or_mlp_grad(const float *params, const float *x, const float *t, int32_t n, int bf16_operands,
                   double *grad);

#ifdef __cplusplus
}
#endif
#endif
