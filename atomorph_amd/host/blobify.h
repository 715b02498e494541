/*
 * Greedy blob detection and dust unification on the host, for the option
 * sets where the reference's result is not the 4-connected components that
 * the device segmentation computes: a colour threshold below 1, a blob size
 * cap, a minimum blob size, or a preferred blob count above 1
 * (thread::blobify_frame / blobify / unify_frame / unify,
 * thread.cpp:225-596).
 */
#ifndef AM_BLOBIFY_H
#define AM_BLOBIFY_H

#include <cstddef>
#include <cstdint>
#include <random>
#include <vector>

namespace am_blobify {

struct options {
    double threshold = 1.0;        // blob_threshold: largest colour distance merged
    size_t max_size = SIZE_MAX;    // blob_max_size: a blob this large stops growing
    size_t min_size = 1;           // blob_min_size: smaller blobs are unified as dust
    size_t number = 1;             // blob_number: stop merging at this many blobs
    size_t box_samples = 10;       // blob_box_samples: dust candidates per choice
    uint16_t box_grip = UINT16_MAX;  // blob_box_grip: dust partner box half-size
};

struct blob_info {
    double x, y, r, g, b, a;  // running means as merged (colours in [0, 1])
    uint64_t group;           // y*65536 + x of the pixel the blob grew from
    size_t size;              // pixels
};

struct frame_result {
    std::vector<int32_t> bop;        // [W*H] ordinal of the pixel's blob, -1: no pixel
    std::vector<blob_info> blobs;    // in the reference's blob order
};

// True when the options make detection differ from 4-connected components.
bool needed(const options &o);

// stored[k]: the W x H RGBA words (red in the low byte) of key frame k as the
// morph stores them (after the colour conversion); a pixel is present when
// its alpha is non-zero.  e1: the worker's random engine, seeded with the
// morph's seed (thread::set_seed, thread.cpp:86-90); detection draws from it
// and leaves it where the reference's blob matching continues.  Frames are
// in key order.  Returns false (and leaves out empty) when W*H exceeds the
// 31-bit pixel indices.
bool segment(const std::vector<const uint32_t *> &stored, int W, int H, const options &o,
             std::default_random_engine &e1, std::vector<frame_result> &out);

}  // namespace am_blobify

extern "C" {
// segment() over nframes frames stored back to back; bop [nframes][W*H];
// nblobs [nframes]; blobs [nframes][W*H] rows of (x, y, r, g, b, a), groups
// [nframes][W*H], sizes [nframes][W*H] (a frame has at most W*H blobs).
// Returns 0, or -1 for an invalid argument.
int am_blobify_frames(const uint32_t *stored, int32_t nframes, int32_t W, int32_t H, double threshold,
                      uint64_t max_size, uint64_t min_size, uint64_t number, uint64_t box_samples, uint32_t box_grip,
                      unsigned seed, int32_t *bop, int32_t *nblobs, double *means, uint64_t *groups,
                      uint64_t *sizes);
}

#endif
