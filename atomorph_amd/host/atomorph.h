/*
 * atomorph.h -- public types and constants of the morph engine, the surface
 * of the reference's atomorph.h (atomorph.h:233-358) that am::morph uses.
 *
 * MI355X build: the engine behind am::morph runs on one MI355X through the
 * C ABI in include/atomorph_hip.h.  The deprecated AM_* API
 * (atomorph.h:28-231) and the OpenCV options are not part of this build.
 */
#ifndef _ATOMORPH_H_
#define _ATOMORPH_H_

// The standard headers a caller of the reference receives through
// atomorph.h (atomorph.h:9-17) and its spline.h -> vec3d.h (vec3d.h:15-16):
// programs written against the reference (its demo/main.cpp uses std::cerr
// and std::thread without including them) compile unchanged.
#include <assert.h>
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <future>
#include <iostream>
#include <limits>
#include <map>
#include <random>
#include <set>
#include <vector>

#include "color.h"

namespace am {

const unsigned RGB    = 0; // Don't change the color space.
const unsigned HSP    = 1; // Convert from RGB to HSP.
const unsigned NONE   = 2; // Color/motion interpolation is off.
const unsigned LINEAR = 3; // Uses linear interpolation.
const unsigned SPLINE = 4; // Spline interpolation (motion only).
const unsigned COSINE = 5; // Cosine interpolation (colors only).
const unsigned PERLIN = 6; // Perlin noise dependent interpolation (colors only).
const unsigned STATE_BLOB_DETECTION   = 0;
const unsigned STATE_BLOB_UNIFICATION = 1;
const unsigned STATE_BLOB_MATCHING    = 2;
const unsigned STATE_ATOM_MORPHING    = 3;
const unsigned STATE_DONE             = 4;
const unsigned char HAS_PIXEL         = 1;
const unsigned char HAS_FLUID         = 2;

/* Refinement schedules of this build (extension, see morph::set_refine). */
const unsigned REFINE_PLAIN   = 0; // one chain per step, like thread::morph
const unsigned REFINE_BATCHED = 1; // all chains per step, one persistent launch
const unsigned REFINE_KNN     = 2; // as BATCHED, partners from nearest neighbours

typedef struct pixel {
    uint16_t x;
    uint16_t y;
    color c;
} pixel;

typedef struct blob {
    size_t index;             // Position in the frame's blob vector.
    std::set<size_t> surface; // Pixel positions (y*65536+x) of the blob.
    std::set<size_t> border;  // Unused by this engine, kept for the layout.
    size_t group = 0;         // Chain this blob belongs to.
    bool unified = false;
    double x, y, r, g, b, a;  // Mean position and colour (colour in [0,1]).
} blob;

typedef struct key_frame {
    std::vector<blob *> blobs;
    std::map<size_t, pixel> pixels;
    std::map<size_t, blob *> owners;
    double x, y, r, g, b, a;
    size_t index = 0;
    size_t first_expansion = 0;
    size_t first_dust = 0;
} frame;

typedef union key_point {
    struct {
        uint16_t x;
        uint16_t y;
        uint8_t x_fract;
        uint8_t y_fract;
        uint8_t flags;
    } s;
    uint64_t word;
} point;

typedef struct blob_chain {
    size_t width = 0;
    size_t height = 0;
    point **points = nullptr;
    size_t **places = nullptr;
    double energy = 0.0;
    size_t max_surface = 0;
} chain;

const size_t WARN_POINTER_SIZE = 1;
const size_t WARN_PIXEL_SIZE = 2;
const size_t WARN_POINT_SIZE = 3;

const unsigned TEXTURE  = 0;
const unsigned AVERAGE  = 1;
const unsigned DISTINCT = 2;

void clear_chain(chain *c);
bool renew_chain(chain *c, size_t width, size_t height);

pixel create_pixel(uint16_t x, uint16_t y, unsigned char r, unsigned char g, unsigned char b, unsigned char a);
pixel create_pixel(uint16_t x, uint16_t y, color c);
inline size_t xy2pos(uint16_t x, uint16_t y) { return (y * (UINT16_MAX + 1) + x); }

inline void point2xy(point pt, float *x, float *y) {
    *x = pt.s.x + pt.s.x_fract / float(UINT8_MAX + 1);
    *y = pt.s.y + pt.s.y_fract / float(UINT8_MAX + 1);
}

inline uint32_t pixel_distance(pixel p1, pixel p2) {
    int32_t xd = p1.x - p2.x;
    int32_t yd = p1.y - p2.y;
    return xd * xd + yd * yd;
}

inline uint32_t approx_point_distance(point p1, point p2) {
    int32_t xd = p1.s.x - p2.s.x;
    int32_t yd = p1.s.y - p2.s.y;
    return xd * xd + yd * yd;
}

/* Squared distance in 1/256-pixel fixed point (atomorph.h:334-339). */
inline uint64_t point_distance(point p1, point p2) {
    int64_t xd = int64_t((UINT8_MAX + 1) * p1.s.x + p1.s.x_fract) - ((UINT8_MAX + 1) * p2.s.x + p2.s.x_fract);
    int64_t yd = int64_t((UINT8_MAX + 1) * p1.s.y + p1.s.y_fract) - ((UINT8_MAX + 1) * p2.s.y + p2.s.y_fract);
    return uint64_t(xd * xd) + uint64_t(yd * yd);
}

inline double distance(double x1, double y1, double x2, double y2) {
    double xd = x1 - x2;
    double yd = y1 - y2;
    return xd * xd + yd * yd;
}

inline bool point_has_pixel(point p) { return (p.s.flags & HAS_PIXEL); }
inline bool point_has_fluid(point p) { return (p.s.flags & HAS_FLUID); }

const char *get_version();
size_t get_warning();
bool uses_opencv();

}  // namespace am

#include "morph.h"

#endif
