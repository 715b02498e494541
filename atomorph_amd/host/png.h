/*
 * png.h -- PNG key frames in and rendered frames out for the command-line
 * driver (the reference's demo uses lodepng: demo/main.cpp:135-213).
 *
 * Decoding accepts every non-interlaced PNG colour type (grey, RGB,
 * palette, grey+alpha, RGBA) at every legal bit depth, with tRNS
 * transparency, and returns RGBA8 (16-bit samples keep their high byte,
 * as lodepng's RGBA8 conversion does).  Encoding writes RGBA8, filter 0,
 * zlib level 6.  Chunk CRCs are checked on input.
 */
#ifndef AM_PNG_H
#define AM_PNG_H

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace am_png {

bool decode_memory(const uint8_t *data, size_t size, std::vector<uint8_t> *rgba, unsigned *width, unsigned *height,
                   std::string *err);
bool decode_file(const std::string &path, std::vector<uint8_t> *rgba, unsigned *width, unsigned *height,
                 std::string *err);
bool encode_file(const std::string &path, const uint8_t *rgba, unsigned width, unsigned height, std::string *err);

}  // namespace am_png

extern "C" {
// C entry points of the codec (for bindings and tests).  decode: RGBA8 into
// out (capacity bytes; call with out = NULL to get *width / *height), 0 on
// success, -1 on error (message from am_png_error).
int am_png_decode_file(const char *path, uint8_t *out, size_t capacity, unsigned *width, unsigned *height);
int am_png_encode_file(const char *path, const uint8_t *rgba, unsigned width, unsigned height);
const char *am_png_error(void);
}

#endif
