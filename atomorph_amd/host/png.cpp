// PNG codec of the command-line driver: chunk parsing, zlib inflate/deflate,
// scanline (un)filtering and conversion to RGBA8, written from the PNG
// specification (ISO/IEC 15948, sections 5-11).
#include "png.h"

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace am_png {
namespace {

const uint8_t kSignature[8] = {137, 80, 78, 71, 13, 10, 26, 10};

uint32_t be32(const uint8_t *p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }
void put32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back(uint8_t(x >> 24));
    v.push_back(uint8_t(x >> 16));
    v.push_back(uint8_t(x >> 8));
    v.push_back(uint8_t(x));
}

bool fail(std::string *err, const std::string &msg) {
    if (err) *err = msg;
    return false;
}

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Reverses the per-scanline filters in place (section 9.2).
bool unfilter(uint8_t *data, size_t rows, size_t row_bytes, size_t bpp, std::string *err) {
    std::vector<uint8_t> zero(row_bytes, 0);
    const uint8_t *prev = zero.data();
    for (size_t r = 0; r < rows; ++r) {
        uint8_t *line = data + r * (row_bytes + 1);
        const uint8_t type = line[0];
        uint8_t *cur = line + 1;
        switch (type) {
            case 0:
                break;
            case 1:
                for (size_t i = bpp; i < row_bytes; ++i) cur[i] = uint8_t(cur[i] + cur[i - bpp]);
                break;
            case 2:
                for (size_t i = 0; i < row_bytes; ++i) cur[i] = uint8_t(cur[i] + prev[i]);
                break;
            case 3:
                for (size_t i = 0; i < row_bytes; ++i) {
                    const int left = i >= bpp ? cur[i - bpp] : 0;
                    cur[i] = uint8_t(cur[i] + ((left + prev[i]) >> 1));
                }
                break;
            case 4:
                for (size_t i = 0; i < row_bytes; ++i) {
                    const int left = i >= bpp ? cur[i - bpp] : 0, upleft = i >= bpp ? prev[i - bpp] : 0;
                    cur[i] = uint8_t(cur[i] + paeth(left, prev[i], upleft));
                }
                break;
            default:
                return fail(err, "unknown scanline filter " + std::to_string(type));
        }
        prev = cur;
    }
    return true;
}

// Raw sample k of a scanline at `depth` bits (1, 2, 4, 8 or 16).
inline unsigned sample(const uint8_t *row, size_t k, unsigned depth) {
    if (depth == 8) return row[k];
    if (depth == 16) return (unsigned(row[2 * k]) << 8) | row[2 * k + 1];
    const size_t bit = k * depth;
    const unsigned shift = 8 - depth - unsigned(bit % 8);
    return (row[bit / 8] >> shift) & ((1u << depth) - 1);
}

// Sample scaled to 8 bits: 16-bit keeps its high byte, low depths are
// stretched to 0..255.
inline uint8_t to8(unsigned v, unsigned depth) {
    if (depth == 8) return uint8_t(v);
    if (depth == 16) return uint8_t(v >> 8);
    return uint8_t(v * 255u / ((1u << depth) - 1));
}

}  // namespace

bool decode_memory(const uint8_t *data, size_t size, std::vector<uint8_t> *rgba, unsigned *width, unsigned *height,
                   std::string *err) {
    if (size < 8 || std::memcmp(data, kSignature, 8) != 0) return fail(err, "not a PNG file");
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    unsigned depth = 0, ctype = 0, interlace = 0;
    bool have_ihdr = false, have_iend = false;
    std::vector<uint8_t> idat, plte, trns;
    while (pos + 12 <= size) {
        const uint32_t len = be32(data + pos);
        if (len > size - pos - 12) return fail(err, "truncated chunk");
        const uint8_t *type = data + pos + 4, *body = data + pos + 8;
        const uint32_t crc = be32(body + len);
        if (uint32_t(crc32(crc32(0L, Z_NULL, 0), type, len + 4)) != crc) return fail(err, "chunk CRC mismatch");
        const std::string t(reinterpret_cast<const char *>(type), 4);
        if (t == "IHDR") {
            if (len != 13) return fail(err, "bad IHDR");
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
            if (body[10] != 0 || body[11] != 0) return fail(err, "unknown compression or filter method");
            have_ihdr = true;
        } else if (t == "PLTE") {
            plte.assign(body, body + len);
        } else if (t == "tRNS") {
            trns.assign(body, body + len);
        } else if (t == "IDAT") {
            idat.insert(idat.end(), body, body + len);
        } else if (t == "IEND") {
            have_iend = true;
            break;
        } else if (!(type[0] & 0x20)) {
            return fail(err, "unknown critical chunk " + t);
        }
        pos += 12 + size_t(len);
    }
    if (!have_ihdr || !have_iend) return fail(err, "missing IHDR or IEND");
    if (w == 0 || h == 0 || w > (1u << 16) || h > (1u << 16)) return fail(err, "unsupported image size");
    if (interlace != 0) return fail(err, "interlaced PNG is not supported");
    unsigned channels = 0;
    bool depth_ok = false;
    switch (ctype) {
        case 0: channels = 1; depth_ok = depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16; break;
        case 2: channels = 3; depth_ok = depth == 8 || depth == 16; break;
        case 3: channels = 1; depth_ok = depth == 1 || depth == 2 || depth == 4 || depth == 8; break;
        case 4: channels = 2; depth_ok = depth == 8 || depth == 16; break;
        case 6: channels = 4; depth_ok = depth == 8 || depth == 16; break;
        default: break;
    }
    if (!channels || !depth_ok) return fail(err, "invalid colour type / bit depth");
    if (ctype == 3 && (plte.empty() || plte.size() % 3 != 0)) return fail(err, "palette image without PLTE");
    const size_t row_bytes = (size_t(w) * channels * depth + 7) / 8;
    const size_t bpp = std::max<size_t>(1, channels * depth / 8);
    std::vector<uint8_t> raw(size_t(h) * (row_bytes + 1));
    uLongf out_len = uLongf(raw.size());
    if (uncompress(raw.data(), &out_len, idat.data(), uLong(idat.size())) != Z_OK || out_len != raw.size())
        return fail(err, "corrupt image data");
    if (!unfilter(raw.data(), h, row_bytes, bpp, err)) return false;

    rgba->assign(size_t(w) * h * 4, 0);
    const size_t npal = plte.size() / 3;
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t *row = raw.data() + size_t(y) * (row_bytes + 1) + 1;
        uint8_t *o = rgba->data() + size_t(y) * w * 4;
        for (uint32_t x = 0; x < w; ++x, o += 4) {
            switch (ctype) {
                case 0: {
                    const unsigned v = sample(row, x, depth);
                    o[0] = o[1] = o[2] = to8(v, depth);
                    o[3] = (trns.size() >= 2 && v == ((unsigned(trns[0]) << 8) | trns[1])) ? 0 : 255;
                    break;
                }
                case 2: {
                    const unsigned r = sample(row, 3 * x, depth), g = sample(row, 3 * x + 1, depth),
                                   b = sample(row, 3 * x + 2, depth);
                    o[0] = to8(r, depth);
                    o[1] = to8(g, depth);
                    o[2] = to8(b, depth);
                    const bool key = trns.size() >= 6 && r == ((unsigned(trns[0]) << 8) | trns[1]) &&
                                     g == ((unsigned(trns[2]) << 8) | trns[3]) &&
                                     b == ((unsigned(trns[4]) << 8) | trns[5]);
                    o[3] = key ? 0 : 255;
                    break;
                }
                case 3: {
                    const unsigned i = sample(row, x, depth);
                    if (i >= npal) return fail(err, "palette index out of range");
                    o[0] = plte[3 * i];
                    o[1] = plte[3 * i + 1];
                    o[2] = plte[3 * i + 2];
                    o[3] = i < trns.size() ? trns[i] : 255;
                    break;
                }
                case 4:
                    o[0] = o[1] = o[2] = to8(sample(row, 2 * x, depth), depth);
                    o[3] = to8(sample(row, 2 * x + 1, depth), depth);
                    break;
                default:
                    for (unsigned c = 0; c < 4; ++c) o[c] = to8(sample(row, 4 * x + c, depth), depth);
                    break;
            }
        }
    }
    *width = w;
    *height = h;
    return true;
}

bool decode_file(const std::string &path, std::vector<uint8_t> *rgba, unsigned *width, unsigned *height,
                 std::string *err) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return fail(err, "cannot open " + path);
    std::vector<uint8_t> buf;
    uint8_t tmp[65536];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    return decode_memory(buf.data(), buf.size(), rgba, width, height, err);
}

bool encode_file(const std::string &path, const uint8_t *rgba, unsigned width, unsigned height, std::string *err) {
    if (width == 0 || height == 0) return fail(err, "empty image");
    const size_t row = size_t(width) * 4;
    std::vector<uint8_t> raw(size_t(height) * (row + 1));
    for (unsigned y = 0; y < height; ++y) {
        raw[y * (row + 1)] = 0;  // filter type None
        std::memcpy(&raw[y * (row + 1) + 1], rgba + y * row, row);
    }
    uLongf zlen = compressBound(uLong(raw.size()));
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), uLong(raw.size()), 6) != Z_OK) return fail(err, "deflate failed");
    z.resize(zlen);
    std::vector<uint8_t> out(kSignature, kSignature + 8);
    auto chunk = [&](const char *type, const uint8_t *body, size_t len) {
        put32(out, uint32_t(len));
        const size_t start = out.size();
        out.insert(out.end(), type, type + 4);
        out.insert(out.end(), body, body + len);
        put32(out, uint32_t(crc32(crc32(0L, Z_NULL, 0), out.data() + start, uInt(len + 4))));
    };
    std::vector<uint8_t> ihdr;
    put32(ihdr, width);
    put32(ihdr, height);
    const uint8_t rest[5] = {8, 6, 0, 0, 0};  // 8-bit RGBA, deflate, adaptive filtering, no interlace
    ihdr.insert(ihdr.end(), rest, rest + 5);
    chunk("IHDR", ihdr.data(), ihdr.size());
    chunk("IDAT", z.data(), z.size());
    chunk("IEND", nullptr, 0);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(err, "cannot create " + path);
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0 && ok) ? true : fail(err, "write failed: " + path);
}

}  // namespace am_png

namespace {
thread_local std::string g_png_error;
}

extern "C" int am_png_decode_file(const char *path, uint8_t *out, size_t capacity, unsigned *width,
                                  unsigned *height) {
    std::vector<uint8_t> img;
    unsigned w = 0, h = 0;
    if (!path || !am_png::decode_file(path, &img, &w, &h, &g_png_error)) return -1;
    if (width) *width = w;
    if (height) *height = h;
    if (out) {
        if (capacity < img.size()) {
            g_png_error = "output buffer too small";
            return -1;
        }
        std::memcpy(out, img.data(), img.size());
    }
    return 0;
}

extern "C" int am_png_encode_file(const char *path, const uint8_t *rgba, unsigned width, unsigned height) {
    if (!path || !rgba) {
        g_png_error = "invalid argument";
        return -1;
    }
    return am_png::encode_file(path, rgba, width, height, &g_png_error) ? 0 : -1;
}

extern "C" const char *am_png_error(void) { return g_png_error.c_str(); }
