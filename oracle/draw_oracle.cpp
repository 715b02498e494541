/*
 * The drawing half of morph::draw_fluid restated for the CPU oracle
 * (TEST INFRASTRUCTURE ONLY, see oracle.h): morph.cpp:1082-1298, which runs
 * after step_fluid has moved the particles.  Per active particle the colour
 * is relaxed towards its target (1082-1097), mapped to the pixel grid
 * (1099-1117) and splatted bilinearly into per-pixel lists (1119-1165);
 * each touched pixel is the weighted mean of its list, summed in particle
 * order (1167-1208); an optional outward feather grows the drawn region by
 * layers of fading alpha (1210-1262); the result is blended over the
 * background (1264-1298).
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <set>
#include <vector>

#include "oracle.h"

namespace {

struct Col {
    uint8_t r, g, b, a;
};

inline uint8_t u8(double v) { return static_cast<uint8_t>(v); }
inline Col col_of(double r, double g, double b, double a) {  // create_color(double...) (color.cpp:22-29)
    return {u8(std::round(r * 255.0)), u8(std::round(g * 255.0)), u8(std::round(b * 255.0)),
            u8(std::round(a * 255.0))};
}
inline size_t xy2pos(uint16_t x, uint16_t y) { return size_t(y) * 65536 + x; }

}  // namespace

extern "C" void or_draw_fluid(double *P, int32_t count, double t, int32_t width, int32_t height,
                              double sparseness, const int32_t *bbox, int32_t feather, int32_t keep_background,
                              const uint8_t *bg, uint8_t *out) {
    const int bx1 = bbox[0], by1 = bbox[1], bx2 = bbox[2], by2 = bbox[3];
    std::map<size_t, std::vector<Col>> colors;
    std::map<size_t, std::vector<double>> weights;
    for (int32_t i = 0; i < count; ++i) {
        double *p = P + 20 * size_t(i);
        if (p[16] == 0.0) continue;  // inactive
        const double s = p[15];
        const double w = (1.0 / (100.0 * std::pow((t - (1.0 * (1.0 - s))), 2.0) + 1.0)) * std::pow(s, 2.0);
        for (int k = 0; k < 4; ++k) p[11 + k] = (w * p[7 + k]) + (1.0 - w) * p[11 + k];
        double px, py;
        if (p[0] <= 10.0) px = 0.0;
        else if (p[0] >= width * sparseness + 10.0) px = (width - 1);
        else px = (p[0] - 10.0) / sparseness;
        if (p[1] <= 10.0) py = 0.0;
        else if (p[1] >= height * sparseness + 10.0) py = (height - 1);
        else py = (p[1] - 10.0) / sparseness;
        const uint16_t x = uint16_t(std::min(int(px), width - 1));
        const uint16_t y = uint16_t(std::min(int(py), height - 1));
        p[17] = 1.0;  // mature
        const double xf = std::modf(px, &px), yf = std::modf(py, &py);
        if (px >= width || py >= height)
            if (px > bx2 || px < bx1 || py > by2 || py < by1) continue;
        const double w11 = ((1.0 - xf) * (1.0 - yf)), w21 = (xf * (1.0 - yf));
        const double w12 = ((1.0 - xf) * yf), w22 = (xf * yf);
        const Col c = col_of(p[11], p[12], p[13], p[14]);
        auto put = [&](size_t pos, double wt) {
            colors[pos].push_back(c);
            weights[pos].push_back(wt);
        };
        if (w11 > 0.0) put(xy2pos(x, y), w11);
        if ((x < bx2 || x + 1 < width) && w21 > 0.0) put(xy2pos(uint16_t(x + 1), y), w21);
        if ((y < by2 || y + 1 < height) && w12 > 0.0) put(xy2pos(x, uint16_t(y + 1)), w12);
        if (w22 > 0.0 && ((y < by2 && x < bx2) || (y + 1 < height && x + 1 < width)))
            put(xy2pos(uint16_t(x + 1), uint16_t(y + 1)), w22);
    }
    // image: the background (or transparent black), then the drawn pixels
    const size_t npx = size_t(width) * height;
    for (size_t i = 0; i < npx; ++i)
        for (int k = 0; k < 4; ++k) out[4 * i + k] = keep_background ? bg[4 * i + k] : 0;
    std::map<size_t, Col> image_buffer, blob;
    for (const auto &kv : colors) {
        const std::vector<Col> &cs = kv.second;
        const std::vector<double> &ws = weights[kv.first];
        double r = 0.0, g = 0.0, b = 0.0, a = 0.0, wsum = 0.0;
        for (size_t k = 0; k < cs.size(); ++k) {
            const double wt = ws[k];
            wsum += wt;
            r += cs[k].r * wt;
            g += cs[k].g * wt;
            b += cs[k].b * wt;
            a += cs[k].a * wt;
        }
        r = std::round(r / wsum);
        g = std::round(g / wsum);
        b = std::round(b / wsum);
        a = std::round(a / wsum);
        if (keep_background) a *= 1.0 - std::pow(t, 24.0);
        const uint16_t x = uint16_t(kv.first % 65536), y = uint16_t(kv.first / 65536);
        const Col c = {u8(r), u8(g), u8(b), u8(a)};
        if (feather > 0) blob[kv.first] = c;
        else image_buffer[size_t(y) * width + x] = c;
    }
    if (feather > 0) {
        for (int layer = 0; layer < feather; ++layer) {
            const double alpha = 1.0 - (double(layer + 1) / double(feather + 1));
            std::set<size_t> border;
            for (const auto &kv : blob) {
                const uint16_t x = uint16_t(kv.first % 65536), y = uint16_t(kv.first / 65536);
                size_t q;
                if (x + 1 < width && !blob.count(q = xy2pos(uint16_t(x + 1), y))) border.insert(q);
                if (y + 1 < height && !blob.count(q = xy2pos(x, uint16_t(y + 1)))) border.insert(q);
                if (x > 0 && !blob.count(q = xy2pos(uint16_t(x - 1), y))) border.insert(q);
                if (y > 0 && !blob.count(q = xy2pos(x, uint16_t(y - 1)))) border.insert(q);
            }
            if (border.empty()) break;
            std::map<size_t, Col> grown;
            for (const size_t pos : border) {
                const uint16_t x = uint16_t(pos % 65536), y = uint16_t(pos / 65536);
                std::set<size_t> comps;
                size_t q;
                if (x + 1 < width && blob.count(q = xy2pos(uint16_t(x + 1), y))) comps.insert(q);
                if (y + 1 < height && blob.count(q = xy2pos(x, uint16_t(y + 1)))) comps.insert(q);
                if (x > 0 && blob.count(q = xy2pos(uint16_t(x - 1), y))) comps.insert(q);
                if (y > 0 && blob.count(q = xy2pos(x, uint16_t(y - 1)))) comps.insert(q);
                double r = 0.0, g = 0.0, b = 0.0, a = 0.0;
                const double w = double(comps.size());
                for (const size_t cpos : comps) {
                    const Col c = blob[cpos];
                    r += (c.r / 255.0) / w;
                    g += (c.g / 255.0) / w;
                    b += (c.b / 255.0) / w;
                    a += (c.a / 255.0) / w;
                }
                a *= alpha;
                grown[pos] = col_of(r, g, b, a);
            }
            for (const auto &kv : grown) blob[kv.first] = kv.second;
        }
        for (const auto &kv : blob) {
            const uint16_t x = uint16_t(kv.first % 65536), y = uint16_t(kv.first / 65536);
            image_buffer[size_t(y) * width + x] = kv.second;
        }
    }
    for (const auto &kv : image_buffer) {
        if (kv.first >= npx) continue;  // the reference's image->at() would refuse it
        Col c = kv.second;
        if (keep_background) {
            const uint8_t *q = bg + 4 * kv.first;
            const double bgr = double(q[0]) / 255.0, bgg = double(q[1]) / 255.0;
            const double bgb = double(q[2]) / 255.0, bga = double(q[3]) / 255.0;
            double r = c.r / 255.0, g = c.g / 255.0, b = c.b / 255.0, a = c.a / 255.0;
            r = a * r + (1.0 - a) * bgr;
            g = a * g + (1.0 - a) * bgg;
            b = a * b + (1.0 - a) * bgb;
            a = bga + (1.0 - bga) * a;
            c = {u8(std::round(r * 255.0)), u8(std::round(g * 255.0)), u8(std::round(b * 255.0)),
                 u8(std::round(a * 255.0))};
        }
        uint8_t *o = out + 4 * kv.first;
        o[0] = c.r; o[1] = c.g; o[2] = c.b; o[3] = c.a;
    }
}
