/*
 * am::color -- RGBA8 colour and the HSP conversions, with the interface of
 * the reference's color.h (color.h:1-33).  The conversions run the HIP
 * library's host build of the kernels' arithmetic (amh_rgb_to_hsp & co.).
 */
#ifndef AM_COLOR_H
#define AM_COLOR_H

#include <cmath>
#include <cstdint>

namespace am {

typedef struct color {
    uint8_t r;
    uint8_t g;
    uint8_t b;
    uint8_t a;
} color;

color create_color(unsigned char r, unsigned char g, unsigned char b, unsigned char a);
color create_color(double r, double g, double b, double a);

/* Euclidean RGBA distance scaled to [0, 1] (color.h:17-24). */
inline double color_distance(color c1, color c2) {
    int16_t rd = c1.r - c2.r;
    int16_t gd = c1.g - c2.g;
    int16_t bd = c1.b - c2.b;
    int16_t ad = c1.a - c2.a;
    return std::sqrt(double(rd * rd + gd * gd + bd * bd + ad * ad)) / 510.0;
}

color rgb_to_hsp(color c);
color hsp_to_rgb(color c);

void RGBtoHSP(double R, double G, double B, double *H, double *S, double *P);
void HSPtoRGB(double H, double S, double P, double *R, double *G, double *B);

inline uint32_t color_word(color c) {
    return uint32_t(c.r) | (uint32_t(c.g) << 8) | (uint32_t(c.b) << 16) | (uint32_t(c.a) << 24);
}
inline color word_color(uint32_t w) {
    color c;
    c.r = uint8_t(w & 255u);
    c.g = uint8_t((w >> 8) & 255u);
    c.b = uint8_t((w >> 16) & 255u);
    c.a = uint8_t(w >> 24);
    return c;
}

}  // namespace am

#endif
