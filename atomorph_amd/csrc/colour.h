// Colour arithmetic of the morph engine, FP64 with the reference's
// operation order (compiled with -ffp-contract=off) so 8-bit results match.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amh {

__host__ __device__ __forceinline__ uint32_t narrow8(double v) { return uint32_t(int32_t(v)) & 0xFFu; }
__host__ __device__ __forceinline__ uint32_t narrow16(double v) { return uint32_t(int32_t(v)) & 0xFFFFu; }
__host__ __device__ __forceinline__ uint32_t rgba_pack(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    return (r & 255u) | ((g & 255u) << 8) | ((b & 255u) << 16) | ((a & 255u) << 24);
}

// Finley's HSP model (color.cpp:71-178, public-domain formulas): perceived
// brightness P = sqrt(.299 R^2 + .587 G^2 + .114 B^2), hue measured within
// the sextant of the largest channel.
__host__ __device__ inline void rgb2hsp(double R, double G, double B, double *H, double *S, double *P) {
    const double wr = 0.299, wg = 0.587, wb = 0.114;
    *P = sqrt(R * R * wr + G * G * wg + B * B * wb);
    if (R == G && R == B) {
        *H = 0.0;
        *S = 0.0;
        return;
    }
    // top: largest channel; after / before: the channels that follow / precede
    // it in R->G->B->R order; hue origins of the two half-sextants.
    double top, after, before, origin_down, origin_up;
    if (R >= G && R >= B) {
        top = R; after = G; before = B; origin_down = 6. / 6.; origin_up = 0. / 6.;
    } else if (G >= R && G >= B) {
        top = G; after = B; before = R; origin_down = 2. / 6.; origin_up = 2. / 6.;
    } else {
        top = B; after = R; before = G; origin_down = 4. / 6.; origin_up = 4. / 6.;
    }
    if (before >= after) {
        *H = origin_down - 1. / 6. * (before - after) / (top - after);
        *S = 1. - after / top;
    } else {
        *H = origin_up + 1. / 6. * (after - before) / (top - before);
        *S = 1. - before / top;
    }
}

__host__ __device__ inline void hsp2rgb(double H, double S, double P, double *R, double *G, double *B) {
    const double w[3] = {0.299, 0.587, 0.114};
    // per sextant: channel index of the largest, middle and smallest value
    int sext;
    if (H < 1. / 6.) sext = 0;
    else if (H < 2. / 6.) sext = 1;
    else if (H < 3. / 6.) sext = 2;
    else if (H < 4. / 6.) sext = 3;
    else if (H < 5. / 6.) sext = 4;
    else sext = 5;
    const int big = (sext == 0 || sext == 5) ? 0 : (sext == 1 || sext == 2) ? 1 : 2;
    const int mid = (sext == 0 || sext == 3) ? 1 : (sext == 1 || sext == 4) ? 0 : 2;
    const int low = 3 - big - mid;
    const double f = (sext & 1) ? 6. * (-H + double(sext + 1) / 6.) : 6. * (H - double(sext) / 6.);
    double c[3];
    const double ratio = 1. - S;  // smallest / largest
    if (ratio > 0.) {
        const double part = 1. + f * (1. / ratio - 1.);
        const double lo = P / sqrt(w[big] / ratio / ratio + w[mid] * part * part + w[low]);
        const double hi = lo / ratio;
        c[low] = lo;
        c[big] = hi;
        c[mid] = lo + f * (hi - lo);
    } else {
        const double hi = sqrt(P * P / (w[big] + w[mid] * f * f));
        c[big] = hi;
        c[mid] = hi * f;
        c[low] = 0.;
    }
    *R = c[0];
    *G = c[1];
    *B = c[2];
}

// color.cpp:31-44
__host__ __device__ inline uint32_t rgb_to_hsp(uint32_t c) {
    double h, s, p;
    rgb2hsp((c & 255u) / 255.0, ((c >> 8) & 255u) / 255.0, ((c >> 16) & 255u) / 255.0, &h, &s, &p);
    return rgba_pack(narrow8(round(h * 255.0)), narrow8(round(s * 255.0)), narrow8(round(p * 255.0)), c >> 24);
}
// color.cpp:46-59 (only the upper bound is clamped)
__host__ __device__ inline uint32_t hsp_to_rgb(uint32_t c) {
    double r, g, b;
    hsp2rgb((c & 255u) / 255.0, ((c >> 8) & 255u) / 255.0, ((c >> 16) & 255u) / 255.0, &r, &g, &b);
    return rgba_pack(narrow8(fmin(round(r * 255.0), 255.0)), narrow8(fmin(round(g * 255.0), 255.0)),
                     narrow8(fmin(round(b * 255.0), 255.0)), c >> 24);
}

}  // namespace amh
