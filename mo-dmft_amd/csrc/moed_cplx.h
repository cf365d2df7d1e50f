// Complex arithmetic and the modulus as the reference build evaluates them (device side).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace moed {

// Complex arithmetic as the reference build evaluates it: textbook division, a real operand
// promoted to (x, 0), no contraction (every product and sum rounded on its own).
struct cxd {
    double re, im;
};
__device__ __forceinline__ cxd cx_div(cxd a, cxd b) {
#pragma clang fp contract(off)
    const double den = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}
__device__ __forceinline__ cxd cx_rdiv(double r, cxd b) { return cx_div({r, 0.0}, b); }
__device__ __forceinline__ cxd cx_mul(cxd a, cxd b) {
#pragma clang fp contract(off)
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// |z| as sqrt(ax^2 + ay^2) followed by one correction step without FMA: C. F. Borges,
// "An Improved Algorithm for hypot(a,b)", ACM TOMS 47 (2021), the fma-free corrected form.
// Range guards keep ax^2 + ay^2 finite and normal; for the inputs of the bath fit the plain
// path is taken.
__device__ __forceinline__ double hypot_corrected(double x, double y) {
#pragma clang fp contract(off)
    x = fabs(x);
    y = fabs(y);
    const double ax = x < y ? y : x;
    const double ay = x < y ? x : y;
    const double EPS = 0x1p-54, SCALE = 0x1p-600;
    auto kernel = [](double a, double b) {
#pragma clang fp contract(off)
        double h = sqrt(a * a + b * b);
        double t1, t2;
        if (h <= 2.0 * b) {
            const double delta = h - b;
            t1 = a * (2.0 * delta - a);
            t2 = (delta - 2.0 * (a - b)) * delta;
        } else {
            const double delta = h - a;
            t1 = 2.0 * delta * (a - 2.0 * b);
            t2 = (4.0 * delta - b) * b + delta * delta;
        }
        h -= (t1 + t2) / (2.0 * h);
        return h;
    };
    if (ax > 0x1p+511) {
        if (ay <= ax * EPS) return ax + ay;
        return kernel(ax * SCALE, ay * SCALE) / SCALE;
    }
    if (ay < 0x1p-511) {
        if (ax >= ay / EPS) return ax + ay;
        return kernel(ax / SCALE, ay / SCALE) * SCALE;
    }
    if (ay <= ax * EPS) return ax + ay;
    return kernel(ax, ay);
}

__device__ __forceinline__ double round_f32(double v) { return (double)(float)v; }

}  // namespace moed
