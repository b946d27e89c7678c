// Complex FP64/FP32 arithmetic for gfx950 device code (and host code of the library).
// Plain structs: every kernel keeps complex values as (re, im) register pairs and
// interleaved (re, im) in memory, the layout of Julia's ComplexF64 arrays.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define HD __host__ __device__ __forceinline__

struct cd {
    double x, y;
};
struct cf {
    float x, y;
};

HD cd mkc(double x, double y) { return cd{x, y}; }
HD cd operator+(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
HD cd operator-(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
HD cd operator-(cd a) { return cd{-a.x, -a.y}; }
HD cd operator*(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
HD cd operator*(double s, cd a) { return cd{s * a.x, s * a.y}; }
HD cd operator*(cd a, double s) { return cd{a.x * s, a.y * s}; }
HD cd operator/(cd a, double s) { return cd{a.x / s, a.y / s}; }
HD cd operator+(cd a, double s) { return cd{a.x + s, a.y}; }
HD cd operator+(double s, cd a) { return cd{s + a.x, a.y}; }
HD cd operator-(double s, cd a) { return cd{s - a.x, -a.y}; }
HD cd operator-(cd a, double s) { return cd{a.x - s, a.y}; }
HD cd& operator+=(cd& a, cd b) { a.x += b.x; a.y += b.y; return a; }
HD cd& operator-=(cd& a, cd b) { a.x -= b.x; a.y -= b.y; return a; }
HD cd conj(cd a) { return cd{a.x, -a.y}; }
HD double cabs1(cd a) { return fabs(a.x) + fabs(a.y); }

// Smith's scaled complex division.
HD cd operator/(cd a, cd b) {
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x;
        double d = b.x + r * b.y;
        return cd{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
    } else {
        double r = b.x / b.y;
        double d = b.y + r * b.x;
        return cd{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
    }
}
HD cd operator/(double a, cd b) { return cd{a, 0.0} / b; }

HD double cabs(cd a) { return hypot(a.x, a.y); }
HD cd cisd(double t) { return cd{cos(t), sin(t)}; }

// Principal square root (Kahan): result has a non-negative real part.
HD cd csqrt_(cd z) {
    if (z.x == 0.0 && z.y == 0.0) return cd{0.0, z.y};
    double t = sqrt((fabs(z.x) + hypot(z.x, z.y)) * 0.5);
    if (z.x >= 0.0) return cd{t, z.y / (2.0 * t)};
    return cd{fabs(z.y) / (2.0 * t), copysign(t, z.y)};
}

// mysqrt of src/PGF.jl:116-120: flip a pure negative-imaginary root to +j.
HD cd mysqrt(cd z) {
    cd a = csqrt_(z);
    if (a.x == 0.0 && a.y < 0.0) a = cd{-a.x, -a.y};
    return a;
}

// Complex tanh (Kahan's formulation), overflow-safe for large |Re z|.
HD cd ctanh_(cd z) {
    double x = z.x, y = z.y;
    if (4.0 * fabs(x) > 710.4758600739439) {  // asinh(DBL_MAX)
        double e = exp(-4.0 * fabs(x));
        return cd{copysign(1.0, x), copysign(4.0 * e * sin(y) * cos(y), y)};
    }
    double t = tan(y);
    double beta = 1.0 + t * t;
    double s = sinh(x);
    double rho = sqrt(1.0 + s * s);
    double den = 1.0 + beta * s * s;
    return cd{(beta * rho * s) / den, t / den};
}

HD cd cexp_(cd z) {
    double e = exp(z.x);
    return cd{e * cos(z.y), e * sin(z.y)};
}

// FP32 helpers for the mixed-precision smooth parts.
HD cf operator+(cf a, cf b) { return cf{a.x + b.x, a.y + b.y}; }
HD cf operator*(cf a, cf b) { return cf{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
HD cf operator*(float s, cf a) { return cf{s * a.x, s * a.y}; }
HD cf mkf(float x, float y) { return cf{x, y}; }
HD cf operator-(cf a, cf b) { return cf{a.x - b.x, a.y - b.y}; }
HD cf operator-(cf a) { return cf{-a.x, -a.y}; }
HD cf &operator+=(cf &a, cf b) { a.x += b.x; a.y += b.y; return a; }
HD cf &operator-=(cf &a, cf b) { a.x -= b.x; a.y -= b.y; return a; }
HD cf conj(cf a) { return cf{a.x, -a.y}; }
HD float cabs1(cf a) { return fabsf(a.x) + fabsf(a.y); }
HD cf operator/(cf a, cf b) {
    if (fabsf(b.x) >= fabsf(b.y)) {
        float r = b.y / b.x;
        float d = b.x + r * b.y;
        return cf{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
    } else {
        float r = b.x / b.y;
        float d = b.y + r * b.x;
        return cf{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
    }
}

// Scalar-type traits used by the templated dense kernels.
template <class C> struct CTraits;
template <> struct CTraits<cd> {
    typedef double R;
    static HD cd zero() { return cd{0.0, 0.0}; }
    static HD cd one() { return cd{1.0, 0.0}; }
};
template <> struct CTraits<cf> {
    typedef float R;
    static HD cf zero() { return cf{0.0f, 0.0f}; }
    static HD cf one() { return cf{1.0f, 0.0f}; }
};
// c += a * b
HD void cfma(cd &c, cd a, cd b) {
    c.x = fma(a.x, b.x, c.x); c.x = fma(-a.y, b.y, c.x);
    c.y = fma(a.x, b.y, c.y); c.y = fma(a.y, b.x, c.y);
}
HD void cfma(cf &c, cf a, cf b) {
    c.x = fmaf(a.x, b.x, c.x); c.x = fmaf(-a.y, b.y, c.x);
    c.y = fmaf(a.x, b.y, c.y); c.y = fmaf(a.y, b.x, c.y);
}
