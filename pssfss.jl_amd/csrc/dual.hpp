// Forward-mode dual numbers (value, derivative along one direction) for the
// adjoint kernels.  Branches and integer decisions are taken on values, as the
// reference's piecewise formulas are; floor / rint have zero derivative.
#pragma once
#include "cplx.hpp"

struct dd {
    double v, d;
};

HD dd mkd(double v, double d = 0.0) { return dd{v, d}; }
HD dd operator+(dd a, dd b) { return dd{a.v + b.v, a.d + b.d}; }
HD dd operator-(dd a, dd b) { return dd{a.v - b.v, a.d - b.d}; }
HD dd operator-(dd a) { return dd{-a.v, -a.d}; }
HD dd operator*(dd a, dd b) { return dd{a.v * b.v, a.d * b.v + a.v * b.d}; }
HD dd operator/(dd a, dd b) {
    double q = a.v / b.v;
    return dd{q, (a.d - q * b.d) / b.v};
}
HD dd operator+(dd a, double b) { return dd{a.v + b, a.d}; }
HD dd operator+(double a, dd b) { return dd{a + b.v, b.d}; }
HD dd operator-(dd a, double b) { return dd{a.v - b, a.d}; }
HD dd operator-(double a, dd b) { return dd{a - b.v, -b.d}; }
HD dd operator*(dd a, double b) { return dd{a.v * b, a.d * b}; }
HD dd operator*(double a, dd b) { return dd{a * b.v, a * b.d}; }
HD dd operator/(dd a, double b) { return dd{a.v / b, a.d / b}; }
HD dd operator/(double a, dd b) {
    double q = a / b.v;
    return dd{q, -q * b.d / b.v};
}
HD dd dsqrt(dd a) {
    double s = sqrt(a.v);
    return dd{s, s > 0 ? 0.5 * a.d / s : 0.0};
}
HD dd dexp(dd a) {
    double e = exp(a.v);
    return dd{e, e * a.d};
}
HD dd dlog(dd a) { return dd{log(a.v), a.d / a.v}; }
HD dd dsin(dd a) { return dd{sin(a.v), cos(a.v) * a.d}; }
HD dd dcos(dd a) { return dd{cos(a.v), -sin(a.v) * a.d}; }
HD dd dfabs(dd a) { return a.v < 0 ? -a : a; }

// complex dual: real and imaginary parts are duals
struct cdd {
    dd x, y;
};
HD cdd mkc(cd a) { return cdd{dd{a.x, 0.0}, dd{a.y, 0.0}}; }
HD cdd mkc(dd re) { return cdd{re, dd{0.0, 0.0}}; }
HD cd val(cdd a) { return cd{a.x.v, a.y.v}; }
HD cd der(cdd a) { return cd{a.x.d, a.y.d}; }
HD cdd operator+(cdd a, cdd b) { return cdd{a.x + b.x, a.y + b.y}; }
HD cdd operator-(cdd a, cdd b) { return cdd{a.x - b.x, a.y - b.y}; }
HD cdd operator-(cdd a) { return cdd{-a.x, -a.y}; }
HD cdd operator*(cdd a, cdd b) { return cdd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
HD cdd operator*(cdd a, cd b) { return cdd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
HD cdd operator*(cd a, cdd b) { return b * a; }
HD cdd operator*(cdd a, dd b) { return cdd{a.x * b, a.y * b}; }
HD cdd operator*(dd a, cdd b) { return cdd{a * b.x, a * b.y}; }
HD cdd operator*(cdd a, double b) { return cdd{a.x * b, a.y * b}; }
HD cdd operator*(double a, cdd b) { return cdd{a * b.x, a * b.y}; }
HD cdd operator*(cd a, dd b) { return cdd{a.x * b, a.y * b}; }
HD cdd operator*(dd a, cd b) { return cdd{a * b.x, a * b.y}; }
HD cdd operator+(cdd a, cd b) { return cdd{a.x + b.x, a.y + b.y}; }
HD cdd operator+(cd a, cdd b) { return b + a; }
HD cdd operator-(cdd a, cd b) { return cdd{a.x - b.x, a.y - b.y}; }
HD cdd operator+(cdd a, dd b) { return cdd{a.x + b, a.y}; }
HD cdd operator+(dd a, cdd b) { return cdd{a + b.x, b.y}; }
HD cdd operator/(cdd a, dd b) { return cdd{a.x / b, a.y / b}; }
HD cdd operator/(cdd a, cdd b) {
    dd den = b.x * b.x + b.y * b.y;
    return cdd{(a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den};
}
HD cdd operator/(cd a, cdd b) { return mkc(a) / b; }
HD cdd operator/(cdd a, cd b) { return a * (1.0 / b); }
HD cdd operator+(double a, cdd b) { return cdd{a + b.x, b.y}; }
HD cdd &operator+=(cdd &a, cdd b) { a = a + b; return a; }
// exp(i t) of a real dual
HD cdd dcis(dd t) { return cdd{dcos(t), dsin(t)}; }
