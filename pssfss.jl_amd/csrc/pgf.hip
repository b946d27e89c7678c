// Periodic Green's function kernels for gfx950:
//   * spectral tables: Kummer-subtracted modal summands (ring-major storage),
//     per-ring convergence sums, square gather, DFT (as GEMMs, linalg.hip),
//     Floquet phase + ghost rows          -- src/PGF.jl:196-365, 478-629
//   * spatial lattice sums integrated over the source triangle for several
//     phasings at once (filljk! / jksums) -- src/Zint.jl:54-125, src/PGF.jl:67-106
//   * analytic singular integrals          -- src/Zint.jl:126-165
#include <type_traits>

#include "internal.hpp"
#include "mfma.hpp"
#include "geom.hpp"

namespace {

constexpr double TWOPI = 6.283185307179586;
constexpr double PI_ = 3.141592653589793;

// The FP32 lattice term pair of squared distance d2 (in units of 1/u): J = e^-a / a
// and K = e^-a with a = sqrt(d2).  One reciprocal square root gives both 1/a and
// a = d2 / sqrt(d2), so a term takes two quarter-rate instructions (rsq, exp)
// instead of three (sqrt, exp, rcp).
__device__ __forceinline__ void term_pair_f32(float d2, float &j, float &k) {
    const float ia = __builtin_amdgcn_rsqf(d2);
    const float e = __expf(-(d2 * ia));
    j = e * ia;
    k = e;
}

// Summands of eqs. (5.19a,b) (electric) or (5.26a,b) (magnetic) for lattice index
// (m, n): transmission-line impedances stepped through the layers from both ends.
__device__ void summand(const TabParams &c, int m, int n, cd &t1, cd &t2) {
#pragma clang fp contract(off)
    double bx = (c.b00[0] + (double)m * c.b1[0]) + (double)n * c.b2[0];
    double by = (c.b00[1] + (double)m * c.b1[1]) + (double)n * c.b2[1];
    double b2 = bx * bx + by * by;
    if (!(b2 >= c.b2min)) b2 = c.b2min;
    double kap2 = b2 + c.u * c.u;
    double kap = sqrt(kap2);
    const int nl = c.nl, s = c.s;
    cd g = mysqrt(b2 - (c.k0sq * c.eps[0]) * c.mu[0]);
    if (!c.magnetic) {
        cd zlte = c.mu[0] / g, zltm = g / c.eps[0];
        for (int i = 1; i < s; ++i) {
            g = mysqrt(b2 - (c.k0sq * c.eps[i]) * c.mu[i]);
            cd z0te = c.mu[i] / g, z0tm = g / c.eps[i];
            cd th = ctanh_(c.width[i] * g);
            zlte = (z0te * (zlte + z0te * th)) / (z0te + zlte * th);
            zltm = (z0tm * (zltm + z0tm * th)) / (z0tm + zltm * th);
        }
        g = mysqrt(b2 - (c.k0sq * c.eps[nl - 1]) * c.mu[nl - 1]);
        cd zrte = c.mu[nl - 1] / g, zrtm = g / c.eps[nl - 1];
        for (int i = nl - 2; i >= s; --i) {
            g = mysqrt(b2 - (c.k0sq * c.eps[i]) * c.mu[i]);
            cd z0te = c.mu[i] / g, z0tm = g / c.eps[i];
            cd th = ctanh_(c.width[i] * g);
            zrte = (z0te * (zrte + z0te * th)) / (z0te + zrte * th);
            zrtm = (z0tm * (zrtm + z0tm * th)) / (z0tm + zrtm * th);
        }
        cd vte = (zlte * zrte) / (zlte + zrte);
        cd vtm = (zltm * zrtm) / (zltm + zrtm);
        vte = (2.0 * vte) / c.mut;
        t1 = vte - (1.0 + c.c3 / kap2) / kap;
        vtm = (2.0 * c.epsb) * vtm;
        t2 = (vtm + ((c.k0sq * c.epsb) * c.mut) * vte) / b2 - (1.0 + c.d3 / kap2) / kap;
    } else {
        cd ylte = g / c.mu[0], yltm = c.eps[0] / g;
        for (int i = 1; i < s; ++i) {
            g = mysqrt(b2 - (c.k0sq * c.eps[i]) * c.mu[i]);
            cd y0te = g / c.mu[i], y0tm = c.eps[i] / g;
            cd th = ctanh_(c.width[i] * g);
            ylte = (y0te * (ylte + y0te * th)) / (y0te + ylte * th);
            yltm = (y0tm * (yltm + y0tm * th)) / (y0tm + yltm * th);
        }
        g = mysqrt(b2 - (c.k0sq * c.eps[nl - 1]) * c.mu[nl - 1]);
        cd yrte = g / c.mu[nl - 1], yrtm = c.eps[nl - 1] / g;
        for (int i = nl - 2; i >= s; --i) {
            g = mysqrt(b2 - (c.k0sq * c.eps[i]) * c.mu[i]);
            cd y0te = g / c.mu[i], y0tm = c.eps[i] / g;
            cd th = ctanh_(c.width[i] * g);
            yrte = (y0te * (yrte + y0te * th)) / (y0te + yrte * th);
            yrtm = (y0tm * (yrtm + y0tm * th)) / (y0tm + yrtm * th);
        }
        t1 = (yltm + yrtm) - (c.f1 + c.f3 / kap2) / kap;
        t2 = ((((yltm + yrtm) * c.k0sq) + ylte) + yrte) / b2 - (c.p1 + c.p3 / kap2) / kap;
    }
}

// k0-derivative helpers on complex duals: the value is the library's own routine,
// the derivative its analytic one (sqrt: dz / 2a; tanh: (1 - t^2) dz).
__device__ __forceinline__ cdd dmysqrt(cdd z) {
    const cd a = mysqrt(val(z));
    const cd d = der(z) / (2.0 * a);
    return cdd{dd{a.x, d.x}, dd{a.y, d.y}};
}
__device__ __forceinline__ cdd dctanh(cdd z) {
    const cd t = ctanh_(val(z));
    const cd d = (1.0 - t * t) * der(z);
    return cdd{dd{t.x, d.x}, dd{t.y, d.y}};
}

// d/dk0 of the summands of eqs. (5.19a,b) (electric) or (5.26a,b) (magnetic) at
// fixed lattice wavevector (fixed phasing) and fixed u, forward mode through the
// same transmission-line recursion as summand().  k0 enters through k0^2 in every
// layer's gamma, the extraction constants (electric: c3, d3; magnetic: f3, p3,
// whose derivatives TabParams holds in dc3, dd3) and the small-beta floor b2min
// (proportional to k0^2).
__device__ void summand_dk0(const TabParams &c, int m, int n, cd &d1, cd &d2) {
    const double bx = (c.b00[0] + (double)m * c.b1[0]) + (double)n * c.b2[0];
    const double by = (c.b00[1] + (double)m * c.b1[1]) + (double)n * c.b2[1];
    const double k0 = sqrt(c.k0sq);
    dd b2 = mkd(bx * bx + by * by);
    if (!(b2.v >= c.b2min)) b2 = dd{c.b2min, 2.0 * c.b2min / k0};
    const dd k0sq = dd{c.k0sq, 2.0 * k0};
    const dd kap2 = b2 + c.u * c.u;
    const dd kap = dsqrt(kap2);
    const int nl = c.nl, s = c.s;
    auto gamma = [&](int i) { return dmysqrt(mkc(b2) - (k0sq * c.eps[i]) * c.mu[i]); };
    // value and derivative of the two extraction constants
    const cdd x3 = c.magnetic ? cdd{dd{c.f3.x, c.dc3.x}, dd{c.f3.y, c.dc3.y}}
                              : cdd{dd{c.c3.x, c.dc3.x}, dd{c.c3.y, c.dc3.y}};
    const cdd y3 = c.magnetic ? cdd{dd{c.p3.x, c.dd3.x}, dd{c.p3.y, c.dd3.y}}
                              : cdd{dd{c.d3.x, c.dd3.x}, dd{c.d3.y, c.dd3.y}};
    cdd t1, t2;
    if (!c.magnetic) {
        cdd g = gamma(0);
        cdd zlte = c.mu[0] / g, zltm = g / c.eps[0];
        for (int i = 1; i < s; ++i) {
            g = gamma(i);
            const cdd z0te = c.mu[i] / g, z0tm = g / c.eps[i];
            const cdd th = dctanh(c.width[i] * g);
            zlte = (z0te * (zlte + z0te * th)) / (z0te + zlte * th);
            zltm = (z0tm * (zltm + z0tm * th)) / (z0tm + zltm * th);
        }
        g = gamma(nl - 1);
        cdd zrte = c.mu[nl - 1] / g, zrtm = g / c.eps[nl - 1];
        for (int i = nl - 2; i >= s; --i) {
            g = gamma(i);
            const cdd z0te = c.mu[i] / g, z0tm = g / c.eps[i];
            const cdd th = dctanh(c.width[i] * g);
            zrte = (z0te * (zrte + z0te * th)) / (z0te + zrte * th);
            zrtm = (z0tm * (zrtm + z0tm * th)) / (z0tm + zrtm * th);
        }
        cdd vte = (zlte * zrte) / (zlte + zrte);
        cdd vtm = (zltm * zrtm) / (zltm + zrtm);
        vte = (2.0 * vte) / c.mut;
        t1 = vte - (1.0 + x3 / kap2) / kap;
        vtm = (2.0 * c.epsb) * vtm;
        t2 = (vtm + ((k0sq * c.epsb) * c.mut) * vte) / b2 - (1.0 + y3 / kap2) / kap;
    } else {
        cdd g = gamma(0);
        cdd ylte = g / c.mu[0], yltm = c.eps[0] / g;
        for (int i = 1; i < s; ++i) {
            g = gamma(i);
            const cdd y0te = g / c.mu[i], y0tm = c.eps[i] / g;
            const cdd th = dctanh(c.width[i] * g);
            ylte = (y0te * (ylte + y0te * th)) / (y0te + ylte * th);
            yltm = (y0tm * (yltm + y0tm * th)) / (y0tm + yltm * th);
        }
        g = gamma(nl - 1);
        cdd yrte = g / c.mu[nl - 1], yrtm = c.eps[nl - 1] / g;
        for (int i = nl - 2; i >= s; --i) {
            g = gamma(i);
            const cdd y0te = g / c.mu[i], y0tm = c.eps[i] / g;
            const cdd th = dctanh(c.width[i] * g);
            yrte = (y0te * (yrte + y0te * th)) / (y0te + yrte * th);
            yrtm = (y0tm * (yrtm + y0tm * th)) / (y0tm + yrtm * th);
        }
        t1 = (yltm + yrtm) - (c.f1 + x3 / kap2) / kap;
        t2 = ((((yltm + yrtm) * k0sq) + ylte) + yrte) / b2 - (c.p1 + y3 / kap2) / kap;
    }
    d1 = der(t1);
    d2 = der(t2);
}

__global__ void k_summands(const TabParams *tp, int64_t cap, int64_t e0, int64_t e1, cd *t1,
                           cd *t2) {
    int64_t e = e0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    const int p = blockIdx.y;
    int r = ring_of(e);
    int m, n;
    ring_mn(r, (int)(e - ring_base(r)), m, n);
    cd a, b;
    summand(tp[p], m, n, a, b);
    t1[p * cap + e] = a;
    t2[p * cap + e] = b;
}

__global__ void k_summands_dk0(const TabParams *tp, int64_t cap, int64_t e0, int64_t e1, cd *t1,
                               cd *t2) {
    int64_t e = e0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    const int p = blockIdx.y;
    int r = ring_of(e);
    int m, n;
    ring_mn(r, (int)(e - ring_base(r)), m, n);
    cd a, b;
    summand_dk0(tp[p], m, n, a, b);
    t1[p * cap + e] = a;
    t2[p * cap + e] = b;
}

// Ring sums in ring order (src/PGF.jl:237-288) and the convergence flag per ring.
__global__ void k_ringsums(int64_t cap, int r0, int r1, int rcap, const cd *t1, const cd *t2,
                           double convtest, int *flags) {
    int r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r > r1) return;
    const int p = blockIdx.y;
    const cd *a = t1 + p * cap, *b = t2 + p * cap;
    int64_t base = ring_base(r);
    int len = r == 0 ? 1 : 8 * r;
    cd s1 = {0.0, 0.0}, s2 = {0.0, 0.0};
    for (int k = 0; k < len; ++k) {
        s1 += a[base + k];
        s2 += b[base + k];
    }
    int f = 0;
    if (r > 0) {
        double q1 = cabs(s1 / a[0]);
        double q2 = cabs(s2 / b[0]);
        f = (q1 < convtest && q2 < convtest) ? 1 : 0;
    }
    flags[(int64_t)p * rcap + r] = f;
}

// S_p = [table1 | table2] as an M x 2M column-major array, entry (i, j) from
// lattice index (i - M/2, j - M/2) (src/PGF.jl:311-312).
__global__ void k_gather_square(const int *Ms, int64_t cap, const cd *t1, const cd *t2, cd *S,
                                int64_t s_stride) {
    const int p = blockIdx.z;
    const int M = Ms[p];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= M || j >= M) return;
    int64_t idx = ring_index(i - M / 2, j - M / 2);
    cd *Sp = S + p * s_stride;
    Sp[i + (int64_t)M * j] = t1[p * cap + idx];
    Sp[i + (int64_t)M * (j + M)] = t2[p * cap + idx];
}

// Unnormalised forward DFT matrix W[j,k] = exp(-2 pi i jk / M), exact integer phase.
__global__ void k_dft_matrix(int M, cd *W) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int k = blockIdx.y * blockDim.y + threadIdx.y;
    if (j >= M || k >= M) return;
    int64_t e = ((int64_t)j * k) % M;
    double ang = -TWOPI * (double)e / (double)M;
    W[j + (int64_t)M * k] = cd{cos(ang), sin(ang)};
}

// Phase factor of eq. (5.32), 1/(2A) or 1/A scaling, and the ghost rows/columns
// at -1 and M (src/PGF.jl:316-357, 593-621).  F holds the two transformed
// tables side by side (M x 2M); tab receives two (M+2)^2 tables.
struct PhaseCtx {
    const cd *F;
    int M, magnetic;
    double psi1, psi2, scale;
    cd c1, c2, c12, g1, g2;
};

__device__ __forceinline__ cd tab_interior(const PhaseCtx &c, int which, int i, int j) {
#pragma clang fp contract(off)
    double pterm = i * (PI_ - c.psi1 / c.M);
    double qterm = j * (PI_ - c.psi2 / c.M);
    cd cf = cisd(pterm + qterm) / c.scale;
    cd v = c.F[i + (int64_t)c.M * (j + which * c.M)];
    return v * cf;
}

__device__ cd tab_value(const PhaseCtx &c, int which, int i, int j) {
    const int M = c.M;
    // Upper ghosts first (they only reference interior values).
    auto up = [&](int ii, int jj) -> cd {
        if (ii < M && jj < M) return tab_interior(c, which, ii, jj);
        if (ii == M && jj == M) return c.c12 * tab_interior(c, which, 0, 0);
        if (ii == M) return c.c1 * tab_interior(c, which, 0, jj);
        return c.c2 * tab_interior(c, which, ii, 0);
    };
    if (i >= 0 && j >= 0) return up(i, j);
    if (i == -1 && j >= 0) return c.g1 * up(M - 1, j);            // row -1, cols 0..M
    // column -1, rows -1..M-1
    cd src = (i == -1) ? c.g1 * up(M - 1, M - 1) : up(i, M - 1);
    return c.g2 * src;
}

__global__ void k_phase_ghost(const int *Ms, const double *psis, const int *mags,
                              const double *areas, const cd *F, int64_t f_stride, cd *tab,
                              int64_t tab_stride) {
    const int p = blockIdx.z;
    PhaseCtx c;
    c.M = Ms[p];
    c.magnetic = mags[p];
    c.psi1 = psis[2 * p];
    c.psi2 = psis[2 * p + 1];
    c.scale = c.magnetic ? areas[p] : 2.0 * areas[p];
    c.F = F + p * f_stride;
    const int E = c.M + 2;
    int i = blockIdx.x * blockDim.x + threadIdx.x - 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y - 1;
    if (i > c.M || j > c.M) return;
    c.c1 = cisd(-c.psi1);
    c.c2 = cisd(-c.psi2);
    if (!c.magnetic) {
        c.c12 = c.c1 * c.c2;
        c.g1 = 1.0 / c.c1;
        c.g2 = 1.0 / c.c2;
    } else {
        c.c12 = cisd(-(c.psi1 + c.psi2));
        c.g1 = cisd(c.psi1);
        c.g2 = cisd(c.psi2);
    }
    cd *out = tab + p * tab_stride;
    if (i == c.M && j == -1) {   // never referenced by the interpolation
        out[(i + 1) + (int64_t)E * (j + 1)] = cd{0.0, 0.0};
        out[(int64_t)E * E + (i + 1) + (int64_t)E * (j + 1)] = cd{0.0, 0.0};
        return;
    }
    out[(i + 1) + (int64_t)E * (j + 1)] = tab_value(c, 0, i, j);
    out[(int64_t)E * E + (i + 1) + (int64_t)E * (j + 1)] = tab_value(c, 1, i, j);
}

// ---------------------------------------------------------------- spatial sums
// One lane per (face-pair class, quadrature point); 9 classes x 7 points per
// wavefront.  Every lane sums the lattice rings of jksums (src/PGF.jl:67-106) for
// F phasings at once: the frequency-independent distance, exp(-d) and 1/d are
// shared by the F phasings, whose lattice phases for |m|,|n| <= JK_RT come from
// LDS.  Each phasing keeps its own ring count (break test at 1e-8 after ring 5).
// MIXED: ring terms (r >= 1) and ring partial sums in FP32, the extracted (0,0)
// term and the running sums in FP64.
template <class T>
struct LatPtT {
    T ox, oy;      // the lattice offset rounded to T
    T lx, ly;      // what that rounding left out (FP32 only; zero in FP64)
    int m, n;
};
typedef LatPtT<float> LatPt;

// lat (MIXED): the lattice offsets of rings 1..JK_RINGMAX in ring order as FP32 hi + lo
// pairs (k_lattice), so that a ring term's distance takes a few FP32 operations on
// values that are uniform over the wavefront instead of FP64 products per lane.
template <int F, bool MIXED>
__global__ __launch_bounds__(256) void k_jk(SheetDev sh, int64_t nufp, double ul, double us1x,
                                            double us1y, double us2x, double us2y,
                                            const double *psis, int nf, cd *jk,
                                            int64_t jk_stride, const LatPt *lat) {
#pragma clang fp contract(off)
    constexpr int RT = JK_RT, W = 2 * JK_RT + 1;
    typedef typename std::conditional<MIXED, cf, cd>::type PT;
    __shared__ PT ph[F][W * W];
    __shared__ double spsi[2 * F];
    for (int t = threadIdx.x; t < 2 * F; t += blockDim.x) spsi[t] = t < 2 * nf ? psis[t] : 0.0;
    for (int t = threadIdx.x; t < F * W * W; t += blockDim.x) {
        int f = t / (W * W), e = t % (W * W);
        int m = e / W - RT, n = e % W - RT;
        double c = 1.0, sn = 0.0;
        if (f < nf) {
            double a = (double)m * psis[2 * f] + (double)n * psis[2 * f + 1];
            c = cos(-a);
            sn = sin(-a);
        }
        PT v;
        v.x = c;
        v.y = sn;
        ph[f][e] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int qp = lane % 7;
    const int64_t iu = ((int64_t)blockIdx.x * 4 + wave) * 9 + lane / 7;
    const bool active = lane < 63 && iu < nufp;
    double urx = 0.0, ury = 0.0;
    if (active) {
        int64_t ifm, ifs;
        class_faces(sh, iu, ifm, ifs);
        const double *rs = sh.fxs + 6 * ifs;
        const double *rm = sh.fxs + 6 * ifm;
        double rs12x = rs[0] - rs[2], rs12y = rs[1] - rs[3];
        double rs31x = rs[4] - rs[0], rs31y = rs[5] - rs[1];
        double rs21x = -rs12x, rs21y = -rs12y;
        double rmcx = ((rm[0] + rm[2]) + rm[4]) / 3.0;
        double rmcy = ((rm[1] + rm[3]) + rm[5]) / 3.0;
        double rtx = (rs[0] + rs21x * QXI[qp]) + rs31x * QETA[qp];
        double rty = (rs[1] + rs21y * QXI[qp]) + rs31y * QETA[qp];
        urx = ul * (rmcx - rtx);
        ury = ul * (rmcy - rty);
    }
    // the observation offset as an FP32 hi + lo pair (MIXED ring terms)
    const float uxh = (float)urx, uyh = (float)ury;
    const float uxl = (float)(urx - (double)uxh), uyl = (float)(ury - (double)uyh);
    cd js[F], ks[F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
        js[f] = cd{0.0, 0.0};
        ks[f] = cd{0.0, 0.0};
    }
    if (active) {
        double arg = sqrt(urx * urx + ury * ury);
        double rterm = exp(-arg);
        double j0;
        if (arg <= 1e-5) j0 = (-arg / 6 + 0.5) * arg - 1;
        else j0 = (rterm - 1) / arg;
#pragma unroll
        for (int f = 0; f < F; ++f) {
            js[f] = cd{j0, 0.0};
            ks[f] = cd{rterm, 0.0};
        }
        unsigned done = 0;
#pragma unroll
        for (int f = 0; f < F; ++f)
            if (f >= nf) done |= 1u << f;
        const unsigned all = (F == 32) ? 0xffffffffu : ((1u << F) - 1u);
        // one phasing at psi = (0, 0) (normal incidence): every lattice phase is
        // exactly 1, so the terms are real and the phase products can go; the
        // sums are bit-identical to the general loop's (e * 1 = e, e * -0 adds 0)
        const bool zp = F == 1 && nf == 1 && spsi[0] == 0.0 && spsi[1] == 0.0;
        for (int r = 1; zp && r <= JK_RINGMAX && done != all; ++r) {
            typename CTraits<PT>::R jr = 0, kr = 0;
            const int len = 8 * r;
            const int64_t rb = ring_base(r);
            for (int k = 0; k < len; ++k) {
                if (MIXED) {
                    const LatPt L = lat[rb + k];
                    const float fx = (uxh - L.ox) + (uxl - L.lx);
                    const float fy = (uyh - L.oy) + (uyl - L.ly);
                    float ea, e;
                    term_pair_f32(fx * fx + fy * fy, ea, e);
                    jr = fmaf(ea, 1.0f, jr);
                    kr = fmaf(e, 1.0f, kr);
                } else {
                    int m, n;
                    ring_mn(r, k, m, n);
                    double dx = urx - ((double)m * us1x + (double)n * us2x);
                    double dy = ury - ((double)m * us1y + (double)n * us2y);
                    double a = sqrt(dx * dx + dy * dy);
                    double e = exp(-a);
                    double ia = 1.0 / a;
                    jr += e * ia;
                    kr += e;
                }
            }
            const cd a = cd{(double)jr, 0.0}, b = cd{(double)kr, 0.0};
            js[0] += a;
            ks[0] += b;
            if (r > 5) {
                double c1 = cabs(a / js[0]), c2 = cabs(b / ks[0]);
                double conv = c1 > c2 ? c1 : c2;
                if (conv < 1e-8) done = all;
            }
        }
        for (int r = 1; !zp && r <= JK_RINGMAX && done != all; ++r) {
            PT jr[F], kr[F];
#pragma unroll
            for (int f = 0; f < F; ++f) {
                jr[f].x = 0; jr[f].y = 0;
                kr[f].x = 0; kr[f].y = 0;
            }
            const int len = 8 * r;
            const int64_t rb = ring_base(r);
            for (int k = 0; k < len; ++k) {
                int m, n;
                double dx = 0, dy = 0;
                float fx = 0, fy = 0;
                if (MIXED) {
                    const LatPt L = lat[rb + k];
                    m = L.m;
                    n = L.n;
                    fx = (uxh - L.ox) + (uxl - L.lx);
                    fy = (uyh - L.oy) + (uyl - L.ly);
                } else {
                    ring_mn(r, k, m, n);
                    dx = urx - ((double)m * us1x + (double)n * us2x);
                    dy = ury - ((double)m * us1y + (double)n * us2y);
                }
                if (MIXED) {
                    float ea, e;
                    term_pair_f32(fx * fx + fy * fy, ea, e);
                    if (r <= RT) {
                        const int off = (m + RT) * W + (n + RT);
#pragma unroll
                        for (int f = 0; f < F; ++f) {
                            const PT p = ph[f][off];
                            jr[f].x = fmaf(ea, p.x, jr[f].x);
                            jr[f].y = fmaf(ea, p.y, jr[f].y);
                            kr[f].x = fmaf(e, p.x, kr[f].x);
                            kr[f].y = fmaf(e, p.y, kr[f].y);
                        }
                    } else {
#pragma unroll
                        for (int f = 0; f < F; ++f) {
                            double a2 = (double)m * spsi[2 * f] + (double)n * spsi[2 * f + 1];
                            float sn, cs;
                            __sincosf((float)(-a2), &sn, &cs);
                            jr[f].x = fmaf(ea, cs, jr[f].x);
                            jr[f].y = fmaf(ea, sn, jr[f].y);
                            kr[f].x = fmaf(e, cs, kr[f].x);
                            kr[f].y = fmaf(e, sn, kr[f].y);
                        }
                    }
                } else {
                    double a = sqrt(dx * dx + dy * dy);
                    double e = exp(-a);
                    double ia = 1.0 / a;
                    if (r <= RT) {
                        const int off = (m + RT) * W + (n + RT);
#pragma unroll
                        for (int f = 0; f < F; ++f) {
                            const PT p = ph[f][off];
                            double tx = e * p.x, ty = e * p.y;
                            jr[f].x += tx * ia;
                            jr[f].y += ty * ia;
                            kr[f].x += tx;
                            kr[f].y += ty;
                        }
                    } else {
#pragma unroll
                        for (int f = 0; f < F; ++f) {
                            double a2 = (double)m * spsi[2 * f] + (double)n * spsi[2 * f + 1];
                            double tx = e * cos(-a2), ty = e * sin(-a2);
                            jr[f].x += tx * ia;
                            jr[f].y += ty * ia;
                            kr[f].x += tx;
                            kr[f].y += ty;
                        }
                    }
                }
            }
#pragma unroll
            for (int f = 0; f < F; ++f) {
                if (!(done & (1u << f))) {
                    cd a = cd{(double)jr[f].x, (double)jr[f].y};
                    cd b = cd{(double)kr[f].x, (double)kr[f].y};
                    js[f] += a;
                    ks[f] += b;
                    if (r > 5) {
                        double c1 = cabs(a / js[f]), c2 = cabs(b / ks[f]);
                        double conv = c1 > c2 ? c1 : c2;
                        if (conv < 1e-8) done |= 1u << f;
                    }
                }
            }
        }
    }
    // Combine the 7 quadrature points of each class in order (src/Zint.jl:117-124).
    const int base_lane = (lane / 7) * 7;
#pragma unroll
    for (int f = 0; f < F; ++f) {
        if (f < nf) {
            cd J = {0, 0}, Jx = {0, 0}, Je = {0, 0}, K = {0, 0}, Kx = {0, 0}, Ke = {0, 0};
#pragma unroll
            for (int q = 0; q < 7; ++q) {
                int src = base_lane + q;
                if (src > 63) src = 63;
                cd jq = cd{__shfl(js[f].x, src, 64), __shfl(js[f].y, src, 64)};
                cd kq = cd{__shfl(ks[f].x, src, 64), __shfl(ks[f].y, src, 64)};
                cd jw = jq * QW[q];
                Jx += jw * QXI[q];
                Je += jw * QETA[q];
                J += jw;
                cd kw = kq * QW[q];
                Kx += kw * QXI[q];
                Ke += kw * QETA[q];
                K += kw;
            }
            if (active && qp == 0) {
                cd *o = jk + f * jk_stride;
                o[0 * nufp + iu] = J;
                o[1 * nufp + iu] = Jx;
                o[2 * nufp + iu] = Je;
                o[3 * nufp + iu] = K;
                o[4 * nufp + iu] = Kx;
                o[5 * nufp + iu] = Ke;
            }
        }
    }
}

// Lattice points summed by the pairs of one lane up to ring r: 8 per ring index.
__device__ __forceinline__ unsigned long long lattice_points(int r) {
    return 4ull * (unsigned long long)r * (unsigned long long)(r + 1);
}

// Sum of a per-lane count over the workgroup, written to out[linear block id].
__device__ void block_count(unsigned long long v, unsigned long long *out) {
    __shared__ unsigned long long part[16];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < nw; ++i) t += part[i];
        out[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
}

__global__ void k_sum_u64(const unsigned long long *a, int64_t n, unsigned long long *total) {
    unsigned long long v = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        v += a[i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(total, v);
}

// Lattice points of rings 0..JK_RINGMAX in ring order: the offset u (m s1 + n s2)
// in the spatial sums' precision T, and (m, n).
// Phasing halves (32 phasings each) per workgroup of the MFMA spatial sums: one.
// Two (512 threads, 64 phasings, the term tile shared by twice the MFMA work) needs
// about 195 VGPRs, runs two waves per SIMD and measured slower on MI355X.
constexpr int JK_MFMA_WH = 1;


template <class T>
__global__ void k_lattice(double us1x, double us1y, double us2x, double us2y, LatPtT<T> *lat) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t E = (int64_t)(2 * JK_RINGMAX + 1) * (2 * JK_RINGMAX + 1);
    if (e >= E) return;
    const int r = ring_of(e);
    int m, n;
    ring_mn(r, (int)(e - ring_base(r)), m, n);
    LatPtT<T> L;
    const double ox = (double)m * us1x + (double)n * us2x;
    const double oy = (double)m * us1y + (double)n * us2y;
    L.ox = (T)ox;
    L.oy = (T)oy;
    L.lx = (T)(ox - (double)L.ox);
    L.ly = (T)(oy - (double)L.oy);
    L.m = m;
    L.n = n;
    lat[e] = L;
}

// MFMA form of the many-phasing spatial sums.  A workgroup of 4 WH wavefronts
// owns 64 rows (9 classes x 7 quadrature points) and 32 WH phasings.  Chunk k
// holds lattice points 1 + 8 k .. 8 k + 8 in ring order.  Per chunk the threads
// form the lattice term pairs (J: e^-a / a, K: e^-a) of the 64 x 8 tile and the
// phases of the 8 x 32 WH tile; the term tile is the A operand and the phase tile
// (cos and sin columns) the B operand of 16x16x4 MFMAs.  The tiles of chunk k + 1
// are formed between the MFMAs of chunk k (double-buffered, one barrier per
// chunk; rows padded against bank conflicts).  Wavefront (wr, wh) accumulates
// rows 16 wr..16 wr + 15 against phasings 32 wh..32 wh + 31; each lane folds the
// ring sums of its 8 (row, phasing) pairs into FP64 running sums at the end of
// ring 5 (rings 1..5 together, in the accumulators) and of every later ring, and
// applies the break test of src/PGF.jl:86-103.  Lattice offsets come from a table
// (k_lattice); phases of rings <= JK_RT are products of per-phasing power tables
// in LDS.  Phasing blocks start at block fblk0; NG = 1 serves a last block with at
// most 16 live phasings, whose second group of 16 would be all padding, with half
// the MFMAs.
template <class T, int WH, int NG>
__device__ __forceinline__ void jk_mfma_body(SheetDev sh, int64_t nufp, double ul,
                                             const LatPtT<T> *lat, const T *phase, int ldp,
                                             int nf, int fblk0, cd *jk, int64_t jk_stride,
                                             unsigned long long *bcnt) {
    typedef MfmaT<T> MT;
    typedef typename MT::V V;
    constexpr bool F32 = std::is_same<T, float>::value;
    constexpr int NT = 256 * WH, FT = 32 * WH, AP = 80, BP = 16 + FT;
    struct Tiles {
        T Aj[2][8][AP], Ak[2][8][AP];
        T Bre[2][8][BP], Bim[2][8][BP];
    };
    struct T2 {
        T x, y;
    };
    constexpr size_t CSB = sizeof(cd) * 64 * 32;
    constexpr size_t LBYTES = sizeof(Tiles) > CSB ? sizeof(Tiles) : CSB;
    __shared__ __align__(16) char lraw[LBYTES];
    __shared__ double urxs[64], urys[64];
    constexpr int PW = 2 * JK_RT + 1;
    __shared__ T P1[2][FT][PW], P2[2][FT][PW];   // [cos/sin][phasing][m + JK_RT]
    Tiles &TL = *reinterpret_cast<Tiles *>(lraw);
    cd(*Cs)[32] = reinterpret_cast<cd(*)[32]>(lraw);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 3, wh = wave >> 2;
    const int64_t u0 = (int64_t)blockIdx.x * 9;
    const int f0 = (fblk0 + (int)blockIdx.y) * FT;
    if (tid < 64) {
        const int r = tid;
        const int64_t iu = u0 + r / 7;
        double urx = 0.0, ury = 0.0;
        if (r < 63 && iu < nufp) {
            const int qp = r % 7;
            int64_t ifm, ifs;
            class_faces(sh, iu, ifm, ifs);
            const double *rs = sh.fxs + 6 * ifs;
            const double *rm = sh.fxs + 6 * ifm;
            double rs12x = rs[0] - rs[2], rs12y = rs[1] - rs[3];
            double rs31x = rs[4] - rs[0], rs31y = rs[5] - rs[1];
            double rs21x = -rs12x, rs21y = -rs12y;
            double rmcx = ((rm[0] + rm[2]) + rm[4]) / 3.0;
            double rmcy = ((rm[1] + rm[3]) + rm[5]) / 3.0;
            double rtx = (rs[0] + rs21x * QXI[qp]) + rs31x * QETA[qp];
            double rty = (rs[1] + rs21y * QXI[qp]) + rs31y * QETA[qp];
            urx = ul * (rmcx - rtx);
            ury = ul * (rmcy - rty);
        }
        urxs[r] = urx;
        urys[r] = ury;
    }
    // exp(-i m psi1) and exp(-i n psi2) for |m|, |n| <= JK_RT, from the phase table's
    // (m, 0) and (0, n) entries; phases of rings <= JK_RT are their products
    for (int t = tid; t < 2 * FT * PW; t += NT) {
        const int which = t / (FT * PW), f = (t / PW) % FT, j = t % PW - JK_RT;
        const int64_t e = which == 0 ? ring_index(j, 0) : ring_index(0, j);
        const T *pp = phase + e * 2 * (int64_t)ldp + 2 * (f0 + f);
        if (which == 0) { P1[0][f][j + JK_RT] = pp[0]; P1[1][f][j + JK_RT] = pp[1]; }
        else { P2[0][f][j + JK_RT] = pp[0]; P2[1][f][j + JK_RT] = pp[1]; }
    }
    __syncthreads();
    // this thread's term row (its lane) relative to the lattice offsets; its
    // phase slot (point pkk, phasing pf)
    // in FP32 both the row's offset and the lattice offsets are pairs hi + lo, so
    // that a distance is their hi difference plus their lo difference: accurate to
    // FP32 relative to the distance itself, also where an observation point lies
    // close to a lattice image of the source point (the hi parts then cancel)
    const T urx = (T)urxs[lane], ury = (T)urys[lane];
    const T urlx = (T)(urxs[lane] - (double)urx), urly = (T)(urys[lane] - (double)ury);
    const int pkk = tid / FT, pf = tid % FT;
    const int pc = lane & 15;
    cd js[2][4], ks[2][4];
    unsigned done = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = 16 * wr + MT::row(lane, e);
        const bool live = r < 63 && u0 + r / 7 < nufp;
        const double arg = sqrt(urxs[r] * urxs[r] + urys[r] * urys[r]);
        const double rterm = exp(-arg);
        const double j0 = arg <= 1e-5 ? (-arg / 6 + 0.5) * arg - 1 : (rterm - 1) / arg;
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            js[g][e] = cd{j0, 0.0};
            ks[g][e] = cd{rterm, 0.0};
            if (!live || f0 + 32 * wh + 16 * g + pc >= nf) done |= 1u << (4 * g + e);
        }
    }
    const T *pcol = phase + 2 * (f0 + pf);
    const int64_t pstride = 2 * (int64_t)ldp;
    const unsigned live0 = ~done & 0xffu;
    unsigned long long npts = 0;     // lattice points of the reference's sums
    int rlast = 0;
    // Chunk k holds lattice points 1 + 8 k .. 8 k + 8 (ring r is chunks
    // r (r - 1) / 2 .. r (r + 1) / 2 - 1).  The tiles of chunk k + 1 are formed
    // while the MFMAs of chunk k run, so that the term arithmetic and the matrix
    // products overlap; lattice data is loaded one chunk earlier still.
    constexpr int NCH = JK_RINGMAX * (JK_RINGMAX + 1) / 2;
    struct T4 {
        T x, y, lx, ly;
    };
    T4 lt[2 / WH];                   // term offsets of the next chunk to be formed
    int2 lq;                         // phase indices of that chunk
    auto load_lat = [&](int k) {
#pragma unroll
        for (int h = 0; h < 2 / WH; ++h) {
            const LatPtT<T> &L = lat[1 + 8 * (int64_t)k + wave + 4 * WH * h];
            lt[h] = T4{L.ox, L.oy, L.lx, L.ly};
        }
        const LatPtT<T> &Q = lat[1 + 8 * (int64_t)k + pkk];
        lq = int2{Q.m, Q.n};
    };
    // the phase tile of chunk k (ring rk) into buffer b, from lq
    auto form_phase = [&](int k, int rk, int b) {
        T br, bi;
        if (rk <= JK_RT) {
            const T c1 = P1[0][pf][lq.x + JK_RT], s1 = P1[1][pf][lq.x + JK_RT];
            const T c2 = P2[0][pf][lq.y + JK_RT], s2 = P2[1][pf][lq.y + JK_RT];
            br = c1 * c2 - s1 * s2;
            bi = c1 * s2 + s1 * c2;
        } else {
            const int64_t e = 1 + 8 * (int64_t)k + pkk;
            br = pcol[e * pstride];
            bi = pcol[e * pstride + 1];
        }
        TL.Bre[b][pkk][pf] = br;
        TL.Bim[b][pkk][pf] = bi;
    };
    // term h of this thread for the chunk in lt, into buffer b
    auto form_term = [&](int b, int h) {
        const int kk = wave + 4 * WH * h;
        if constexpr (F32) {
            const float dx = (urx - lt[h].x) + (urlx - lt[h].lx);
            const float dy = (ury - lt[h].y) + (urly - lt[h].ly);
            float tj, tk;
            term_pair_f32(dx * dx + dy * dy, tj, tk);
            TL.Aj[b][kk][lane] = tj;
            TL.Ak[b][kk][lane] = tk;
        } else {
            const double dx = urx - lt[h].x, dy = ury - lt[h].y;
            const double a = sqrt(dx * dx + dy * dy);
            const double ex = exp(-a);
            TL.Aj[b][kk][lane] = ex * (1.0 / a);
            TL.Ak[b][kk][lane] = ex;
        }
    };
    static_assert(WH == 1, "the interleaved chunk loop forms two terms per thread");
    if (__syncthreads_or(done != 0xffu)) {
        load_lat(0);
        form_phase(0, 1, 0);
        form_term(0, 0);
        form_term(0, 1);
        load_lat(1);
        __syncthreads();
        V aJ[2][2], aK[2][2];        // [g][cos/sin]: this ring's sums
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                aJ[g][c] = V{0, 0, 0, 0};
                aK[g][c] = V{0, 0, 0, 0};
            }
        int r = 1, kend = 1;         // ring r ends with chunk kend - 1
        for (int k = 0; k < NCH; ++k) {
            const int b = k & 1;
            // chunk whose tiles are formed during this one (the last chunk's tiles
            // are formed again at the very end, into a buffer nobody reads)
            const int kn = k + 1 < NCH ? k + 1 : NCH - 1;
            form_phase(kn, k + 1 < kend ? r : r + 1, b ^ 1);
            T a_j[2], a_k[2], bre[2][2], bim[2][2];
#pragma unroll
            for (int k4 = 0; k4 < 2; ++k4) {
                const int kr = 4 * k4 + (lane >> 4);
                a_j[k4] = TL.Aj[b][kr][16 * wr + pc];
                a_k[k4] = TL.Ak[b][kr][16 * wr + pc];
#pragma unroll
                for (int g = 0; g < 2; ++g) {
                    bre[k4][g] = TL.Bre[b][kr][32 * wh + 16 * g + pc];
                    bim[k4][g] = TL.Bim[b][kr][32 * wh + 16 * g + pc];
                }
            }
            // the term tile of chunk kn and the MFMAs of chunk k, interleaved
            // (one MFMA, then two VALU instructions) in two halves whose eight
            // MFMAs each use distinct accumulators
#pragma unroll
            for (int k4 = 0; k4 < 2; ++k4) {
                form_term(b ^ 1, k4);
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    aJ[g][0] = MT::mma(a_j[k4], bre[k4][g], aJ[g][0]);
                    aJ[g][1] = MT::mma(a_j[k4], bim[k4][g], aJ[g][1]);
                    aK[g][0] = MT::mma(a_k[k4], bre[k4][g], aK[g][0]);
                    aK[g][1] = MT::mma(a_k[k4], bim[k4][g], aK[g][1]);
                }
                if (k4 == 1) load_lat(kn + 1 < NCH ? kn + 1 : NCH - 1);
                if (k4 == 0) __builtin_amdgcn_sched_group_barrier(0x100, 12, 0);
#pragma unroll
                for (int i = 0; i < 4 * NG; ++i) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                    __builtin_amdgcn_sched_group_barrier(0x002, F32 ? 2 : 6, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            __syncthreads();
            if (k + 1 < kend) continue;
            // ring r complete: folded into the FP64 running sums (in FP32 every ring
            // on its own, which keeps the FP32 accumulation short) and, after ring 5,
            // tested; FP64 accumulates rings 1..4 with ring 5.
            if (r >= 5 || F32) {
#pragma unroll
                for (int g = 0; g < NG; ++g)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const unsigned bit = 1u << (4 * g + e);
                        if (!(done & bit)) {
                            const cd a = cd{(double)aJ[g][0][e], (double)aJ[g][1][e]};
                            const cd bb = cd{(double)aK[g][0][e], (double)aK[g][1][e]};
                            js[g][e] += a;
                            ks[g][e] += bb;
                            if (r > 5) {
                                // |ring/sum| < 1e-8 for J and K, as squared magnitudes
                                const double aj = a.x * a.x + a.y * a.y;
                                const double sj = js[g][e].x * js[g][e].x + js[g][e].y * js[g][e].y;
                                const double ak = bb.x * bb.x + bb.y * bb.y;
                                const double sk = ks[g][e].x * ks[g][e].x + ks[g][e].y * ks[g][e].y;
                                if (aj < 1e-16 * sj && ak < 1e-16 * sk) {
                                    done |= bit;
                                    npts += lattice_points(r);
                                }
                            }
                        }
                    }
                rlast = r;
                if (!__syncthreads_or(done != 0xffu)) break;
#pragma unroll
                for (int g = 0; g < 2; ++g)
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        aJ[g][c] = V{0, 0, 0, 0};
                        aK[g][c] = V{0, 0, 0, 0};
                    }
            }
            ++r;
            kend += r;
        }
    }
    if (bcnt) {
        npts += (unsigned long long)__popc(live0 & ~done) * lattice_points(rlast);
        block_count(npts, bcnt);
    }
    // J then K, each phasing half in turn: per-row sums through LDS, combined over
    // the 7 quadrature points (src/Zint.jl:117-124)
    for (int which = 0; which < 2; ++which)
        for (int hh = 0; hh < WH; ++hh) {
            __syncthreads();
            if (wh == hh) {
#pragma unroll
                for (int g = 0; g < 2; ++g)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int r = 16 * wr + MT::row(lane, e);
                        Cs[r][16 * g + pc] = which == 0 ? js[g][e] : ks[g][e];
                    }
            }
            __syncthreads();
            for (int t = tid; t < 9 * 32; t += NT) {
                const int u = t / 32, f = t % 32;
                const int64_t iu = u0 + u;
                const int fg = f0 + 32 * hh + f;
                if (iu >= nufp || fg >= nf) continue;
                cd S = {0, 0}, Sx = {0, 0}, Se = {0, 0};
#pragma unroll
                for (int q = 0; q < 7; ++q) {
                    const cd w = Cs[7 * u + q][f] * QW[q];
                    Sx += w * QXI[q];
                    Se += w * QETA[q];
                    S += w;
                }
                cd *o = jk + (int64_t)fg * jk_stride + (int64_t)(3 * which) * nufp;
                o[iu] = S;
                o[nufp + iu] = Sx;
                o[2 * nufp + iu] = Se;
            }
        }
}

template <int WH, int NG>
__global__ __launch_bounds__(256 * WH) __attribute__((amdgpu_waves_per_eu(3))) void k_jk_mfma(
    SheetDev sh, int64_t nufp, double ul, const LatPt *lat, const float *phase, int ldp, int nf,
    int fblk0, cd *jk, int64_t jk_stride, unsigned long long *bcnt) {
    jk_mfma_body<float, WH, NG>(sh, nufp, ul, lat, phase, ldp, nf, fblk0, jk, jk_stride, bcnt);
}

template <int NG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_jk_mfma64(
    SheetDev sh, int64_t nufp, double ul, const LatPtT<double> *lat, const double *phase,
    int ldp, int nf, int fblk0, cd *jk, int64_t jk_stride, unsigned long long *bcnt) {
    jk_mfma_body<double, 1, NG>(sh, nufp, ul, lat, phase, ldp, nf, fblk0, jk, jk_stride, bcnt);
}

// Lattice phase table exp(-i (m psi1 + n psi2)) over rings 0..JK_RINGMAX in ring
// order, one column per phasing (zero beyond nf), interleaved (re, im).
template <class T>
__global__ void k_phase_table(const double *psis, int nf, int ldp, T *phase) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    const int64_t E = (int64_t)(2 * JK_RINGMAX + 1) * (2 * JK_RINGMAX + 1);
    if (e >= E) return;
    int r = ring_of(e);
    int m, n;
    ring_mn(r, (int)(e - ring_base(r)), m, n);
    double c = 0.0, sn = 0.0;
    if (f < nf) {
        double a = (double)m * psis[2 * f] + (double)n * psis[2 * f + 1];
        c = cos(-a);
        sn = sin(-a);
    }
    T *p = phase + (size_t)e * (2 * ldp) + 2 * f;
    p[0] = (T)c;
    p[1] = (T)sn;
}

// Closed-form singular integrals of eqs. (B.3) over the source triangle,
// src/Zint.jl:126-165 (closed = true: every class).
__global__ void k_singular(SheetDev sh, int64_t nufp, double ul, double *rho_r, double *rinv) {
#pragma clang fp contract(off)
    int64_t iu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (iu >= nufp) return;
    int64_t ifm, ifs;
    class_faces(sh, iu, ifm, ifs);
    const double *rsp = sh.fxs + 6 * ifs, *rm = sh.fxs + 6 * ifm;
    double rs[3][2] = {{rsp[0], rsp[1]}, {rsp[2], rsp[3]}, {rsp[4], rsp[5]}};
    double rs32x = rs[2][0] - rs[1][0], rs32y = rs[2][1] - rs[1][1];
    double rs12x = rs[0][0] - rs[1][0], rs12y = rs[0][1] - rs[1][1];
    double area2 = (-rs32y) * rs12x + rs32x * rs12y;
    double unz = 1.0;
    if (area2 < 0) { area2 = fabs(area2); unz = -1.0; }
    double rmcx = ((rm[0] + rm[2]) + rm[4]) / 3.0;
    double rmcy = ((rm[1] + rm[3]) + rm[5]) / 3.0;
    double rrx = 0, rry = 0, ri = 0;
    for (int i = 0; i < 3; ++i) {
        int ip1 = (i + 1) % 3;
        double lx = rs[ip1][0] - rs[i][0], ly = rs[ip1][1] - rs[i][1];
        double ln = sqrt(lx * lx + ly * ly);
        lx /= ln;
        ly /= ln;
        double ux = (-unz) * (-ly), uy = (-unz) * lx;
        double p0 = ux * (rs[i][0] - rmcx) + uy * (rs[i][1] - rmcy);
        double p0sq = p0 * p0;
        double lp = lx * (rs[ip1][0] - rmcx) + ly * (rs[ip1][1] - rmcy);
        double lm = lx * (rs[i][0] - rmcx) + ly * (rs[i][1] - rmcy);
        double ledge = lp - lm;
        double pplus = sqrt(p0sq + lp * lp), pminus = sqrt(p0sq + lm * lm);
        double factor;
        if (fabs(p0 / ledge) < 1e-4) {
            factor = lp * pplus - lm * pminus;
        } else {
            factor = p0 * log((pplus + lp) / (pminus + lm));
            ri += factor;
            factor = (factor * p0 + pplus * lp) - pminus * lm;
        }
        rrx += ux * factor;
        rry += uy * factor;
    }
    rrx *= 0.5;
    rry *= 0.5;
    rrx /= area2;
    rry /= area2;
    ri /= (area2 * ul);
    rho_r[2 * iu] = rrx;
    rho_r[2 * iu + 1] = rry;
    rinv[iu] = ri;
}

}  // namespace

void launch_summands(const TabParams *d_tp, int npts, int64_t cap, int r0, int r1, cd *t1,
                     cd *t2, hipStream_t st) {
    int64_t e0 = r0 == 0 ? 0 : (int64_t)(2 * r0 - 1) * (2 * r0 - 1);
    int64_t e1 = (int64_t)(2 * r1 + 1) * (2 * r1 + 1);
    if (e1 > cap) throw ArgError("ring capacity exceeded");
    int64_t cnt = e1 - e0;
    dim3 grid((unsigned)((cnt + 127) / 128), (unsigned)npts);
    hipLaunchKernelGGL(k_summands, grid, dim3(128), 0, st, d_tp, cap, e0, e1, t1, t2);
    HIPCHECK(hipGetLastError());
}

void launch_summands_dk0(const TabParams *d_tp, int npts, int64_t cap, int r0, int r1, cd *t1,
                         cd *t2, hipStream_t st) {
    int64_t e0 = r0 == 0 ? 0 : (int64_t)(2 * r0 - 1) * (2 * r0 - 1);
    int64_t e1 = (int64_t)(2 * r1 + 1) * (2 * r1 + 1);
    if (e1 > cap) throw ArgError("ring capacity exceeded");
    int64_t cnt = e1 - e0;
    dim3 grid((unsigned)((cnt + 127) / 128), (unsigned)npts);
    hipLaunchKernelGGL(k_summands_dk0, grid, dim3(128), 0, st, d_tp, cap, e0, e1, t1, t2);
    HIPCHECK(hipGetLastError());
}

void launch_ringsums(const TabParams *, int npts, int64_t cap, int r0, int r1, int rcap,
                     const cd *t1, const cd *t2, double convtest, int *flags, hipStream_t st) {
    int cnt = r1 - r0 + 1;
    dim3 grid((unsigned)((cnt + 63) / 64), (unsigned)npts);
    hipLaunchKernelGGL(k_ringsums, grid, dim3(64), 0, st, cap, r0, r1, rcap, t1, t2, convtest,
                       flags);
    HIPCHECK(hipGetLastError());
}

void launch_gather_square(int npts, const int *d_M, int Mmax, int64_t cap, const cd *t1,
                          const cd *t2, cd *S, int64_t s_stride, hipStream_t st) {
    dim3 blk(16, 16);
    dim3 grid((unsigned)((Mmax + 15) / 16), (unsigned)((Mmax + 15) / 16), (unsigned)npts);
    hipLaunchKernelGGL(k_gather_square, grid, blk, 0, st, d_M, cap, t1, t2, S, s_stride);
    HIPCHECK(hipGetLastError());
}

void launch_dft_matrix(int M, cd *W, hipStream_t st) {
    dim3 blk(16, 16);
    dim3 grid((unsigned)((M + 15) / 16), (unsigned)((M + 15) / 16));
    hipLaunchKernelGGL(k_dft_matrix, grid, blk, 0, st, M, W);
    HIPCHECK(hipGetLastError());
}

void launch_phase_ghost(int npts, const int *d_M, const double *d_psi, const int *d_mag,
                        const double *d_area, const cd *F, int64_t f_stride, cd *tab,
                        int64_t tab_stride, int Mmax, hipStream_t st) {
    dim3 blk(16, 16);
    dim3 grid((unsigned)((Mmax + 2 + 15) / 16), (unsigned)((Mmax + 2 + 15) / 16),
              (unsigned)npts);
    hipLaunchKernelGGL(k_phase_ghost, grid, blk, 0, st, d_M, d_psi, d_mag, d_area, F, f_stride,
                       tab, tab_stride);
    HIPCHECK(hipGetLastError());
}

// Few phasings: the per-lane kernel (phases in LDS, F per launch).  Many
// phasings: the tiled kernel, whose lattice-term tile is shared by 32 phasings.
void launch_jk(const SheetDev &sh, int64_t nufp, double ulocal, const double *us1,
               const double *us2, const double *d_psi, int nf, cd *jk, int64_t jk_stride,
               int mixed, hipStream_t st, unsigned long long *total) {
    if (nf < 1) return;
    if (nf < JK_MFMA_MIN) {
        const int per = mixed ? 16 : 8;
        dim3 grid((unsigned)((nufp + 35) / 36));
        // the mixed kernels read the lattice offsets from a table (stream-ordered
        // allocation: nothing waits)
        LatPt *lat = nullptr;
        if (mixed) {
            const int64_t E = (int64_t)(2 * JK_RINGMAX + 1) * (2 * JK_RINGMAX + 1);
            if (hipMallocAsync((void **)&lat, (size_t)E * sizeof(LatPt), st) != hipSuccess) {
                (void)hipGetLastError();
                throw HipError("hipMallocAsync of the lattice offset table failed");
            }
            hipLaunchKernelGGL(k_lattice<float>, dim3((unsigned)((E + 255) / 256)), dim3(256), 0,
                               st, us1[0], us1[1], us2[0], us2[1], lat);
            HIPCHECK(hipGetLastError());
        }
        for (int g = 0; g < nf; g += per) {
            const int n = nf - g < per ? nf - g : per;
            const double *ps = d_psi + 2 * g;
            cd *o = jk + (int64_t)g * jk_stride;
#define JKL(FF, MX)                                                                        \
    hipLaunchKernelGGL((k_jk<FF, MX>), grid, dim3(256), 0, st, sh, nufp, ulocal, us1[0],  \
                       us1[1], us2[0], us2[1], ps, n, o, jk_stride, lat)
            if (mixed) {
                if (n <= 1) JKL(1, true); else if (n <= 2) JKL(2, true);
                else if (n <= 4) JKL(4, true); else if (n <= 8) JKL(8, true);
                else JKL(16, true);
            } else {
                if (n <= 1) JKL(1, false); else if (n <= 2) JKL(2, false);
                else if (n <= 4) JKL(4, false); else JKL(8, false);
            }
#undef JKL
            HIPCHECK(hipGetLastError());
        }
        if (lat) HIPCHECK(hipFreeAsync(lat, st));
        return;
    }
    // phasings per workgroup: 32 JK_MFMA_WH (FP32) or 32 (FP64)
    const int FT = mixed ? 32 * JK_MFMA_WH : 32;
    const int ldp = (nf + FT - 1) / FT * FT;
    const int64_t E = (int64_t)(2 * JK_RINGMAX + 1) * (2 * JK_RINGMAX + 1);
    const size_t elt = mixed ? sizeof(float) : sizeof(double);
    dim3 pg((unsigned)((E + 255) / 256), (unsigned)ldp);
    dim3 grid((unsigned)((nufp + 8) / 9), (unsigned)(ldp / FT));
    const int64_t nblk = (int64_t)grid.x * grid.y;
    const size_t phase_bytes = (size_t)E * 2 * ldp * elt;
    const size_t lat_bytes = (size_t)E * (mixed ? sizeof(LatPtT<float>) : sizeof(LatPtT<double>));
    const size_t cnt_bytes = total ? (size_t)nblk * sizeof(unsigned long long) : 0;
    void *ph = nullptr;
    if (hipMallocAsync(&ph, phase_bytes + lat_bytes + cnt_bytes, st) != hipSuccess) {
        (void)hipGetLastError();
        throw HipError("hipMallocAsync of the lattice phase table failed");
    }
    void *lat = (char *)ph + phase_bytes;
    unsigned long long *bcnt =
        total ? reinterpret_cast<unsigned long long *>((char *)ph + phase_bytes + lat_bytes)
              : nullptr;
    const dim3 lg((unsigned)((E + 255) / 256));
    if (mixed) {
        hipLaunchKernelGGL(k_phase_table<float>, pg, dim3(256), 0, st, d_psi, nf, ldp, (float *)ph);
        hipLaunchKernelGGL(k_lattice<float>, lg, dim3(256), 0, st, us1[0], us1[1], us2[0], us2[1],
                           (LatPtT<float> *)lat);
    } else {
        hipLaunchKernelGGL(k_phase_table<double>, pg, dim3(256), 0, st, d_psi, nf, ldp,
                           (double *)ph);
        hipLaunchKernelGGL(k_lattice<double>, lg, dim3(256), 0, st, us1[0], us1[1], us2[0], us2[1],
                           (LatPtT<double> *)lat);
    }
    HIPCHECK(hipGetLastError());
    // a last block with at most 16 live phasings runs the half-width kernel
    const int nblk_f = (int)grid.y;
    const bool half_tail = (!mixed || JK_MFMA_WH == 1) && nf - (nblk_f - 1) * FT <= 16;
    const dim3 gmain(grid.x, (unsigned)(half_tail ? nblk_f - 1 : nblk_f)), gtail(grid.x, 1);
    unsigned long long *btail = bcnt ? bcnt + (int64_t)(nblk_f - 1) * grid.x : nullptr;
    if (mixed) {
        const LatPtT<float> *lf = (const LatPtT<float> *)lat;
        const float *pf = (const float *)ph;
        if (gmain.y > 0)
            hipLaunchKernelGGL((k_jk_mfma<JK_MFMA_WH, 2>), gmain, dim3(256 * JK_MFMA_WH), 0, st,
                               sh, nufp, ulocal, lf, pf, ldp, nf, 0, jk, jk_stride, bcnt);
        if (half_tail) {
            HIPCHECK(hipGetLastError());
            hipLaunchKernelGGL((k_jk_mfma<JK_MFMA_WH, 1>), gtail, dim3(256 * JK_MFMA_WH), 0, st,
                               sh, nufp, ulocal, lf, pf, ldp, nf, nblk_f - 1, jk, jk_stride,
                               btail);
        }
    } else {
        const LatPtT<double> *ld = (const LatPtT<double> *)lat;
        const double *pd = (const double *)ph;
        if (gmain.y > 0)
            hipLaunchKernelGGL((k_jk_mfma64<2>), gmain, dim3(256), 0, st, sh, nufp, ulocal, ld,
                               pd, ldp, nf, 0, jk, jk_stride, bcnt);
        if (half_tail) {
            HIPCHECK(hipGetLastError());
            hipLaunchKernelGGL((k_jk_mfma64<1>), gtail, dim3(256), 0, st, sh, nufp, ulocal, ld, pd,
                               ldp, nf, nblk_f - 1, jk, jk_stride, btail);
        }
    }
    HIPCHECK(hipGetLastError());
    if (total) {
        const unsigned sg =
            (unsigned)std::max<int64_t>(1, std::min<int64_t>((nblk + 255) / 256, 1024));
        hipLaunchKernelGGL(k_sum_u64, dim3(sg), dim3(256), 0, st, bcnt, nblk, total);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipFreeAsync(ph, st));
}

void launch_singular(const SheetDev &sh, int64_t nufp, double ulocal, double *rho_r,
                     double *rinv, hipStream_t st) {
    dim3 grid((unsigned)((nufp + 255) / 256));
    hipLaunchKernelGGL(k_singular, grid, dim3(256), 0, st, sh, nufp, ulocal, rho_r, rinv);
    HIPCHECK(hipGetLastError());
}
