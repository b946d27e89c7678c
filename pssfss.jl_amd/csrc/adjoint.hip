// Adjoint (sensitivity) kernels for gfx950: derivatives of the sheet GSM with
// respect to mesh geometry, without storing dZ.  Each geometry direction is
// carried through the fill's ingredients in forward mode (dual numbers):
//   * spatial sums (jksums / filljk!, src/PGF.jl:67-106, src/Zint.jl:54-125)
//   * analytic singular integrals (src/Zint.jl:126-165)
//   * class terms of fillz (src/FillZY.jl:50-245) with the table interpolation
//     (src/PGF.jl:392-443), contracted on the fly against the adjoint and forward
//     solutions: sum_{m,s} dZ_ms W_ms with W = A X^T of rank <= #modes
//   * RWG Fourier transforms (rwgbfft!, src/RWG.jl:428-512)
// plus the transposed triangular solves Z^T lambda = g with the LU factors.
// Integer and branch decisions are taken on values, as the value kernels take
// them, so derivatives are those of the same piecewise-smooth formulas.
#include "internal.hpp"
#include "dual.hpp"
#include "geom.hpp"

namespace {

constexpr double PI_ = 3.141592653589793;
constexpr double TWOPI = 6.283185307179586;
constexpr double FOURPI = 12.566370614359172;
constexpr double MU0 = 1.25663706212e-6;
constexpr double C0 = 299792458.0;

struct DVec {
    dd x, y;
};

// Face vertices as duals: values from the sheet, derivatives from the direction.
__device__ __forceinline__ void face_dual(const double *fx, const double *ft, DVec v[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        v[a].x = dd{fx[2 * a], ft[2 * a]};
        v[a].y = dd{fx[2 * a + 1], ft[2 * a + 1]};
    }
}

// ------------------------------------------------------------ spatial sums
// One lane per (class, quadrature point), 9 classes x 7 points per wavefront; one
// phasing.  Output per direction: J, J_xi, J_eta, K, K_xi, K_eta per class as duals.
// The same spatial sums with the derivatives along up to ND directions at once:
// the distances, exponentials and phases are those of the value and are shared;
// direction d adds d(dist) = (r . dr_d) / |r| and the chain rule of each term.
// ftan holds nd blocks of the faces' vertex directions (ftan_stride doubles
// apart); out holds nd blocks of 6 nufp duals (out_stride apart), each as the
// single-direction kernel writes them.
template <int ND>
__global__ __launch_bounds__(256) void k_jk_tan_multi(SheetDev sh, int64_t nufp, double ul,
                                                      double us1x, double us1y, double us2x,
                                                      double us2y, double psi1, double psi2,
                                                      const double *ftan, int64_t ftan_stride,
                                                      int nd, cdd *out, int64_t out_stride) {
    __shared__ cd sj[4][64][ND + 1], sk[4][64][ND + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int qp = lane % 7;
    const int64_t iu = ((int64_t)blockIdx.x * 4 + wave) * 9 + lane / 7;
    const bool active = lane < 63 && iu < nufp;
    cd js = {0, 0}, ks = {0, 0}, jsd[ND], ksd[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) jsd[d] = ksd[d] = cd{0, 0};
    if (active) {
        int64_t ifm, ifs;
        class_faces(sh, iu, ifm, ifs);
        const double *fs = sh.fxs + 6 * ifs, *fm = sh.fxs + 6 * ifm;
        const double rmcx = ((fm[0] + fm[2]) + fm[4]) / 3.0;
        const double rmcy = ((fm[1] + fm[3]) + fm[5]) / 3.0;
        const double rtx = (fs[0] + (fs[2] - fs[0]) * QXI[qp]) + (fs[4] - fs[0]) * QETA[qp];
        const double rty = (fs[1] + (fs[3] - fs[1]) * QXI[qp]) + (fs[5] - fs[1]) * QETA[qp];
        const double ux = ul * (rmcx - rtx), uy = ul * (rmcy - rty);
        double dux[ND], duy[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            dux[d] = duy[d] = 0.0;
            if (d < nd) {
                const double *ts = ftan + d * ftan_stride + 6 * ifs;
                const double *tm = ftan + d * ftan_stride + 6 * ifm;
                const double dmx = ((tm[0] + tm[2]) + tm[4]) / 3.0;
                const double dmy = ((tm[1] + tm[3]) + tm[5]) / 3.0;
                const double dtx = (ts[0] + (ts[2] - ts[0]) * QXI[qp]) + (ts[4] - ts[0]) * QETA[qp];
                const double dty = (ts[1] + (ts[3] - ts[1]) * QXI[qp]) + (ts[5] - ts[1]) * QETA[qp];
                dux[d] = ul * (dmx - dtx);
                duy[d] = ul * (dmy - dty);
            }
        }
        // the extracted (0, 0) term
        const double arg = sqrt(ux * ux + uy * uy);
        const double rterm = exp(-arg);
        const bool small = arg <= 1e-5;
        const double j0 = small ? (-arg / 6.0 + 0.5) * arg - 1.0 : (rterm - 1.0) / arg;
        const double dj0 = small ? -arg / 3.0 + 0.5 : (-rterm / arg) - (rterm - 1.0) / (arg * arg);
        js = cd{j0, 0.0};
        ks = cd{rterm, 0.0};
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const double darg = arg > 0 ? (ux * dux[d] + uy * duy[d]) / arg : 0.0;
            jsd[d] = cd{dj0 * darg, 0.0};
            ksd[d] = cd{-rterm * darg, 0.0};
        }
        const bool zero_phase = psi1 == 0.0 && psi2 == 0.0;
        for (int r = 1; r <= JK_RINGMAX; ++r) {
            cd jr = {0, 0}, kr = {0, 0}, jrd[ND], krd[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d) jrd[d] = krd[d] = cd{0, 0};
            for (int k = 0; k < 8 * r; ++k) {
                int m, n;
                ring_mn(r, k, m, n);
                const double dx = ux - ((double)m * us1x + (double)n * us2x);
                const double dy = uy - ((double)m * us1y + (double)n * us2y);
                const double a = sqrt(dx * dx + dy * dy);
                const double e = exp(-a);
                const double ea = e / a;
                // d(ea) = -(e + ea) da / a, d(e) = -e da, da = (dx ddx + dy ddy) / a
                const double ia = 1.0 / a;
                const double fj = -(e + ea) * ia * ia, fk = -e * ia;
                cd ph = {1.0, 0.0};
                if (!zero_phase) ph = cisd(-((double)m * psi1 + (double)n * psi2));
                jr += ph * ea;
                kr += ph * e;
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const double dd_ = dx * dux[d] + dy * duy[d];
                    jrd[d] += ph * (fj * dd_);
                    krd[d] += ph * (fk * dd_);
                }
            }
            js += jr;
            ks += kr;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                jsd[d] += jrd[d];
                ksd[d] += krd[d];
            }
            if (r > 5) {
                double c1 = cabs(jr / js), c2 = cabs(kr / ks);
                if ((c1 > c2 ? c1 : c2) < 1e-8) break;
            }
        }
    }
    sj[wave][lane][0] = js;
    sk[wave][lane][0] = ks;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        sj[wave][lane][d + 1] = jsd[d];
        sk[wave][lane][d + 1] = ksd[d];
    }
    __syncthreads();
    if (active && qp == 0) {
        const int base = lane;        // lanes base..base+6 hold this class
        for (int d = 0; d < nd; ++d) {
            cdd J = mkc(cd{0, 0}), Jx = J, Je = J, K = J, Kx = J, Ke = J;
            for (int q = 0; q < 7; ++q) {
                const cd jv = sj[wave][base + q][0], jt = sj[wave][base + q][d + 1];
                const cd kv = sk[wave][base + q][0], kt = sk[wave][base + q][d + 1];
                cdd jw = cdd{dd{jv.x, jt.x}, dd{jv.y, jt.y}} * QW[q];
                Jx += jw * QXI[q];
                Je += jw * QETA[q];
                J += jw;
                cdd kw = cdd{dd{kv.x, kt.x}, dd{kv.y, kt.y}} * QW[q];
                Kx += kw * QXI[q];
                Ke += kw * QETA[q];
                K += kw;
            }
            cdd *o = out + d * out_stride;
            o[0 * nufp + iu] = J;
            o[1 * nufp + iu] = Jx;
            o[2 * nufp + iu] = Je;
            o[3 * nufp + iu] = K;
            o[4 * nufp + iu] = Kx;
            o[5 * nufp + iu] = Ke;
        }
    }
}

// ------------------------------------------------------- singular integrals
__global__ void k_singular_tan(SheetDev sh, int64_t nufp, double ul, const double *ftan,
                               dd *rho_r, dd *rinv) {
    int64_t iu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (iu >= nufp) return;
    int64_t ifm, ifs;
    class_faces(sh, iu, ifm, ifs);
    DVec rs[3], rm[3];
    face_dual(sh.fxs + 6 * ifs, ftan + 6 * ifs, rs);
    face_dual(sh.fxs + 6 * ifm, ftan + 6 * ifm, rm);
    dd rs32x = rs[2].x - rs[1].x, rs32y = rs[2].y - rs[1].y;
    dd rs12x = rs[0].x - rs[1].x, rs12y = rs[0].y - rs[1].y;
    dd area2 = (-rs32y) * rs12x + rs32x * rs12y;
    double unz = 1.0;
    if (area2.v < 0) { area2 = -area2; unz = -1.0; }
    dd rmcx = ((rm[0].x + rm[1].x) + rm[2].x) / 3.0;
    dd rmcy = ((rm[0].y + rm[1].y) + rm[2].y) / 3.0;
    dd rrx = mkd(0), rry = mkd(0), ri = mkd(0);
    for (int i = 0; i < 3; ++i) {
        int ip1 = (i + 1) % 3;
        dd lx = rs[ip1].x - rs[i].x, ly = rs[ip1].y - rs[i].y;
        dd ln = dsqrt(lx * lx + ly * ly);
        lx = lx / ln;
        ly = ly / ln;
        dd ux = (-unz) * (-ly), uy = (-unz) * lx;
        dd p0 = ux * (rs[i].x - rmcx) + uy * (rs[i].y - rmcy);
        dd p0sq = p0 * p0;
        dd lp = lx * (rs[ip1].x - rmcx) + ly * (rs[ip1].y - rmcy);
        dd lm = lx * (rs[i].x - rmcx) + ly * (rs[i].y - rmcy);
        dd ledge = lp - lm;
        dd pplus = dsqrt(p0sq + lp * lp), pminus = dsqrt(p0sq + lm * lm);
        dd factor;
        if (fabs(p0.v / ledge.v) < 1e-4) {
            factor = lp * pplus - lm * pminus;
        } else {
            factor = p0 * dlog((pplus + lp) / (pminus + lm));
            ri = ri + factor;
            factor = (factor * p0 + pplus * lp) - pminus * lm;
        }
        rrx = rrx + ux * factor;
        rry = rry + uy * factor;
    }
    rrx = rrx * 0.5;
    rry = rry * 0.5;
    rrx = rrx / area2;
    rry = rry / area2;
    ri = ri / (area2 * ul);
    rho_r[2 * iu] = rrx;
    rho_r[2 * iu + 1] = rry;
    rinv[iu] = ri;
}

// --------------------------------------------------------- class terms
__device__ cdd sigma_interp_d(const cd *tab, int M, const double *b1, const double *b2,
                              double psi1, double psi2, dd rx, dd ry) {
    dd x1o = (b1[0] * rx + b1[1] * ry) / TWOPI;
    if (fabs(x1o.v) < 1e-8) x1o = mkd(0.0);
    dd x2o = (b2[0] * rx + b2[1] * ry) / TWOPI;
    if (fabs(x2o.v) < 1e-8) x2o = mkd(0.0);
    dd x1 = x1o - floor(x1o.v);
    dd x2 = x2o - floor(x2o.v);
    double msh = rint(x1o.v - x1.v), nsh = rint(x2o.v - x2.v);
    int m = (int)(M * x1.v);
    dd p = M * x1 - (double)m;
    int n = (int)(M * x2.v);
    dd q = M * x2 - (double)n;
    if (m > M - 1) m = M - 1;
    if (n > M - 1) n = M - 1;
    dd p2 = p * p, q2 = q * q, pq = p * q;
    const int64_t E = M + 2;
#define T(a, b) tab[(int64_t)((a) + 1) + E * (int64_t)((b) + 1)]
    cdd sm = (0.5 * (q2 - q)) * T(m, n - 1);
    sm = sm + (0.5 * (p2 - p)) * T(m - 1, n);
    sm = sm + (((1.0 + pq) - p2) - q2) * T(m, n);
    sm = sm + (0.5 * ((p2 - 2.0 * pq) + p)) * T(m + 1, n);
    sm = sm + (0.5 * ((q2 - 2.0 * pq) + q)) * T(m, n + 1);
    sm = sm + pq * T(m + 1, n + 1);
#undef T
    if (msh != 0 || nsh != 0) sm = sm * cisd(-(msh * psi1 + nsh * psi2));
    return sm;
}

// One lane per face pair (every pair its own class): derivative of the nine
// flag-free terms w[iml][isl] along the direction, contracted as
//   sum conj-flagged dw * sum_c A_o[m, c] X[s, c]
// for each objective o.  Per-workgroup partial sums go to part[block][o].  The
// terms are k_fill_classes' (electric or magnetic, src/FillZY.jl:183-200 and
// :385-400) in dual numbers; self pairs of a resistive face add the derivative of
// the surface-resistance term of k_fill_gather (src/FillZY.jl:221-231), which moves
// with the face's area and edge lengths.
template <bool MAGNETIC>
__global__ __launch_bounds__(256) void k_class_tan(SheetDev sh, const FillParams *fps,
                                                   const cdd *jkd, const dd *rho_r,
                                                   const dd *rinv_, const double *ftan_m,
                                                   const cd *A, int64_t sAo, const cd *X,
                                                   int ncol, int nobj, int has_fr, cd *part) {
    __shared__ cd red[4][4];
    const int64_t iu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const FillParams &P = *fps;
    const int64_t nbf = sh.nbf;
    cd acc[4] = {cd{0, 0}, cd{0, 0}, cd{0, 0}, cd{0, 0}};
    if (iu < sh.nufp) {
        const double EPS0 = 1.0 / (MU0 * (C0 * C0));
        int64_t ifm, ifs;
        class_faces(sh, iu, ifm, ifs);
        DVec rs[3], rm[3];
        face_dual(sh.fxm + 6 * ifs, ftan_m + 6 * ifs, rs);
        face_dual(sh.fxm + 6 * ifm, ftan_m + 6 * ifm, rm);
        const dd rmcx = ((rm[0].x + rm[1].x) + rm[2].x) / 3.0;
        const dd rmcy = ((rm[0].y + rm[1].y) + rm[2].y) / 3.0;
        cdd I1x = mkc(cd{0, 0}), I1e = I1x, I1 = I1x, I2 = I1x;
        for (int i = 0; i < 7; ++i) {
            dd rtx = (rs[0].x + (rs[1].x - rs[0].x) * QXI[i]) + (rs[2].x - rs[0].x) * QETA[i];
            dd rty = (rs[0].y + (rs[1].y - rs[0].y) * QXI[i]) + (rs[2].y - rs[0].y) * QETA[i];
            dd rdx = rmcx - rtx, rdy = rmcy - rty;
            cdd s1 = sigma_interp_d(P.tab1, P.M, P.b1, P.b2, P.psi1, P.psi2, rdx, rdy);
            cdd s2 = sigma_interp_d(P.tab2, P.M, P.b1, P.b2, P.psi1, P.psi2, rdx, rdy);
            cdd sw = s1 * QW[i];
            I1x += sw * QXI[i];
            I1e += sw * QETA[i];
            I1 += sw;
            I2 += s2 * QW[i];
        }
        const int64_t nu = sh.nufp;
        const cdd J = jkd[iu], Jx = jkd[nu + iu], Je = jkd[2 * nu + iu];
        const cdd K = jkd[3 * nu + iu], Kx = jkd[4 * nu + iu], Ke = jkd[5 * nu + iu];
        const dd ri = rinv_[iu];
        const dd rrx = rho_r[2 * iu], rry = rho_r[2 * iu + 1];
        const cdd I1z = (I1 - I1x) - I1e, Jz = (J - Jx) - Je, Kz = (K - Kx) - Ke;
        dd rc2[3][2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            rc2[i][0] = 0.5 * (rmcx - rm[i].x);
            rc2[i][1] = 0.5 * (rmcy - rm[i].y);
        }
        const double u = P.u;
        cdd phi_core, Phi;
        cd asf;
        if (!MAGNETIC) {
            phi_core = ((FOURPI * I2) + u * (ri + J)) + (P.d3 / u) * K;
            Phi = (P.psf_base * (1.0 / ((TWOPI * P.omega) * EPS0))) * phi_core;
            asf = (MU0 / FOURPI) * P.mut;
        } else {
            phi_core = ((P.I2fact * I2) + u * (ri + J)) + P.KPfact * K;
            Phi = P.psf_base * phi_core;
            asf = (-EPS0 / PI_) * P.epsb;
        }
        const cd iw = MAGNETIC ? cd{0.0, -P.omega} : cd{0.0, P.omega};
        cd dw[3][3];
#pragma unroll
        for (int isl = 0; isl < 3; ++isl) {
            const dd sx = rrx + ((u * ri) * (rmcx - rs[isl].x));
            const dd sy = rry + ((u * ri) * (rmcy - rs[isl].y));
            cdd Ai[2];
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const dd r0 = d ? rs[0].y : rs[0].x, r1 = d ? rs[1].y : rs[1].x;
                const dd r2 = d ? rs[2].y : rs[2].x, ri_ = d ? rs[isl].y : rs[isl].x;
                cdd I1i = ((r0 * I1x + r1 * I1e) + r2 * I1z) - ri_ * I1;
                cdd Ji = ((r0 * Jx + r1 * Je) + r2 * Jz) - ri_ * J;
                cdd Ki = ((r0 * Kx + r1 * Ke) + r2 * Kz) - ri_ * K;
                if (!MAGNETIC)
                    Ai[d] = asf * ((((FOURPI * I1i) + (d ? sy : sx)) + u * Ji) + (P.c3 / u) * Ki);
                else
                    Ai[d] = asf * ((((P.I1fact * I1i) + (d ? sy : sx)) + u * Ji) + P.KFfact * Ki);
            }
#pragma unroll
            for (int iml = 0; iml < 3; ++iml) {
                const cdd dot = rc2[iml][0] * Ai[0] + rc2[iml][1] * Ai[1];
                dw[iml][isl] = der(iw * dot - Phi);
            }
        }
        // surface resistance of a self pair: q = R / (48 area), edge lengths squared
        // l2[i] of the edge opposite vertex i
        double dr[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (!MAGNETIC && has_fr && ifm == ifs && sh.fr[ifs] != 0) {
            const dd r32x = rs[2].x - rs[1].x, r32y = rs[2].y - rs[1].y;
            const dd r12x = rs[0].x - rs[1].x, r12y = rs[0].y - rs[1].y;
            const dd r13x = rs[0].x - rs[2].x, r13y = rs[0].y - rs[2].y;
            const dd area48 = 48.0 * dfabs(0.5 * (r32x * r12y - r32y * r12x));
            const dd l2[3] = {r32x * r32x + r32y * r32y, r13x * r13x + r13y * r13y,
                              r12x * r12x + r12y * r12y};
            const dd q = sh.fr[ifs] / area48;
#pragma unroll
            for (int lm = 0; lm < 3; ++lm)
#pragma unroll
                for (int ls = 0; ls < 3; ++ls) {
                    if (lm == ls) {
                        const int nx = ls == 2 ? 0 : ls + 1, pv = ls == 0 ? 2 : ls - 1;
                        dr[lm][ls] = (q * (3.0 * (l2[nx] + l2[pv]) - l2[ls])).d;
                    } else {
                        const int t3 = 3 - lm - ls;
                        dr[lm][ls] = (q * ((l2[ls] + l2[lm]) - 3.0 * l2[t3])).d;
                    }
                }
        }
        for (int lm = 0; lm < 3; ++lm) {
            const int m = sh.fbf[3 * ifm + lm];
            if (m < 0) continue;
            const cd ffm = conj(P.ff[sh.feci[3 * ifm + lm]]);
            const cd hm = sh.fsg[3 * ifm + lm] > 0 ? ffm : -ffm;
            for (int ls = 0; ls < 3; ++ls) {
                const int s = sh.fbf[3 * ifs + ls];
                if (s < 0) continue;
                const cd ffs = P.ff[sh.feci[3 * ifs + ls]];
                const cd hs = sh.fsg[3 * ifs + ls] > 0 ? ffs : -ffs;
                cd z = hm * (hs * dw[lm][ls]);
                if (lm == ls) z = z + dr[lm][ls];
                else z = z + (hs * hm) * dr[lm][ls];
                for (int o = 0; o < nobj; ++o) {
                    cd wms = {0, 0};
                    for (int c = 0; c < ncol; ++c)
                        wms += A[o * sAo + m + nbf * c] * X[s + nbf * c];
                    acc[o] += z * wms;
                }
            }
        }
    }
    for (int o = 0; o < 4; ++o) {
        for (int off = 32; off > 0; off >>= 1) {
            acc[o].x += __shfl_xor(acc[o].x, off, 64);
            acc[o].y += __shfl_xor(acc[o].y, off, 64);
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int o = 0; o < 4; ++o) red[w][o] = acc[o];
    __syncthreads();
    if (threadIdx.x < nobj) {
        cd t = {0, 0};
        for (int i = 0; i < 4; ++i) t += red[i][threadIdx.x];
        part[(int64_t)blockIdx.x * nobj + threadIdx.x] = t;
    }
}

__global__ void k_sum_cd(const cd *part, int64_t nblk, int nobj, cd *out) {
    const int o = blockIdx.x;
    cd v = {0, 0};
    for (int64_t i = threadIdx.x; i < nblk; i += blockDim.x) v += part[i * nobj + o];
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
    }
    __shared__ cd red[16];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        cd t = {0, 0};
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
        out[o] = t;
    }
}

// ------------------------------------------------------ RWG transforms
__device__ __forceinline__ dd sj0d(dd x) {
    if (fabs(x.v) < 1e-3) { dd x2 = x * x; return 1.0 - x2 / 6.0 * (1.0 - x2 / 20.0); }
    return dsin(x) / x;
}
__device__ __forceinline__ dd sj1d(dd x) {
    if (fabs(x.v) < 1e-3) return x * (1.0 / 3 - x * x / 30.0);
    return (dsin(x) / x - dcos(x)) / x;
}

// Derivative of one face's transform (local edge i), cf. face_ft of the value path.
__device__ void face_ft_tan(const double *fxm, const double *ftm, int i, const BfftJob &Jb,
                            cd ff, cd out[2]) {
    const double k0 = Jb.k[0], k1 = Jb.k[1];
    DVec r[3];
    face_dual(fxm, ftm, r);
    DVec lv[3], rc[3];
    double lmax = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int nx = (a + 1) % 3;
        lv[a].x = r[nx].x - r[a].x;
        lv[a].y = r[nx].y - r[a].y;
        rc[a].x = r[a].x + 0.5 * lv[a].x;
        rc[a].y = r[a].y + 0.5 * lv[a].y;
        double ln = sqrt(lv[a].x.v * lv[a].x.v + lv[a].y.v * lv[a].y.v);
        if (ln > lmax) lmax = ln;
    }
    double kmagsq = k0 * k0 + k1 * k1;
    double kmag = sqrt(kmagsq);
    if (kmag * lmax < 1e-4) {
        dd cx = ((r[0].x + r[1].x) + r[2].x) / 3.0, cy = ((r[0].y + r[1].y) + r[2].y) / 3.0;
        cdd fac = dcis(k0 * cx + k1 * cy) * ff;
        out[0] = der((0.5 * (cx - r[i].x)) * fac);
        out[1] = der((0.5 * (cy - r[i].y)) * fac);
        return;
    }
    double kfx = 2 * k0 / kmagsq, kfy = 2 * k1 / kmagsq;
    dd darea = lv[0].x * lv[1].y - lv[0].y * lv[1].x;
    dd denom = darea * kmagsq;
    cdd cs[2] = {mkc(cd{0, 0}), mkc(cd{0, 0})};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        dd zlk = lv[a].x * k1 - lv[a].y * k0;
        cdd ct1[2] = {cdd{mkd(-kfx), rc[a].x}, cdd{mkd(-kfy), rc[a].y}};
        dd dk = 0.5 * (k0 * lv[a].x + k1 * lv[a].y);
        dd j0 = sj0d(dk), j1 = sj1d(dk);
        dd rt2[2] = {((0.5 * zlk) * j1) * lv[a].x, ((0.5 * zlk) * j1) * lv[a].y};
        cdd cph = dcis(k0 * rc[a].x + k1 * rc[a].y);
        dd zh[2] = {-lv[a].y, lv[a].x};
        cdd cjr[2] = {cdd{mkd(0), r[i].x}, cdd{mkd(0), r[i].y}};
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            cdd ct3 = (zh[d] + zlk * (ct1[d] - cjr[d])) * j0;
            cs[d] += cph * (ct3 - mkc(rt2[d]));
        }
    }
    cdd fac = mkc(ff) / mkc(denom);
    out[0] = der(cs[0] * fac);
    out[1] = der(cs[1] * fac);
}

__global__ void k_bfft_tan(SheetDev sh, const int32_t *bf_face, const int32_t *bf_loc,
                           const BfftJob *jobs, const double *ftan_m, cd *dft) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int job = blockIdx.y;
    if (b >= sh.nbf) return;
    const BfftJob Jb = jobs[job];
    cd acc[2] = {cd{0, 0}, cd{0, 0}};
    for (int t = 0; t < 2; ++t) {
        const int f = bf_face[2 * b + t], i = bf_loc[2 * b + t];
        cd c[2];
        face_ft_tan(sh.fxm + 6 * f, ftan_m + 6 * f, i, Jb, Jb.ff[sh.feci[3 * f + i]], c);
        if (sh.fsg[3 * f + i] > 0) { acc[0] += c[0]; acc[1] += c[1]; }
        else { acc[0] -= c[0]; acc[1] -= c[1]; }
    }
    cd *o = dft + ((int64_t)job * sh.nbf + b) * 2;
    o[0] = acc[0];
    o[1] = acc[1];
}

// g[:, col] = t_col . ft[job(col)]  (observation vectors of the extraction)
__global__ void k_obs(int64_t nbf, const cd *ft, const int32_t *col_job, const double *col_t,
                      cd *g, int64_t ld) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    if (n >= nbf) return;
    const cd *f = ft + ((int64_t)col_job[col] * nbf + n) * 2;
    g[n + ld * col] = col_t[2 * col] * f[0] + col_t[2 * col + 1] * f[1];
}

}  // namespace

// ===================================================================== launchers
void launch_jk_tan_multi(const SheetDev &sh, double ul, const double *us1, const double *us2,
                         double psi1, double psi2, const double *ftan_s, int64_t ftan_stride,
                         int nd, cdd *out, int64_t out_stride, hipStream_t st) {
    if (nd < 1 || nd > 4) throw ArgError("adjoint: 1..4 directions per tangent pass");
    dim3 grid((unsigned)((sh.nufp + 35) / 36));
    hipLaunchKernelGGL(k_jk_tan_multi<4>, grid, dim3(256), 0, st, sh, sh.nufp, ul, us1[0],
                       us1[1], us2[0], us2[1], psi1, psi2, ftan_s, ftan_stride, nd, out,
                       out_stride);
    HIPCHECK(hipGetLastError());
}

void launch_singular_tan(const SheetDev &sh, double ul, const double *ftan_s, dd *rho_r,
                         dd *rinv, hipStream_t st) {
    dim3 grid((unsigned)((sh.nufp + 255) / 256));
    hipLaunchKernelGGL(k_singular_tan, grid, dim3(256), 0, st, sh, sh.nufp, ul, ftan_s, rho_r,
                       rinv);
    HIPCHECK(hipGetLastError());
}

void launch_class_tan(const SheetDev &sh, const FillParams *d_fp, const cdd *jkd,
                      const dd *rho_r, const dd *rinv, const double *ftan_m, const cd *A,
                      int64_t sAo, const cd *X, int ncol, int nobj, int mag, int has_fr,
                      cd *part, cd *out, hipStream_t st) {
    if (nobj < 1 || nobj > 4) throw ArgError("adjoint: 1..4 objectives");
    if (!sh.identity) throw ArgError("adjoint: face-pair classes must be the identity");
    const int64_t nblk = (sh.nufp + 255) / 256;
    if (mag)
        hipLaunchKernelGGL(k_class_tan<true>, dim3((unsigned)nblk), dim3(256), 0, st, sh, d_fp,
                           jkd, rho_r, rinv, ftan_m, A, sAo, X, ncol, nobj, 0, part);
    else
        hipLaunchKernelGGL(k_class_tan<false>, dim3((unsigned)nblk), dim3(256), 0, st, sh, d_fp,
                           jkd, rho_r, rinv, ftan_m, A, sAo, X, ncol, nobj, has_fr, part);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_sum_cd, dim3((unsigned)nobj), dim3(1024), 0, st, part, nblk, nobj, out);
    HIPCHECK(hipGetLastError());
}

void launch_bfft_tan(const SheetDev &sh, const int32_t *bf_face, const int32_t *bf_loc,
                     const BfftJob *d_jobs, int njobs, const double *ftan_m, cd *dft,
                     hipStream_t st) {
    dim3 grid((unsigned)((sh.nbf + 255) / 256), (unsigned)njobs);
    hipLaunchKernelGGL(k_bfft_tan, grid, dim3(256), 0, st, sh, bf_face, bf_loc, d_jobs, ftan_m,
                       dft);
    HIPCHECK(hipGetLastError());
}

void launch_obs(int64_t nbf, int ncol, const cd *ft, const int32_t *d_col_job,
                const double *d_col_t, cd *g, int64_t ld, hipStream_t st) {
    dim3 grid((unsigned)((nbf + 255) / 256), (unsigned)ncol);
    hipLaunchKernelGGL(k_obs, grid, dim3(256), 0, st, nbf, ft, d_col_job, d_col_t, g, ld);
    HIPCHECK(hipGetLastError());
}
