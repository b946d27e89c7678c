// MoM matrix fill for gfx950 in two phases without atomics.  Phase 1: one lane
// per unique-face-pair class interpolates the two spectral tables at the 7
// quadrature points (zint), combines them with the cached spatial integrals and
// the analytic singular terms, and stores the 9 flag-free local edge-pair terms.
// Phase 2: one lane per Z entry gathers the terms of its <= 4 face pairs.  fillz: src/FillZY.jl:50-245; filly: :273-457;
// zint: src/Zint.jl:218-245; Sigma interpolation: src/PGF.jl:392-443.
#include <algorithm>

#include "internal.hpp"

namespace {

constexpr double TWOPI = 6.283185307179586;
constexpr double FOURPI = 12.566370614359172;
constexpr double PI_ = 3.141592653589793;
constexpr double MU0 = 1.25663706212e-6;
constexpr double C0 = 299792458.0;

__constant__ double QXI[7] = {0.33333333333333330, 0.10128650732345633, 0.79742698535308730,
                              0.10128650732345633, 0.47014206410511505, 0.05971587178976989,
                              0.47014206410511505};
__constant__ double QETA[7] = {0.33333333333333333, 0.79742698535308730, 0.10128650732345633,
                               0.10128650732345633, 0.05971587178976989, 0.47014206410511505,
                               0.47014206410511505};
__constant__ double QW[7] = {0.1125, 0.06296959027241358, 0.06296959027241358,
                             0.06296959027241358, 0.06619707639425308, 0.06619707639425308,
                             0.06619707639425308};

// 6-point interpolation (AMS-55 25.2.67) with the cell and shift decisions taken on
// the same IEEE values as the reference (no contraction).  The Floquet shift
// cis(-(msh psi1 + nsh psi2)) for |msh|, |nsh| <= 1 comes from the point's 9-entry
// table (FillParams::shift9, computed on the host).
__device__ cd sigma_interp(const cd *tab, int M, const double *b1, const double *b2,
                           double psi1, double psi2, const cd *shift9, double rx, double ry) {
#pragma clang fp contract(off)
    double x1o = (b1[0] * rx + b1[1] * ry) / TWOPI;
    if (fabs(x1o) < 1e-8) x1o = 0.0;
    double x2o = (b2[0] * rx + b2[1] * ry) / TWOPI;
    if (fabs(x2o) < 1e-8) x2o = 0.0;
    // Julia's mod(x, 1): x - floor(x) rounds the same real number once, as
    // fmod(x, 1) (exact) followed by +1 for negative x does.
    double x1 = x1o - floor(x1o);
    double x2 = x2o - floor(x2o);
    double msh = rint(x1o - x1), nsh = rint(x2o - x2);
    int m = (int)(M * x1);
    double p = M * x1 - m;
    int n = (int)(M * x2);
    double q = M * x2 - n;
    if (m > M - 1) m = M - 1;
    if (n > M - 1) n = M - 1;
    double p2 = p * p, q2 = q * q, pq = p * q;
    const int64_t E = M + 2;
#define T(a, b) tab[(int64_t)((a) + 1) + E * (int64_t)((b) + 1)]
    cd sm = (0.5 * (q2 - q)) * T(m, n - 1);
    sm = sm + (0.5 * (p2 - p)) * T(m - 1, n);
    sm = sm + (((1 + pq) - p2) - q2) * T(m, n);
    sm = sm + (0.5 * ((p2 - 2 * pq) + p)) * T(m + 1, n);
    sm = sm + (0.5 * ((q2 - 2 * pq) + q)) * T(m, n + 1);
    sm = sm + pq * T(m + 1, n + 1);
#undef T
    if (msh != 0 || nsh != 0) {
        if (fabs(msh) <= 1 && fabs(nsh) <= 1)
            sm = sm * shift9[(int)(3 * (msh + 1) + (nsh + 1))];
        else
            sm = sm * cisd(-(msh * psi1 + nsh * psi2));
    }
    return sm;
}

// Thread t -> class.  With every face pair its own class, a wavefront takes an
// 8 x 8 block of (match, source) faces, so that its interpolation points (match
// centroids minus source points) cluster in the tables; otherwise thread t takes
// class t.  Threads needed: fill_class_threads().
__device__ __forceinline__ bool fill_class_of(const SheetDev &sh, int64_t t, int64_t &iu,
                                              int64_t &ifm, int64_t &ifs) {
    if (sh.identity) {
        const int64_t nf = sh.nface, nb = (nf + 7) >> 3;
        const int64_t tile = t >> 6, w = t & 63;
        const int64_t bm = tile / nb, bs = tile - bm * nb;
        ifm = 8 * bm + (w >> 3);
        ifs = 8 * bs + (w & 7);
        if (bm >= nb || ifm >= nf || ifs >= nf) return false;
        iu = ifm * nf + ifs;
        return true;
    }
    iu = t;
    if (iu >= sh.nufp) return false;
    class_faces(sh, iu, ifm, ifs);
    return true;
}

int64_t fill_class_threads(const SheetDev &sh) {
    if (!sh.identity) return sh.nufp;
    const int64_t nb = (sh.nface + 7) / 8;
    return nb * nb * 64;
}

// Phase 1: one lane per (class, point).  The contribution of local edge pair
// (iml of the match face, isl of the source face) to Z is
//     conj(h_m) h_s w[iml][isl]
// where h = +-ff[cell code] is the Floquet/sign flag of each basis-function half
// and w is flag-free and the same for every member of a translation class.  The
// nine w are written to plane (3 iml + isl) of G, coalesced in the class index.
template <bool MAGNETIC>
__global__ __launch_bounds__(256) void k_fill_classes(SheetDev sh, const FillParams *fps,
                                                      const double *rho_r, const double *rinv_,
                                                      cd *G0, int64_t gstride) {
#pragma clang fp contract(off)
    const FillParams &P = fps[blockIdx.y];
    int64_t iu, ifm, ifs;
    if (!fill_class_of(sh, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, iu, ifm, ifs)) return;
    cd *G = G0 + (int64_t)blockIdx.y * gstride;
    const double EPS0 = 1.0 / (MU0 * (C0 * C0));
    double rs[3][2], rm[3][2];
    {
        const double *a = sh.fxm + 6 * ifs, *b = sh.fxm + 6 * ifm;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            rs[i][0] = a[2 * i]; rs[i][1] = a[2 * i + 1];
            rm[i][0] = b[2 * i]; rm[i][1] = b[2 * i + 1];
        }
    }
    const double rmcx = ((rm[0][0] + rm[1][0]) + rm[2][0]) / 3;
    const double rmcy = ((rm[0][1] + rm[1][1]) + rm[2][1]) / 3;
    cd I1x = {0, 0}, I1e = {0, 0}, I1 = {0, 0}, I2 = {0, 0};
    for (int i = 0; i < 7; ++i) {
        double rtx = (rs[0][0] + (rs[1][0] - rs[0][0]) * QXI[i]) + (rs[2][0] - rs[0][0]) * QETA[i];
        double rty = (rs[0][1] + (rs[1][1] - rs[0][1]) * QXI[i]) + (rs[2][1] - rs[0][1]) * QETA[i];
        double rdx = rmcx - rtx, rdy = rmcy - rty;
        cd s1 = sigma_interp(P.tab1, P.M, P.b1, P.b2, P.psi1, P.psi2, P.shift9, rdx, rdy);
        cd s2 = sigma_interp(P.tab2, P.M, P.b1, P.b2, P.psi1, P.psi2, P.shift9, rdx, rdy);
        cd sw = s1 * QW[i];
        I1x += sw * QXI[i];
        I1e += sw * QETA[i];
        I1 += sw;
        I2 += s2 * QW[i];
    }
    const int64_t nu = sh.nufp;
    const cd J = P.jk[iu], Jx = P.jk[nu + iu], Je = P.jk[2 * nu + iu];
    const cd K = P.jk[3 * nu + iu], Kx = P.jk[4 * nu + iu], Ke = P.jk[5 * nu + iu];
    const double ri = rinv_[iu];
    const double rrx = rho_r[2 * iu], rry = rho_r[2 * iu + 1];
    const cd I1z = (I1 - I1x) - I1e, Jz = (J - Jx) - Je, Kz = (K - Kx) - Ke;
    double rc2[3][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        rc2[i][0] = 0.5 * (rmcx - rm[i][0]);
        rc2[i][1] = 0.5 * (rmcy - rm[i][1]);
    }
    // Edge-independent parts of A_i and Phi_i (src/FillZY.jl:183-200, :385-400).
    const double u = P.u;
    cd phi_core;
    if (!MAGNETIC) phi_core = ((FOURPI * I2) + u * (ri + J)) + (P.d3 / u) * K;
    else phi_core = ((P.I2fact * I2) + u * (ri + J)) + P.KPfact * K;
    const cd iw = MAGNETIC ? cd{0.0, -P.omega} : cd{0.0, P.omega};
    cd Phi, asf;
    if (!MAGNETIC) {
        Phi = (P.psf_base * (1.0 / ((TWOPI * P.omega) * EPS0))) * phi_core;
        asf = (MU0 / FOURPI) * P.mut;
    } else {
        Phi = P.psf_base * phi_core;
        asf = (-EPS0 / PI_) * P.epsb;
    }
#pragma unroll
    for (int isl = 0; isl < 3; ++isl) {
        const double singx = rrx + ((u * ri) * (rmcx - rs[isl][0]));
        const double singy = rry + ((u * ri) * (rmcy - rs[isl][1]));
        const double sing[2] = {singx, singy};
        cd Ai[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            cd I1i = ((rs[0][d] * I1x + rs[1][d] * I1e) + rs[2][d] * I1z) - rs[isl][d] * I1;
            cd Ji = ((rs[0][d] * Jx + rs[1][d] * Je) + rs[2][d] * Jz) - rs[isl][d] * J;
            cd Ki = ((rs[0][d] * Kx + rs[1][d] * Ke) + rs[2][d] * Kz) - rs[isl][d] * K;
            if (!MAGNETIC)
                Ai[d] = asf * ((((FOURPI * I1i) + sing[d]) + u * Ji) + (P.c3 / u) * Ki);
            else
                Ai[d] = asf * ((((P.I1fact * I1i) + sing[d]) + u * Ji) + P.KFfact * Ki);
        }
#pragma unroll
        for (int iml = 0; iml < 3; ++iml) {
            const cd dot = rc2[iml][0] * Ai[0] + rc2[iml][1] * Ai[1];
            G[(int64_t)(3 * iml + isl) * nu + iu] = iw * dot - Phi;
        }
    }
}

// d/dk0 of the class terms of k_fill_classes at fixed phasing and u.  k0 enters
// through the two spectral tables (P.dtab1, P.dtab2 hold their derivatives,
// interpolated with the same weights), the spatial-sum coefficients (electric c3,
// d3; magnetic KFfact, KPfact; derivatives in P.dc3, P.dd3) and omega = c k0; the
// spatial sums, the singular integrals and the Floquet flags do not depend on k0.
// Planes as G.
template <bool MAGNETIC>
__global__ __launch_bounds__(256) void k_fill_classes_dk0(SheetDev sh, const FillParams *fps,
                                                          const double *rho_r,
                                                          const double *rinv_, cd *G) {
    int64_t iu, ifm, ifs;
    if (!fill_class_of(sh, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, iu, ifm, ifs)) return;
    const FillParams &P = *fps;
    const double EPS0 = 1.0 / (MU0 * (C0 * C0));
    double rs[3][2], rm[3][2];
    {
        const double *a = sh.fxm + 6 * ifs, *b = sh.fxm + 6 * ifm;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            rs[i][0] = a[2 * i]; rs[i][1] = a[2 * i + 1];
            rm[i][0] = b[2 * i]; rm[i][1] = b[2 * i + 1];
        }
    }
    const double rmcx = ((rm[0][0] + rm[1][0]) + rm[2][0]) / 3;
    const double rmcy = ((rm[0][1] + rm[1][1]) + rm[2][1]) / 3;
    cd I1x = {0, 0}, I1e = {0, 0}, I1 = {0, 0}, I2 = {0, 0};
    cd D1x = {0, 0}, D1e = {0, 0}, D1 = {0, 0}, D2 = {0, 0};
    for (int i = 0; i < 7; ++i) {
        double rtx = (rs[0][0] + (rs[1][0] - rs[0][0]) * QXI[i]) + (rs[2][0] - rs[0][0]) * QETA[i];
        double rty = (rs[0][1] + (rs[1][1] - rs[0][1]) * QXI[i]) + (rs[2][1] - rs[0][1]) * QETA[i];
        double rdx = rmcx - rtx, rdy = rmcy - rty;
        const cd s1 =
            sigma_interp(P.tab1, P.M, P.b1, P.b2, P.psi1, P.psi2, P.shift9, rdx, rdy) * QW[i];
        const cd s2 = sigma_interp(P.tab2, P.M, P.b1, P.b2, P.psi1, P.psi2, P.shift9, rdx, rdy);
        const cd t1 =
            sigma_interp(P.dtab1, P.M, P.b1, P.b2, P.psi1, P.psi2, P.shift9, rdx, rdy) * QW[i];
        const cd t2 = sigma_interp(P.dtab2, P.M, P.b1, P.b2, P.psi1, P.psi2, P.shift9, rdx, rdy);
        I1x += s1 * QXI[i];
        I1e += s1 * QETA[i];
        I1 += s1;
        I2 += s2 * QW[i];
        D1x += t1 * QXI[i];
        D1e += t1 * QETA[i];
        D1 += t1;
        D2 += t2 * QW[i];
    }
    const int64_t nu = sh.nufp;
    const cd J = P.jk[iu], Jx = P.jk[nu + iu], Je = P.jk[2 * nu + iu];
    const cd K = P.jk[3 * nu + iu], Kx = P.jk[4 * nu + iu], Ke = P.jk[5 * nu + iu];
    const double ri = rinv_[iu];
    const double rrx = rho_r[2 * iu], rry = rho_r[2 * iu + 1];
    const cd I1z = (I1 - I1x) - I1e, D1z = (D1 - D1x) - D1e;
    const cd Jz = (J - Jx) - Je, Kz = (K - Kx) - Ke;
    double rc2[3][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        rc2[i][0] = 0.5 * (rmcx - rm[i][0]);
        rc2[i][1] = 0.5 * (rmcy - rm[i][1]);
    }
    const double u = P.u, w = P.omega, dw = C0;
    cd phi_core, dphi_core, dPhi, asf;
    if (!MAGNETIC) {
        phi_core = ((FOURPI * I2) + u * (ri + J)) + (P.d3 / u) * K;
        dphi_core = (FOURPI * D2) + (P.dd3 / u) * K;
        // Phi = pf phi_core / omega
        const cd pf = P.psf_base * (1.0 / (TWOPI * EPS0));
        dPhi = pf * ((1.0 / w) * dphi_core - (dw / (w * w)) * phi_core);
        asf = (MU0 / FOURPI) * P.mut;
    } else {
        phi_core = ((P.I2fact * I2) + u * (ri + J)) + P.KPfact * K;
        dphi_core = (P.I2fact * D2) + P.dd3 * K;
        // Phi = psf_base phi_core, psf_base proportional to 1 / omega
        dPhi = P.psf_base * (dphi_core - (dw / w) * phi_core);
        asf = (-EPS0 / PI_) * P.epsb;
    }
    const double sgn = MAGNETIC ? -1.0 : 1.0;
    const cd iw = cd{0.0, sgn * w}, idw = cd{0.0, sgn * dw};
#pragma unroll
    for (int isl = 0; isl < 3; ++isl) {
        const double singx = rrx + ((u * ri) * (rmcx - rs[isl][0]));
        const double singy = rry + ((u * ri) * (rmcy - rs[isl][1]));
        const double sing[2] = {singx, singy};
        cd Ai[2], dAi[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            cd I1i = ((rs[0][d] * I1x + rs[1][d] * I1e) + rs[2][d] * I1z) - rs[isl][d] * I1;
            cd D1i = ((rs[0][d] * D1x + rs[1][d] * D1e) + rs[2][d] * D1z) - rs[isl][d] * D1;
            cd Ji = ((rs[0][d] * Jx + rs[1][d] * Je) + rs[2][d] * Jz) - rs[isl][d] * J;
            cd Ki = ((rs[0][d] * Kx + rs[1][d] * Ke) + rs[2][d] * Kz) - rs[isl][d] * K;
            if (!MAGNETIC) {
                Ai[d] = asf * ((((FOURPI * I1i) + sing[d]) + u * Ji) + (P.c3 / u) * Ki);
                dAi[d] = asf * ((FOURPI * D1i) + (P.dc3 / u) * Ki);
            } else {
                Ai[d] = asf * ((((P.I1fact * I1i) + sing[d]) + u * Ji) + P.KFfact * Ki);
                dAi[d] = asf * ((P.I1fact * D1i) + P.dc3 * Ki);
            }
        }
#pragma unroll
        for (int iml = 0; iml < 3; ++iml) {
            const cd dot = rc2[iml][0] * Ai[0] + rc2[iml][1] * Ai[1];
            const cd ddot = rc2[iml][0] * dAi[0] + rc2[iml][1] * dAi[1];
            G[(int64_t)(3 * iml + isl) * nu + iu] = (idw * dot + iw * ddot) - dPhi;
        }
    }
}

// Phase 2: one lane per Z entry (point, row m, column s).  Z[m][s] sums the
// contributions of the (<= 4) face pairs formed by the two halves of m and the
// two halves of s (the scatter of src/FillZY.jl:201-240 turned into a gather),
// plus the surface-resistance term of self-triangle pairs (src/FillZY.jl:221-231).
template <bool MAGNETIC>
__global__ __launch_bounds__(256) void k_fill_gather(SheetDev sh, const FillParams *fps,
                                                     const int32_t *bf_face,
                                                     const int32_t *bf_loc,
                                                     const int32_t *fpcls, const cd *G0,
                                                     int64_t gstride, int has_fr) {
#pragma clang fp contract(off)
    const int64_t nbf = sh.nbf;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nbf * nbf) return;
    const FillParams &P = fps[blockIdx.y];
    const cd *G = G0 + (int64_t)blockIdx.y * gstride;
    const int64_t m = e / nbf, s = e - m * nbf;
    const int64_t nu = sh.nufp, nf = sh.nface;
    cd acc = {0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int fm = bf_face[2 * m + a], lm = bf_loc[2 * m + a];
        const cd ffm = conj(P.ff[sh.feci[3 * fm + lm]]);
        const cd hm = sh.fsg[3 * fm + lm] > 0 ? ffm : -ffm;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int fs = bf_face[2 * s + b], ls = bf_loc[2 * s + b];
            const cd ffs = P.ff[sh.feci[3 * fs + ls]];
            const cd hs = sh.fsg[3 * fs + ls] > 0 ? ffs : -ffs;
            const int64_t cls = sh.identity ? (int64_t)fm * nf + fs
                                            : (int64_t)fpcls[(int64_t)fm + nf * fs];
            cd zc = hm * (hs * G[(int64_t)(3 * lm + ls) * nu + cls]);
            if (!MAGNETIC && has_fr && fm == fs && sh.fr[fm] != 0) {
                const double *v = sh.fxm + 6 * fs;
                const double r32x = v[4] - v[2], r32y = v[5] - v[3];
                const double r12x = v[0] - v[2], r12y = v[1] - v[3];
                const double r13x = v[0] - v[4], r13y = v[1] - v[5];
                const double area48 = 48 * fabs(0.5 * (r32x * r12y - r32y * r12x));
                double l2[3];
                l2[0] = r32x * r32x + r32y * r32y;
                l2[1] = r13x * r13x + r13y * r13y;
                l2[2] = r12x * r12x + r12y * r12y;
                const double q = sh.fr[fs] / area48;
                if (lm == ls) {
                    const int nx = ls == 2 ? 0 : ls + 1, pv = ls == 0 ? 2 : ls - 1;
                    zc = zc + q * (3 * (l2[nx] + l2[pv]) - l2[ls]);
                } else {
                    const int t3 = 3 - lm - ls;
                    zc = zc + ((q * hs) * hm) * ((l2[ls] + l2[lm]) - 3 * l2[t3]);
                }
            }
            acc += zc;
        }
    }
    P.Z[m * (P.ldz ? P.ldz : nbf) + s] = acc;     // row-major Z[m][s]
}

__global__ void k_pad_identity(cd *Z0, int64_t n, int64_t ld, int64_t sZ) {
    cd *Z = Z0 + (int64_t)blockIdx.y * sZ;
    const int64_t pad = ld - n;
    // rows [0, n): columns [n, ld); rows [n, ld): every column
    const int64_t cnt = n * pad + pad * ld;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cnt;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t r, c;
        if (t < n * pad) {
            r = t / pad;
            c = n + t % pad;
        } else {
            r = n + (t - n * pad) / ld;
            c = (t - n * pad) % ld;
        }
        Z[r * ld + c] = r == c ? cd{1.0, 0.0} : cd{0.0, 0.0};
    }
}

__global__ void k_zero(cd *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = cd{0.0, 0.0};
}

}  // namespace

void launch_zero(cd *p, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
    HIPCHECK(hipGetLastError());
}

void launch_pad_identity(cd *Z, int64_t n, int64_t ld, int64_t sZ, int npts, hipStream_t st) {
    if (n >= ld || npts <= 0) return;
    const int64_t cnt = (ld - n) * (n + ld);
    const int64_t blocks = std::min<int64_t>((cnt + 255) / 256, 4096);
    hipLaunchKernelGGL(k_pad_identity, dim3((unsigned)blocks, (unsigned)npts), dim3(256), 0, st, Z,
                       n, ld, sZ);
    HIPCHECK(hipGetLastError());
}

void launch_fill_classes(const SheetDev &sh, const FillParams *d_fp, int npts,
                         const double *rho_r, const double *rinv, int magnetic, cd *G,
                         int64_t gstride, hipStream_t st) {
    dim3 grid((unsigned)((fill_class_threads(sh) + 255) / 256), (unsigned)npts);
    if (magnetic)
        hipLaunchKernelGGL(k_fill_classes<true>, grid, dim3(256), 0, st, sh, d_fp, rho_r, rinv,
                           G, gstride);
    else
        hipLaunchKernelGGL(k_fill_classes<false>, grid, dim3(256), 0, st, sh, d_fp, rho_r, rinv,
                           G, gstride);
    HIPCHECK(hipGetLastError());
}

void launch_fill_classes_dk0(const SheetDev &sh, const FillParams *d_fp, const double *rho_r,
                             const double *rinv, int magnetic, cd *G, hipStream_t st) {
    dim3 grid((unsigned)((fill_class_threads(sh) + 255) / 256));
    if (magnetic)
        hipLaunchKernelGGL(k_fill_classes_dk0<true>, grid, dim3(256), 0, st, sh, d_fp, rho_r,
                           rinv, G);
    else
        hipLaunchKernelGGL(k_fill_classes_dk0<false>, grid, dim3(256), 0, st, sh, d_fp, rho_r,
                           rinv, G);
    HIPCHECK(hipGetLastError());
}

void launch_fill_gather(const SheetDev &sh, const FillParams *d_fp, int npts,
                        const int32_t *bf_face, const int32_t *bf_loc, const int32_t *fpcls,
                        const cd *G, int64_t gstride, int has_fr, int magnetic, hipStream_t st) {
    const int64_t n2 = sh.nbf * sh.nbf;
    dim3 grid((unsigned)((n2 + 255) / 256), (unsigned)npts);
    if (magnetic)
        hipLaunchKernelGGL(k_fill_gather<true>, grid, dim3(256), 0, st, sh, d_fp, bf_face, bf_loc,
                           fpcls, G, gstride, has_fr);
    else
        hipLaunchKernelGGL(k_fill_gather<false>, grid, dim3(256), 0, st, sh, d_fp, bf_face,
                           bf_loc, fpcls, G, gstride, has_fr);
    HIPCHECK(hipGetLastError());
}
