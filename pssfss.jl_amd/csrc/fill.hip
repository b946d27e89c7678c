// MoM matrix fill for gfx950: one lane per unique-face-pair class.  Each lane
// interpolates the two spectral tables at the 7 quadrature points (zint), combines
// them with the cached spatial integrals and the analytic singular terms, and
// adds the <= 9 edge-pair contributions of every member face pair of the class
// into Z with FP64 atomics.  fillz: src/FillZY.jl:50-245; filly: :273-457;
// zint: src/Zint.jl:218-245; Sigma interpolation: src/PGF.jl:392-443.
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
// the same IEEE values as the reference (no contraction).
__device__ cd sigma_interp(const cd *tab, int M, const double *b1, const double *b2,
                           double psi1, double psi2, double rx, double ry) {
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
    if (msh != 0 || nsh != 0) sm = sm * cisd(-(msh * psi1 + nsh * psi2));
    return sm;
}

__device__ __forceinline__ void atomic_add_cd(cd *p, cd v) {
    unsafeAtomicAdd(&p->x, v.x);
    unsafeAtomicAdd(&p->y, v.y);
}

template <bool MAGNETIC>
__global__ __launch_bounds__(256) void k_fill(SheetDev sh, const FillParams *fps,
                                              const double *rho_r, const double *rinv_) {
#pragma clang fp contract(off)
    const int64_t iu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (iu >= sh.nufp) return;
    const FillParams &P = fps[blockIdx.y];
    const double EPS0 = 1.0 / (MU0 * (C0 * C0));
    int64_t ifm, ifs;
    class_faces(sh, iu, ifm, ifs);
    double rs[3][2], rm[3][2];
    {
        const double *a = sh.fxm + 6 * ifs, *b = sh.fxm + 6 * ifm;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            rs[i][0] = a[2 * i]; rs[i][1] = a[2 * i + 1];
            rm[i][0] = b[2 * i]; rm[i][1] = b[2 * i + 1];
        }
    }
    const double rs32x = rs[2][0] - rs[1][0], rs32y = rs[2][1] - rs[1][1];
    const double rs12x = rs[0][0] - rs[1][0], rs12y = rs[0][1] - rs[1][1];
    const double area = 0.5 * (rs32x * rs12y - rs32y * rs12x);
    const double area48 = 48 * fabs(area);
    const bool self_tri = (ifm == ifs);
    double ls[3] = {0, 0, 0};
    if (!MAGNETIC && self_tri) {
        double r13x = rs[0][0] - rs[2][0], r13y = rs[0][1] - rs[2][1];
        ls[0] = sqrt(rs32x * rs32x + rs32y * rs32y);
        ls[1] = sqrt(r13x * r13x + r13y * r13y);
        ls[2] = sqrt(rs12x * rs12x + rs12y * rs12y);
    }
    const double rmcx = ((rm[0][0] + rm[1][0]) + rm[2][0]) / 3;
    const double rmcy = ((rm[0][1] + rm[1][1]) + rm[2][1]) / 3;
    cd I1x = {0, 0}, I1e = {0, 0}, I1 = {0, 0}, I2 = {0, 0};
    for (int i = 0; i < 7; ++i) {
        double rtx = (rs[0][0] + (rs[1][0] - rs[0][0]) * QXI[i]) + (rs[2][0] - rs[0][0]) * QETA[i];
        double rty = (rs[0][1] + (rs[1][1] - rs[0][1]) * QXI[i]) + (rs[2][1] - rs[0][1]) * QETA[i];
        double rdx = rmcx - rtx, rdy = rmcy - rty;
        cd s1 = sigma_interp(P.tab1, P.M, P.b1, P.b2, P.psi1, P.psi2, rdx, rdy);
        cd s2 = sigma_interp(P.tab2, P.M, P.b1, P.b2, P.psi1, P.psi2, rdx, rdy);
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
    // Edge-independent parts of A_i and Phi_i.
    const double u = P.u;
    cd phi_core;
    if (!MAGNETIC) phi_core = ((FOURPI * I2) + u * (ri + J)) + (P.d3 / u) * K;
    else phi_core = ((P.I2fact * I2) + u * (ri + J)) + P.KPfact * K;
    const cd iw = MAGNETIC ? cd{0.0, -P.omega} : cd{0.0, P.omega};
    const cd psf_scale = P.psf_base * (1.0 / ((TWOPI * P.omega) * EPS0));
    const int64_t nmem = sh.identity ? 1 : sh.uoff[iu + 1] - sh.uoff[iu];
    const int64_t nbf = sh.nbf;
    for (int64_t im_ = 0; im_ < nmem; ++im_) {
        int64_t fm = ifm, fs = ifs;
        if (!sh.identity) {
            const int64_t fp = sh.umem[sh.uoff[iu] + im_];
            fm = fp % sh.nface;
            fs = fp / sh.nface;
        }
        for (int isl = 0; isl < 3; ++isl) {
            const int sbf = sh.fbf[3 * fs + isl];
            if (sbf < 0) continue;
            const cd ffs = P.ff[sh.feci[3 * fs + isl]];
            const cd sflag = sh.fsg[3 * fs + isl] > 0 ? ffs : -ffs;
            const double singx = rrx + ((u * ri) * (rmcx - rs[isl][0]));
            const double singy = rry + ((u * ri) * (rmcy - rs[isl][1]));
            cd I1i[2], Ji[2], Ki[2];
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                I1i[d] = ((rs[0][d] * I1x + rs[1][d] * I1e) + rs[2][d] * I1z) - rs[isl][d] * I1;
                Ji[d] = ((rs[0][d] * Jx + rs[1][d] * Je) + rs[2][d] * Jz) - rs[isl][d] * J;
                Ki[d] = ((rs[0][d] * Kx + rs[1][d] * Ke) + rs[2][d] * Kz) - rs[isl][d] * K;
            }
            const double sing[2] = {singx, singy};
            cd Ai[2], Phi;
            if (!MAGNETIC) {
                cd asf = ((MU0 / FOURPI) * sflag) * P.mut;
                cd psf = psf_scale * sflag;
#pragma unroll
                for (int d = 0; d < 2; ++d)
                    Ai[d] = asf * ((((FOURPI * I1i[d]) + sing[d]) + u * Ji[d]) + (P.c3 / u) * Ki[d]);
                Phi = psf * phi_core;
            } else {
                cd fsf = ((-EPS0 / PI_) * P.epsb) * sflag;
                cd ysf = P.psf_base * sflag;
#pragma unroll
                for (int d = 0; d < 2; ++d)
                    Ai[d] = fsf * ((((P.I1fact * I1i[d]) + sing[d]) + u * Ji[d]) + P.KFfact * Ki[d]);
                Phi = ysf * phi_core;
            }
            const int se = sh.fedge[3 * fs + isl];
            for (int iml = 0; iml < 3; ++iml) {
                const int mbf = sh.fbf[3 * fm + iml];
                if (mbf < 0) continue;
                const cd ffm = conj(P.ff[sh.feci[3 * fm + iml]]);
                const cd mflag = sh.fsg[3 * fm + iml] > 0 ? ffm : -ffm;
                const cd dot = rc2[iml][0] * Ai[0] + rc2[iml][1] * Ai[1];
                cd zc = mflag * (iw * dot - Phi);
                if (!MAGNETIC && self_tri && sh.fr[fm] != 0) {
                    static const int nxt[3] = {1, 2, 0}, prv[3] = {2, 0, 1};
                    static const int third[3][3] = {{-1, 2, 1}, {2, -1, 0}, {1, 0, -1}};
                    if (sh.fedge[3 * fm + iml] == se) {
                        double zl = sh.fr[fs] / area48 *
                                    (3 * (ls[nxt[isl]] * ls[nxt[isl]] + ls[prv[isl]] * ls[prv[isl]]) -
                                     ls[isl] * ls[isl]);
                        zc = zc + zl;
                    } else {
                        int t3 = third[isl][iml];
                        cd zlc = (((sh.fr[fs] / area48) * sflag) * mflag) *
                                 ((ls[isl] * ls[isl] + ls[iml] * ls[iml]) - 3 * ls[t3] * ls[t3]);
                        zc = zc + zlc;
                    }
                }
                atomic_add_cd(P.Z + (int64_t)mbf * nbf + (int64_t)sbf, zc);  // row-major Z[m][s]
            }
        }
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

void launch_fill(const SheetDev &sh, const FillParams *d_fp, int npts, const double *rho_r,
                 const double *rinv, int magnetic, hipStream_t st) {
    dim3 grid((unsigned)((sh.nufp + 255) / 256), (unsigned)npts);
    if (magnetic)
        hipLaunchKernelGGL(k_fill<true>, grid, dim3(256), 0, st, sh, d_fp, rho_r, rinv);
    else
        hipLaunchKernelGGL(k_fill<false>, grid, dim3(256), 0, st, sh, d_fp, rho_r, rinv);
    HIPCHECK(hipGetLastError());
}
