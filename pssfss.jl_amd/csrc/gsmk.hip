// Basis-function Fourier transforms, right-hand sides, Floquet-mode projection
// and the batched GSM cascade for gfx950.
//   rwgbfft!: src/RWG.jl:428-512 (McInturff & Simon closed form, small-k branch)
//   RHS (7.39) and extraction (6.18)/(6.37): src/PSSFSS.jl:431-456, 557-584
//   cascade (Redheffer star product, eq. 3.35): src/GSMs.jl:125-141
#include "internal.hpp"

namespace {

__device__ __forceinline__ double sj0(double x) {
    if (fabs(x) < 1e-3) { double x2 = x * x; return 1 - x2 / 6 * (1 - x2 / 20); }
    return sin(x) / x;
}
__device__ __forceinline__ double sj1(double x) {
    if (fabs(x) < 1e-3) return x * (1.0 / 3 - x * x / 30);
    return (sin(x) / x - cos(x)) / x;
}

// Contribution of one face (local edge i) to the transform of its basis function.
__device__ void face_ft(const double *fxm, int i, const BfftJob &J, cd ff, cd out[2]) {
#pragma clang fp contract(off)
    const double k0 = J.k[0], k1 = J.k[1];
    double r[3][2], lv[3][2], rc[3][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) { r[a][0] = fxm[2 * a]; r[a][1] = fxm[2 * a + 1]; }
    double lmax = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int nx = (a + 1) % 3;
        lv[a][0] = r[nx][0] - r[a][0];
        lv[a][1] = r[nx][1] - r[a][1];
        rc[a][0] = r[a][0] + 0.5 * lv[a][0];
        rc[a][1] = r[a][1] + 0.5 * lv[a][1];
        double ln = sqrt(lv[a][0] * lv[a][0] + lv[a][1] * lv[a][1]);
        if (ln > lmax) lmax = ln;
    }
    double kmagsq = k0 * k0 + k1 * k1;
    double kmag = sqrt(kmagsq);
    if (kmag * lmax < 1e-4) {
        double cx = ((r[0][0] + r[1][0]) + r[2][0]) / 3, cy = ((r[0][1] + r[1][1]) + r[2][1]) / 3;
        cd fac = cisd(k0 * cx + k1 * cy) * ff;
        out[0] = (0.5 * (cx - r[i][0])) * fac;
        out[1] = (0.5 * (cy - r[i][1])) * fac;
        return;
    }
    double kfx = 2 * k0 / kmagsq, kfy = 2 * k1 / kmagsq;
    double darea = lv[0][0] * lv[1][1] - lv[0][1] * lv[1][0];
    double denom = darea * kmagsq;
    double zlk[3], j0v[3], rt2[3][2];
    cd ct1[3][2], cph[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        zlk[a] = lv[a][0] * k1 - lv[a][1] * k0;
        ct1[a][0] = cd{0.0 * rc[a][0], rc[a][0]} - kfx;
        ct1[a][1] = cd{0.0 * rc[a][1], rc[a][1]} - kfy;
        double dk = 0.5 * (k0 * lv[a][0] + k1 * lv[a][1]);
        j0v[a] = sj0(dk);
        double j1 = sj1(dk);
        rt2[a][0] = ((0.5 * zlk[a]) * j1) * lv[a][0];
        rt2[a][1] = ((0.5 * zlk[a]) * j1) * lv[a][1];
        cph[a] = cisd(k0 * rc[a][0] + k1 * rc[a][1]);
    }
    cd cjr[2] = {cd{0.0 * r[i][0], r[i][0]}, cd{0.0 * r[i][1], r[i][1]}};
    cd cs[2] = {cd{0, 0}, cd{0, 0}};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double zh[2] = {-lv[a][1], lv[a][0]};
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            cd ct3 = (zh[d] + zlk[a] * (ct1[a][d] - cjr[d])) * j0v[a];
            cs[d] += cph[a] * (ct3 - rt2[a][d]);
        }
    }
    cd fac = ff / denom;
    out[0] = cs[0] * fac;
    out[1] = cs[1] * fac;
}

// ft[job][bf][2]: one lane per (basis function, job).
__global__ void k_bfft(SheetDev sh, const int32_t *bf_face, const int32_t *bf_loc,
                       const BfftJob *jobs, cd *ft) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int job = blockIdx.y;
    if (b >= sh.nbf) return;
    const BfftJob J = jobs[job];
    cd acc[2] = {cd{0, 0}, cd{0, 0}};
    int f0 = bf_face[2 * b], f1 = bf_face[2 * b + 1];
    int i0 = bf_loc[2 * b], i1 = bf_loc[2 * b + 1];
    // Faces in ascending order, as the reference's face loop accumulates.
    int fa = f0, ia = i0, fb = f1, ib = i1;
    if (f1 < f0 || (f1 == f0 && i1 < i0)) { fa = f1; ia = i1; fb = f0; ib = i0; }
    for (int t = 0; t < 2; ++t) {
        int f = t == 0 ? fa : fb, i = t == 0 ? ia : ib;
        cd c[2];
        face_ft(sh.fxm + 6 * f, i, J, J.ff[sh.feci[3 * f + i]], c);
        if (sh.fsg[3 * f + i] > 0) { acc[0] += c[0]; acc[1] += c[1]; }
        else { acc[0] -= c[0]; acc[1] -= c[1]; }
    }
    cd *o = ft + ((int64_t)job * sh.nbf + b) * 2;
    o[0] = acc[0];
    o[1] = acc[1];
}

// rhs[:, col] = conj(ft[job]) . src  (Julia's dot conjugates its first argument).
__global__ void k_rhs(int64_t nbf, const cd *ft, const int32_t *col_job, const cd *col_src,
                      cd *rhs, int64_t ld) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    if (n >= nbf) return;
    const cd *f = ft + ((int64_t)col_job[col] * nbf + n) * 2;
    cd s0 = col_src[2 * col], s1 = col_src[2 * col + 1];
    rhs[n + ld * col] = conj(f[0]) * s0 + conj(f[1]) * s1;
}

// out[col][job][2] = sum_n X[n, col] * ft[job][n]
__global__ __launch_bounds__(256) void k_project(int64_t nbf, const cd *X, int64_t ldx,
                                                 const cd *ft, cd *out, int njobs) {
    const int col = blockIdx.y, job = blockIdx.x;
    const cd *x = X + ldx * col;
    const cd *f = ft + (int64_t)job * nbf * 2;
    cd a0 = {0, 0}, a1 = {0, 0};
    for (int64_t n = threadIdx.x; n < nbf; n += blockDim.x) {
        cd xv = x[n];
        a0 += xv * f[2 * n];
        a1 += xv * f[2 * n + 1];
    }
    __shared__ cd s0[256], s1[256];
    s0[threadIdx.x] = a0;
    s1[threadIdx.x] = a1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cd *o = out + ((int64_t)col * njobs + job) * 2;
        o[0] = s0[0];
        o[1] = s1[0];
    }
}

// ------------------------------------------------------------------ cascade
constexpr int NC = 34;

// X <- M^{-1} X (n x n system, ncol columns) by Gaussian elimination with partial
// pivoting in LDS; M is destroyed.
__device__ void wg_solve(cd (*M)[NC + 1], cd (*X)[NC + 1], int n, int ncol) {
    __shared__ int piv;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int j = 0; j < n; ++j) {
        if (tid == 0) {
            int p = j;
            double best = cabs(M[j][j]);
            for (int i = j + 1; i < n; ++i) {
                double v = cabs(M[i][j]);
                if (v > best) { best = v; p = i; }
            }
            piv = p;
        }
        __syncthreads();
        int p = piv;
        if (p != j) {
            for (int c = tid; c < n; c += nt) { cd t = M[j][c]; M[j][c] = M[p][c]; M[p][c] = t; }
            for (int c = tid; c < ncol; c += nt) { cd t = X[j][c]; X[j][c] = X[p][c]; X[p][c] = t; }
        }
        __syncthreads();
        cd d = M[j][j];
        for (int t = tid; t < (n - j - 1) * (n + ncol); t += nt) {
            int i = j + 1 + t / (n + ncol), c = t % (n + ncol);
            cd l = M[i][j] / d;
            if (c < n) { if (c > j) M[i][c] -= l * M[j][c]; }
            else X[i][c - n] -= l * X[j][c - n];
        }
        __syncthreads();
        for (int i = j + 1 + tid; i < n; i += nt) M[i][j] = cd{0, 0};
        __syncthreads();
    }
    for (int j = n - 1; j >= 0; --j) {
        for (int c = tid; c < ncol; c += nt) X[j][c] = X[j][c] / M[j][j];
        __syncthreads();
        for (int t = tid; t < j * ncol; t += nt) {
            int i = t / ncol, c = t % ncol;
            X[i][c] -= M[i][j] * X[j][c];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_cascade(int n1, int n2, int n3, const cd *a0,
                                                 const cd *b0, cd *c0) {
    __shared__ cd M[NC][NC + 1];
    __shared__ cd X[NC][NC + 1];
    __shared__ cd T[NC][NC + 1];
    const int64_t e = blockIdx.x;
    const int64_t la = (int64_t)n1 * n1 + 2 * (int64_t)n1 * n2 + (int64_t)n2 * n2;
    const int64_t lb = (int64_t)n2 * n2 + 2 * (int64_t)n2 * n3 + (int64_t)n3 * n3;
    const int64_t lc = (int64_t)n1 * n1 + 2 * (int64_t)n1 * n3 + (int64_t)n3 * n3;
    const cd *a = a0 + e * la, *b = b0 + e * lb;
    cd *c = c0 + e * lc;
    const cd *a11 = a, *a12 = a + n1 * n1, *a21 = a12 + n1 * n2, *a22 = a21 + n2 * n1;
    const cd *b11 = b, *b12 = b + n2 * n2, *b21 = b12 + n2 * n3, *b22 = b21 + n3 * n2;
    cd *c11 = c, *c12 = c + n1 * n1, *c21 = c12 + n1 * n3, *c22 = c21 + n3 * n1;
    const int tid = threadIdx.x, nt = blockDim.x;
    // gprod1 = (I - a22 b11) \ a21
    for (int t = tid; t < n2 * n2; t += nt) {
        int i = t % n2, j = t / n2;
        cd s = {i == j ? 1.0 : 0.0, 0.0};
        for (int l = 0; l < n2; ++l) s -= a22[i + n2 * l] * b11[l + n2 * j];
        M[i][j] = s;
    }
    for (int t = tid; t < n2 * n1; t += nt) X[t % n2][t / n2] = a21[t];
    __syncthreads();
    wg_solve(M, X, n2, n1);
    // s21 = b21 gprod1; s11 = a11 + a12 (b11 gprod1)
    for (int t = tid; t < n3 * n1; t += nt) {
        int i = t % n3, j = t / n3;
        cd s = {0, 0};
        for (int l = 0; l < n2; ++l) s += b21[i + n3 * l] * X[l][j];
        c21[i + n3 * j] = s;
    }
    for (int t = tid; t < n2 * n1; t += nt) {
        int i = t % n2, j = t / n2;
        cd s = {0, 0};
        for (int l = 0; l < n2; ++l) s += b11[i + n2 * l] * X[l][j];
        T[i][j] = s;
    }
    __syncthreads();
    for (int t = tid; t < n1 * n1; t += nt) {
        int i = t % n1, j = t / n1;
        cd s = {0, 0};
        for (int l = 0; l < n2; ++l) s += a12[i + n1 * l] * T[l][j];
        c11[i + n1 * j] = a11[i + n1 * j] + s;
    }
    __syncthreads();
    // gprod2 = (I - b11 a22) \ b12
    for (int t = tid; t < n2 * n2; t += nt) {
        int i = t % n2, j = t / n2;
        cd s = {i == j ? 1.0 : 0.0, 0.0};
        for (int l = 0; l < n2; ++l) s -= b11[i + n2 * l] * a22[l + n2 * j];
        M[i][j] = s;
    }
    for (int t = tid; t < n2 * n3; t += nt) X[t % n2][t / n2] = b12[t];
    __syncthreads();
    wg_solve(M, X, n2, n3);
    for (int t = tid; t < n1 * n3; t += nt) {
        int i = t % n1, j = t / n1;
        cd s = {0, 0};
        for (int l = 0; l < n2; ++l) s += a12[i + n1 * l] * X[l][j];
        c12[i + n1 * j] = s;
    }
    for (int t = tid; t < n2 * n3; t += nt) {
        int i = t % n2, j = t / n2;
        cd s = {0, 0};
        for (int l = 0; l < n2; ++l) s += a22[i + n2 * l] * X[l][j];
        T[i][j] = s;
    }
    __syncthreads();
    for (int t = tid; t < n3 * n3; t += nt) {
        int i = t % n3, j = t / n3;
        cd s = {0, 0};
        for (int l = 0; l < n2; ++l) s += b21[i + n3 * l] * T[l][j];
        c22[i + n3 * j] = b22[i + n3 * j] + s;
    }
}

}  // namespace

void launch_bfft(const SheetDev &sh, const int32_t *bf_face, const int32_t *bf_loc,
                 const BfftJob *d_jobs, int njobs, cd *ft, hipStream_t st) {
    if (njobs <= 0) return;
    dim3 grid((unsigned)((sh.nbf + 255) / 256), (unsigned)njobs);
    hipLaunchKernelGGL(k_bfft, grid, dim3(256), 0, st, sh, bf_face, bf_loc, d_jobs, ft);
    HIPCHECK(hipGetLastError());
}

void launch_rhs(int64_t nbf, int ncol, const cd *ft, const int32_t *d_col_job,
                const cd *d_col_src, cd *rhs, int64_t ld, hipStream_t st) {
    if (ncol <= 0) return;
    dim3 grid((unsigned)((nbf + 255) / 256), (unsigned)ncol);
    hipLaunchKernelGGL(k_rhs, grid, dim3(256), 0, st, nbf, ft, d_col_job, d_col_src, rhs, ld);
    HIPCHECK(hipGetLastError());
}

void launch_project(int64_t nbf, int ncol, int njobs, const cd *X, int64_t ldx, const cd *ft,
                    cd *out, hipStream_t st) {
    if (ncol <= 0 || njobs <= 0) return;
    dim3 grid((unsigned)njobs, (unsigned)ncol);
    hipLaunchKernelGGL(k_project, grid, dim3(256), 0, st, nbf, X, ldx, ft, out, njobs);
    HIPCHECK(hipGetLastError());
}

void cascade_batched_dev(int nb, int n1, int n2, int n3, const cd *a, const cd *b, cd *c,
                         hipStream_t st) {
    if (n1 > NC || n2 > NC || n3 > NC)
        throw ArgError("cascade: at most 34 modes per port");
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_cascade, dim3((unsigned)nb), dim3(256), 0, st, n1, n2, n3, a, b, c);
    HIPCHECK(hipGetLastError());
}
