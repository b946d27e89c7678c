// Dense complex linear algebra for gfx950: batched GEMM, blocked LU with partial
// pivoting, triangular solves and the FP64 refinement helpers.
//
// Matrices of the MoM path are stored ROW-major on the device (the fill writes
// Z[m][s] at m*n + s): row interchanges, panel rows and the U12 block rows are
// then contiguous.  The factorisation is LAPACK's right-looking zgetrf with the
// pivot of largest |re|+|im| (izamax), applied to whole rows, so P*A = L*U with
// the usual packed L\U storage -- the lu! / ldiv! pair of src/PSSFSS.jl:401,439.
#include "internal.hpp"
#include "mfma.hpp"

namespace {

constexpr int NB = 64;  // panel width
// Panels with more rows than this are factored by 8-column leaves (panel_recursive).
constexpr int64_t RECURSIVE_PANEL_ROWS = 3072;

// ------------------------------------------------------------------ GEMM (NN)
// C = alpha*A*B + beta*C, column-major complex, batched by blockIdx.z, on the
// matrix cores.  A complex product is a real product with K doubled: with
// (re, im) interleaved along k, Re C = [Ar, -Ai] [Br; Bi] and
// Im C = [Ai, Ar] [Br; Bi], so B feeds the 16x16x4 MFMA in its memory layout
// and each 4-deep real step covers two complex k.  64x64 tile per workgroup,
// 32x32 per wavefront, BK = 16 complex k per LDS stage, double-buffered.
template <class C>
__global__ __launch_bounds__(256) void k_gemm_mfma(int m, int n, int k, C alpha,
                                                   const C *__restrict__ A, int64_t lda,
                                                   int64_t sA, const C *__restrict__ B,
                                                   int64_t ldb, int64_t sB, C beta, C *Cm,
                                                   int64_t ldc, int64_t sC) {
    typedef typename CTraits<C>::R R;
    typedef MfmaT<R> MT;
    typedef typename MT::V V;
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ C As[2][BK][BM];
    __shared__ C Bs[2][BK][BN];
    const int64_t bz = blockIdx.z;
    A += bz * sA;
    B += bz * sB;
    Cm += bz * sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
    const C zero = CTraits<C>::zero();
    V acc[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < 2; ++c) acc[a][b][c] = V{0, 0, 0, 0};
    C ra[4], rb[4];
#define GEMM_LOAD(K0)                                                                          \
    _Pragma("unroll") for (int q = 0; q < 4; ++q) {                                           \
        const int t = tid + 256 * q;                                                           \
        const int i = t % BM, kk = t / BM;                                                     \
        ra[q] = zero;                                                                          \
        if (m0 + i < m && (K0) + kk < k) ra[q] = A[(m0 + i) + lda * ((K0) + kk)];              \
        const int kb = t % BK, j = t / BK;                                                     \
        rb[q] = zero;                                                                          \
        if ((K0) + kb < k && n0 + j < n) rb[q] = B[((K0) + kb) + ldb * (n0 + j)];              \
    }
#define GEMM_STORE(BUF)                                                                        \
    _Pragma("unroll") for (int q = 0; q < 4; ++q) {                                           \
        const int t = tid + 256 * q;                                                           \
        As[BUF][t / BM][t % BM] = ra[q];                                                       \
        Bs[BUF][t % BK][t / BK] = rb[q];                                                       \
    }
    const int nk = (k + BK - 1) / BK;
    GEMM_LOAD(0)
    GEMM_STORE(0)
    __syncthreads();
    const int c = (lane >> 4) & 1;        // re / im half of the doubled k
    const int kh = lane >> 5;             // which of the two complex k of a step
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) { GEMM_LOAD((kt + 1) * BK) }
#pragma unroll
        for (int s2 = 0; s2 < BK / 2; ++s2) {
            const int kc = 2 * s2 + kh;
            R are[2], aim[2], bop[2];
#pragma unroll
            for (int ti = 0; ti < 2; ++ti) {
                const C a = As[buf][kc][32 * wm + 16 * ti + (lane & 15)];
                are[ti] = c == 0 ? a.x : -a.y;
                aim[ti] = c == 0 ? a.y : a.x;
            }
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) {
                const C b = Bs[buf][kc][32 * wn + 16 * tj + (lane & 15)];
                bop[tj] = c == 0 ? b.x : b.y;
            }
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj) {
                    acc[ti][tj][0] = MT::mma(are[ti], bop[tj], acc[ti][tj][0]);
                    acc[ti][tj][1] = MT::mma(aim[ti], bop[tj], acc[ti][tj][1]);
                }
        }
        if (kt + 1 < nk) { GEMM_STORE(buf ^ 1) }
        __syncthreads();
    }
#undef GEMM_LOAD
#undef GEMM_STORE
    const bool use_beta = !(beta.x == 0 && beta.y == 0);
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t i = m0 + 32 * wm + 16 * ti + MT::row(lane, e);
                const int64_t j = n0 + 32 * wn + 16 * tj + (lane & 15);
                if (i < m && j < n) {
                    C v;
                    v.x = acc[ti][tj][0][e];
                    v.y = acc[ti][tj][1][e];
                    C *p = Cm + i + ldc * j;
                    C o = alpha * v;
                    if (use_beta) o += beta * (*p);
                    *p = o;
                }
            }
}

// ------------------------------------------------------------------- LU panel
// Factor columns [k, k+kb) of rows [k, n) of a row-major matrix with partial
// pivoting; swaps are applied inside the panel only (k_rowswap does the rest).
// The pivot search of column j+1 is fused into the rank-1 update pass of column j.
template <class C>
__global__ __launch_bounds__(1024) void k_panel(int64_t n, int64_t k, int kb, C *A0, int64_t sA,
                                                int32_t *ipiv0, int64_t sP) {
    typedef typename CTraits<C>::R R;
    C *A = A0 + (int64_t)blockIdx.x * sA;
    int32_t *ipiv = ipiv0 + (int64_t)blockIdx.x * sP;
    const int64_t lda = n;
    __shared__ C prow[NB];
    __shared__ R rv[16];
    __shared__ int ri[16];
    __shared__ int pivs;
    const int tid = threadIdx.x, nt = blockDim.x, nw = nt >> 6;
    R best = (R)-1;
    int bi = 0x7fffffff;
    for (int64_t i = k + tid; i < n; i += nt) {
        R v = cabs1(A[i * lda + k]);
        if (v > best) { best = v; bi = (int)i; }
    }
    for (int j = 0; j < kb; ++j) {
        const int64_t col = k + j;
        for (int off = 32; off > 0; off >>= 1) {
            R ov = __shfl_xor(best, off, 64);
            int oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        if ((tid & 63) == 0) { rv[tid >> 6] = best; ri[tid >> 6] = bi; }
        __syncthreads();
        if (tid == 0) {
            R b = rv[0];
            int bx = ri[0];
            for (int w = 1; w < nw; ++w)
                if (rv[w] > b || (rv[w] == b && ri[w] < bx)) { b = rv[w]; bx = ri[w]; }
            pivs = bx;
            ipiv[col] = bx;
        }
        __syncthreads();
        const int64_t p = pivs;
        if (p != col) {
            for (int c = tid; c < kb; c += nt) {
                C t = A[col * lda + k + c];
                A[col * lda + k + c] = A[p * lda + k + c];
                A[p * lda + k + c] = t;
            }
        }
        __syncthreads();
        for (int c = tid; c < kb; c += nt) prow[c] = A[col * lda + k + c];
        __syncthreads();
        const C piv = prow[j];
        const bool nz = !(piv.x == 0 && piv.y == 0);
        const C rp = nz ? CTraits<C>::one() / piv : CTraits<C>::zero();
        best = (R)-1;
        bi = 0x7fffffff;
        const bool next = j + 1 < kb;
        for (int64_t i = col + 1 + tid; i < n; i += nt) {
            C *row = A + i * lda + k;
            C l = nz ? row[j] * rp : row[j];
            row[j] = l;
            C ml = -l;
            for (int c = j + 1; c < kb; ++c) cfma(row[c], ml, prow[c]);
            if (next) {
                R v = cabs1(row[j + 1]);
                if (v > best) { best = v; bi = (int)i; }
            }
        }
        __syncthreads();
    }
}

// Panel factorisation with the rows resident in registers, W columns at a time.
// Interchanges inside the panel are logical: perm[i] (LDS) is the physical row
// (relative to k) holding logical row i, so a pivot swap exchanges two perm
// entries and two register rows and moves no memory.  After each W-column
// sub-block is factored, its U rows for the remaining panel columns are solved
// against the W x W unit-lower block, and the remaining panel columns of the
// rows below take the rank-W update.  At the end the panel rows are written back
// in logical order, so the matrix holds exactly what LAPACK's sequential
// interchanges (recorded in ipiv) produce.  Rows per thread: R (m <= 1024 R).
template <class C, int R, int W>
__global__ __launch_bounds__(1024) void k_panel_reg(int64_t n, int64_t k, int kb, C *A0,
                                                    int64_t sA, int32_t *ipiv0, int64_t sP) {
    typedef typename CTraits<C>::R Rl;
    constexpr int NT = 1024;
    C *A = A0 + (int64_t)blockIdx.x * sA;
    int32_t *ipiv = ipiv0 + (int64_t)blockIdx.x * sP;
    const int m = (int)(n - k);
    __shared__ int perm[R * NT];
    __shared__ C prow[W];
    __shared__ C Ld[W][W];
    __shared__ C Us[W][NB];
    __shared__ C xch[2][W];
    __shared__ Rl rv[16];
    __shared__ int ri[16];
    __shared__ int pivs;
    const int tid = threadIdx.x;
    const C zero = CTraits<C>::zero();
    for (int i = tid; i < m; i += NT) perm[i] = i;
    __syncthreads();
    C x[R][W];
    for (int c0 = 0; c0 < kb; c0 += W) {
        const int wb = kb - c0 < W ? kb - c0 : W;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tid + NT * r;
            const bool live = i >= c0 && i < m;
            const C *row = A + (k + (live ? perm[i] : 0)) * n + k + c0;
#pragma unroll
            for (int c = 0; c < W; ++c) x[r][c] = (live && c < wb) ? row[c] : zero;
        }
#pragma unroll
        for (int j = 0; j < W; ++j) {
            if (j >= wb) break;
            const int col = c0 + j;
            Rl best = (Rl)-1;
            int bi = 0x7fffffff;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = tid + NT * r;
                if (i >= col && i < m) {
                    Rl v = cabs1(x[r][j]);
                    if (v > best) { best = v; bi = i; }
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                Rl ov = __shfl_xor(best, off, 64);
                int oi = __shfl_xor(bi, off, 64);
                if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
            }
            if ((tid & 63) == 0) { rv[tid >> 6] = best; ri[tid >> 6] = bi; }
            __syncthreads();
            if (tid == 0) {
                Rl b = rv[0];
                int bx = ri[0];
                for (int w = 1; w < NT / 64; ++w)
                    if (rv[w] > b || (rv[w] == b && ri[w] < bx)) { b = rv[w]; bx = ri[w]; }
                pivs = bx;
                ipiv[k + col] = (int32_t)(k + bx);
            }
            __syncthreads();
            const int p = pivs;
            const int oc = col % NT, sc = col / NT, op = p % NT, sp = p / NT;
            if (p != col) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (tid == op && r == sp)
#pragma unroll
                        for (int c = 0; c < W; ++c) xch[0][c] = x[r][c];
                    if (tid == oc && r == sc)
#pragma unroll
                        for (int c = 0; c < W; ++c) xch[1][c] = x[r][c];
                }
                if (tid == 0) { int t = perm[col]; perm[col] = perm[p]; perm[p] = t; }
                __syncthreads();
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (tid == oc && r == sc)
#pragma unroll
                        for (int c = 0; c < W; ++c) x[r][c] = xch[0][c];
                    if (tid == op && r == sp)
#pragma unroll
                        for (int c = 0; c < W; ++c) x[r][c] = xch[1][c];
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (tid == oc && r == sc)
#pragma unroll
                    for (int c = 0; c < W; ++c) prow[c] = x[r][c];
            __syncthreads();
            const C piv = prow[j];
            const bool nz = !(piv.x == 0 && piv.y == 0);
            const C rp = nz ? CTraits<C>::one() / piv : zero;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = tid + NT * r;
                if (i > col && i < m) {
                    C l = nz ? x[r][j] * rp : x[r][j];
                    x[r][j] = l;
                    C ml = -l;
#pragma unroll
                    for (int c = j + 1; c < W; ++c) cfma(x[r][c], ml, prow[c]);
                }
            }
        }
        // write the sub-block back; the diagonal block goes to LDS as well
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tid + NT * r;
            if (i >= c0 && i < m) {
                C *row = A + (k + perm[i]) * n + k + c0;
#pragma unroll
                for (int c = 0; c < W; ++c)
                    if (c < wb) row[c] = x[r][c];
                if (i < c0 + wb)
#pragma unroll
                    for (int c = 0; c < W; ++c) Ld[i - c0][c] = x[r][c];
            }
        }
        __syncthreads();
        const int cr0 = c0 + wb;
        if (cr0 < kb) {
            // U rows of this sub-block for panel columns [cr0, kb): unit-lower solve
            if (tid < kb - cr0) {
                const int cc = cr0 + tid;
                C u[W];
#pragma unroll
                for (int j = 0; j < W; ++j) {
                    if (j < wb) {
                        C v = A[(k + perm[c0 + j]) * n + k + cc];
#pragma unroll
                        for (int l = 0; l < j; ++l) cfma(v, -Ld[j][l], u[l]);
                        u[j] = v;
                        A[(k + perm[c0 + j]) * n + k + cc] = v;
                        Us[j][cc] = v;
                    }
                }
            }
            __syncthreads();
            // rank-wb update of the remaining panel columns of the rows below
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = tid + NT * r;
                if (i >= cr0 && i < m) {
                    C *row = A + (k + perm[i]) * n + k;
                    for (int cc = cr0; cc < kb; ++cc) {
                        C v = row[cc];
#pragma unroll
                        for (int j = 0; j < W; ++j)
                            if (j < wb) cfma(v, -x[r][j], Us[j][cc]);
                        row[cc] = v;
                    }
                }
            }
            __syncthreads();
        }
    }
    // Panel rows into logical order, W columns at a time.
    for (int c0 = 0; c0 < kb; c0 += W) {
        const int wb = kb - c0 < W ? kb - c0 : W;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tid + NT * r;
            if (i < m) {
                const C *row = A + (k + perm[i]) * n + k + c0;
#pragma unroll
                for (int c = 0; c < W; ++c)
                    if (c < wb) x[r][c] = row[c];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tid + NT * r;
            if (i < m) {
                C *row = A + (k + i) * n + k + c0;
#pragma unroll
                for (int c = 0; c < W; ++c)
                    if (c < wb) row[c] = x[r][c];
            }
        }
        __syncthreads();
    }
}

// Apply the panel's interchanges to the rest of each row (columns outside the
// panel), in pivot order.
template <class C>
__global__ void k_rowswap(int64_t n, int64_t k, int kb, C *A0, int64_t sA,
                          const int32_t *ipiv0, int64_t sP, int64_t cbeg, int64_t cend,
                          int64_t sbeg, int64_t send) {
    // columns [cbeg, cend) without [sbeg, send), in pivot order of rows k..k+kb-1
    C *A = A0 + (int64_t)blockIdx.y * sA;
    const int32_t *ipiv = ipiv0 + (int64_t)blockIdx.y * sP;
    const int64_t c0 = cbeg + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t skip = send - sbeg;
    if (c0 >= cend - skip) return;
    const int64_t c = c0 < sbeg ? c0 : c0 + skip;
    for (int j = 0; j < kb; ++j) {
        const int64_t r = k + j, p = ipiv[r];
        if (p != r) {
            C t = A[r * n + c];
            A[r * n + c] = A[p * n + c];
            A[p * n + c] = t;
        }
    }
}

// U12 = L11^{-1} A12 (unit lower L11) for columns [k + kb, cend), one thread per
// column.
template <class C>
__global__ __launch_bounds__(64) void k_trsm_u12(int64_t n, int64_t k, int kb, C *A0,
                                                 int64_t sA, int64_t cend) {
    C *A = A0 + (int64_t)blockIdx.y * sA;
    __shared__ C L[NB][NB];
    for (int t = threadIdx.x; t < kb * kb; t += blockDim.x) {
        int r = t / kb, l = t % kb;
        L[r][l] = A[(k + r) * n + k + l];
    }
    __syncthreads();
    const int64_t c = k + kb + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cend) return;
    C x[NB];
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        if (r < kb) {
            C v = A[(k + r) * n + c];
#pragma unroll
            for (int l = 0; l < r; ++l) cfma(v, -L[r][l], x[l]);
            x[r] = v;
            A[(k + r) * n + c] = v;
        }
    }
}

// Row permutation of the factorisation: perm[i] = original row now at row i,
// from LAPACK's sequential interchanges (one workgroup per system, LDS).
__global__ __launch_bounds__(256) void k_make_perm(int64_t n, const int32_t *ipiv0, int64_t sP,
                                                   int32_t *perm0) {
    extern __shared__ int32_t pm[];
    const int32_t *ipiv = ipiv0 + (int64_t)blockIdx.x * sP;
    int32_t *perm = perm0 + (int64_t)blockIdx.x * sP;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) pm[i] = (int32_t)i;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int64_t j = 0; j < n; ++j) {
            int32_t p = ipiv[j];
            if (p != j) { int32_t t = pm[j]; pm[j] = pm[p]; pm[p] = t; }
        }
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) perm[i] = pm[i];
}

template <class C>
__device__ __forceinline__ C shfl_c(C v, int src) {
    C r;
    r.x = __shfl(v.x, src, 64);
    r.y = __shfl(v.y, src, 64);
    return r;
}

template <class C>
__device__ __forceinline__ C wave_sum(C v) {
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
    }
    return v;
}

// P B, then L y = b (unit lower) and U x = y for nrhs <= NR columns of a
// column-major B, one workgroup per system, left-looking: the rows of block b
// first take the whole dot product with the already solved part (lanes stride
// the contiguous row, RW rows per wavefront, U column chunks in flight, one
// cross-lane reduction per row and right-hand side), then the 64x64 diagonal
// block is solved one wavefront per right-hand side.  XL: the solution is also
// kept in LDS (n * NR entries of dynamic shared memory) for the dot products.
// Grid: x = system (sel, optional, lists them), y = group of NR right-hand sides.
template <class C, int NR, int RW, int U, bool XL>
__global__ __launch_bounds__(1024) void k_trsv_ll(int64_t n, int nrhs, const C *LU0, int64_t sA,
                                                  const int32_t *perm0, int64_t sP, C *B0,
                                                  int64_t sB, C *T0, const int32_t *sel) {
    const int64_t sys = sel ? sel[blockIdx.x] : blockIdx.x;
    // right-hand-side group of this workgroup: columns [NR y, NR y + NR)
    const int64_t cg = (int64_t)blockIdx.y * NR;
    nrhs = nrhs - (int)cg < NR ? nrhs - (int)cg : NR;
    const C *LU = LU0 + sys * sA;
    const int32_t *perm = perm0 + sys * sP;
    C *B = B0 + sys * sB + cg * n;
    C *T = T0 + sys * sB + cg * n;
    __shared__ C D[NB][NB + 1];
    __shared__ C Y[NR][NB];
    extern __shared__ char xlds_raw[];
    C *XS = reinterpret_cast<C *>(xlds_raw);      // [NR][n] when XL
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    const C zero = CTraits<C>::zero();
    for (int64_t t = tid; t < n * nrhs; t += blockDim.x) {
        int64_t i = t % n, c = t / n;
        T[i + n * c] = B[perm[i] + n * c];
    }
    __syncthreads();
    for (int64_t t = tid; t < n * nrhs; t += blockDim.x) B[t] = T[t];
    __syncthreads();
    const C *X = XL ? XS : B;
    const int64_t nblk = (n + NB - 1) / NB;
    for (int pass = 0; pass < 2; ++pass) {
        const bool fwd = pass == 0;
        for (int64_t q = 0; q < nblk; ++q) {
            const int64_t bb = fwd ? q : nblk - 1 - q;
            const int64_t b0 = bb * NB;
            const int bs = (int)min((int64_t)NB, n - b0);
            for (int t = tid; t < bs * bs; t += blockDim.x) {
                int r = t / bs, l = t % bs;
                D[r][l] = LU[(b0 + r) * n + b0 + l];
            }
            const int64_t j0 = fwd ? 0 : b0 + bs, j1 = fwd ? b0 : n;
            for (int r0 = wave * RW; r0 < bs; r0 += nw * RW) {
                C acc[RW][NR];
#pragma unroll
                for (int a = 0; a < RW; ++a)
#pragma unroll
                    for (int c = 0; c < NR; ++c) acc[a][c] = zero;
                const C *rows[RW];
#pragma unroll
                for (int a = 0; a < RW; ++a)
                    rows[a] = LU + (b0 + (r0 + a < bs ? r0 + a : r0)) * n;
                int64_t j = j0 + lane;
                for (; j + 64 * (U - 1) < j1; j += 64 * U) {
                    C l[U][RW];
#pragma unroll
                    for (int u = 0; u < U; ++u)
#pragma unroll
                        for (int a = 0; a < RW; ++a) l[u][a] = rows[a][j + 64 * u];
#pragma unroll
                    for (int u = 0; u < U; ++u)
#pragma unroll
                        for (int c = 0; c < NR; ++c) {
                            if (c < nrhs) {
                                const C xv = X[j + 64 * u + n * c];
#pragma unroll
                                for (int a = 0; a < RW; ++a) cfma(acc[a][c], l[u][a], xv);
                            }
                        }
                }
                for (; j < j1; j += 64) {
#pragma unroll
                    for (int a = 0; a < RW; ++a) {
                        const C lv = rows[a][j];
#pragma unroll
                        for (int c = 0; c < NR; ++c)
                            if (c < nrhs) cfma(acc[a][c], lv, X[j + n * c]);
                    }
                }
#pragma unroll
                for (int a = 0; a < RW; ++a)
#pragma unroll
                    for (int c = 0; c < NR; ++c) {
                        C v = wave_sum(acc[a][c]);
                        if (lane == c && c < nrhs && r0 + a < bs)
                            Y[c][r0 + a] = B[b0 + r0 + a + n * c] - v;
                    }
            }
            __syncthreads();
            if (wave < nrhs) {
                C y = lane < bs ? Y[wave][lane] : zero;
                if (fwd) {
                    for (int r = 0; r < bs; ++r) {
                        C xr = shfl_c(y, r);
                        if (lane > r && lane < bs) cfma(y, -D[lane][r], xr);
                    }
                } else {
                    for (int r = bs - 1; r >= 0; --r) {
                        if (lane == r) y = y / D[r][r];
                        C xr = shfl_c(y, r);
                        if (lane < r) cfma(y, -D[lane][r], xr);
                    }
                }
                if (lane < bs) {
                    B[b0 + lane + n * wave] = y;
                    if (XL) XS[b0 + lane + n * wave] = y;
                }
            }
            __syncthreads();
        }
    }
}

// R = B - A X for row-major A, column-major X/B/R (n x nrhs), one wavefront per row;
// also max |R| and max |X| per (system, rhs), as monotone bit patterns.
template <int NR>
__global__ __launch_bounds__(256) void k_residual(int64_t n, int nrhs, const cd *A0, int64_t sA,
                                                  const cd *X0, const cd *B0, int64_t sB,
                                                  cd *R0, unsigned long long *rmax,
                                                  unsigned long long *xmax, const int32_t *sel) {
    const int64_t b = sel ? sel[blockIdx.y] : blockIdx.y;
    const cd *A = A0 + b * sA;
    const cd *X = X0 + b * sB;
    const cd *Bv = B0 + b * sB;
    cd *R = R0 + b * sB;
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    cd acc[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = cd{0.0, 0.0};
    const cd *row = A + i * n;
    for (int64_t j = lane; j < n; j += 64) {
        cd a = row[j];
#pragma unroll
        for (int c = 0; c < NR; ++c)
            if (c < nrhs) cfma(acc[c], a, X[j + n * c]);
    }
#pragma unroll
    for (int c = 0; c < NR; ++c) {
        for (int off = 32; off > 0; off >>= 1) {
            acc[c].x += __shfl_xor(acc[c].x, off, 64);
            acc[c].y += __shfl_xor(acc[c].y, off, 64);
        }
    }
    if (lane < nrhs && lane < NR) {
        cd a = cd{0.0, 0.0};
#pragma unroll
        for (int c = 0; c < NR; ++c)
            if (c == lane) a = acc[c];
        cd r = Bv[i + n * lane] - a;
        R[i + n * lane] = r;
        double rm = fmax(fabs(r.x), fabs(r.y));
        cd xv = X[i + n * lane];
        double xm = fmax(fabs(xv.x), fabs(xv.y));
        atomicMax(rmax + b * nrhs + lane, (unsigned long long)__double_as_longlong(rm));
        atomicMax(xmax + b * nrhs + lane, (unsigned long long)__double_as_longlong(xm));
    }
}

// Infinity norm (largest row sum of |a_ij|) of each row-major n x n system, as
// a monotone bit pattern through atomicMax; one wavefront per row.
__global__ __launch_bounds__(256) void k_rownorm(int64_t n, const cd *A0, int64_t sA,
                                                 unsigned long long *out) {
    const int64_t b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const cd *row = A0 + b * sA + i * n;
    double s = 0;
    for (int64_t j = lane; j < n; j += 64) s += cabs(row[j]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) atomicMax(out + b, (unsigned long long)__double_as_longlong(s));
}

__global__ void k_cd_to_cf(const cd *a, cf *b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) b[i] = cf{(float)a[i].x, (float)a[i].y};
}

__global__ void k_add_corr_sel(cd *x0, const cf *d0, int64_t sB, const int32_t *sel) {
    const int64_t sys = sel[blockIdx.y];
    cd *x = x0 + sys * sB;
    const cf *d = d0 + sys * sB;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < sB;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = x[i] + cd{(double)d[i].x, (double)d[i].y};
}

__global__ void k_add_corr(cd *x, const cf *d, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) x[i] = x[i] + cd{(double)d[i].x, (double)d[i].y};
}

unsigned grid1(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

template <class C>
void gemm_t(int m, int n, int k, C alpha, const C *A, int64_t lda, int64_t sA, const C *B,
            int64_t ldb, int64_t sB, C beta, C *Cm, int64_t ldc, int64_t sC, int batch,
            hipStream_t st) {
    if (m <= 0 || n <= 0 || batch <= 0) return;
    dim3 grid((unsigned)((m + 63) / 64), (unsigned)((n + 63) / 64), (unsigned)batch);
    hipLaunchKernelGGL(k_gemm_mfma<C>, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, sA, B, ldb,
                       sB, beta, Cm, ldc, sC);
    HIPCHECK(hipGetLastError());
}

// Register-resident panel for panels of up to 12288 rows, the streaming panel
// beyond.  W is chosen so the R x W register block needs no scratch.
template <class C>
void launch_panel(int64_t n, int64_t k, int kb, C *A, int64_t sA, int32_t *ipiv, int64_t sP,
                  int batch, hipStream_t st) {
    const int64_t m = n - k;
    constexpr bool F = sizeof(C) == sizeof(cf);
    dim3 g((unsigned)batch), b(1024);
#define PR(RR, WW) hipLaunchKernelGGL((k_panel_reg<C, RR, WW>), g, b, 0, st, n, k, kb, A, sA, ipiv, sP)
    if constexpr (F) {
        if (m <= 1024) PR(1, 8);
        else if (m <= 2048) PR(2, 8);
        else if (m <= 3072) PR(3, 8);
        else if (m <= 6144) PR(6, 4);
        else if (m <= 12288) PR(12, 2);
        else hipLaunchKernelGGL(k_panel<C>, g, b, 0, st, n, k, kb, A, sA, ipiv, sP);
    } else {
        if (m <= 1024) PR(1, 8);
        else if (m <= 2048) PR(2, 4);
        else if (m <= 3072) PR(3, 4);
        else if (m <= 6144) PR(6, 2);
        else if (m <= 12288) PR(12, 1);
        else hipLaunchKernelGGL(k_panel<C>, g, b, 0, st, n, k, kb, A, sA, ipiv, sP);
    }
#undef PR
    HIPCHECK(hipGetLastError());
}

template <class C>
void rowswap_t(int64_t n, int64_t k, int kb, C *A, int64_t sA, const int32_t *ipiv, int64_t sP,
               int64_t cbeg, int64_t cend, int64_t sbeg, int64_t send, int batch,
               hipStream_t st) {
    const int64_t cols = (cend - cbeg) - (send - sbeg);
    if (cols <= 0) return;
    dim3 g((unsigned)((cols + 255) / 256), (unsigned)batch);
    hipLaunchKernelGGL(k_rowswap<C>, g, dim3(256), 0, st, n, k, kb, A, sA, ipiv, sP, cbeg, cend,
                       sbeg, send);
    HIPCHECK(hipGetLastError());
}

template <class C>
void trsm_u12_t(int64_t n, int64_t k, int kb, C *A, int64_t sA, int64_t cend, int batch,
                hipStream_t st) {
    const int64_t cols = cend - (k + kb);
    if (cols <= 0) return;
    dim3 g((unsigned)((cols + 63) / 64), (unsigned)batch);
    hipLaunchKernelGGL(k_trsm_u12<C>, g, dim3(64), 0, st, n, k, kb, A, sA, cend);
    HIPCHECK(hipGetLastError());
}

// A[r0:n, c0:c1] -= A[r0:n, kk:kk+kw] A[kk:kk+kw, c0:c1] (row-major, r0 = c0 = kk + kw)
// as a column-major GEMM on the transposed views.
template <class C>
void update_t(int64_t n, int64_t kk, int kw, int64_t c1, C *A, int64_t sA, int batch,
              hipStream_t st) {
    const int64_t r0 = kk + kw;
    const int64_t ncols = c1 - r0, nrows = n - r0;
    if (ncols <= 0 || nrows <= 0) return;
    C mone, one;
    mone.x = -1; mone.y = 0;
    one.x = 1; one.y = 0;
    gemm_t<C>((int)ncols, (int)nrows, kw, mone, A + kk * n + r0, n, sA, A + r0 * n + kk, n, sA,
              one, A + r0 * n + r0, n, sA, batch, st);
}

// Panels taller than the register kernel's sweet spot: 8-column leaves factored by
// the register kernel, each followed by its interchanges on the panel's other
// columns, the U rows it needs and a rank-8 GEMM update of the panel's remaining
// columns (multi-workgroup launches, so a few large systems still fill the chip).
template <class C>
void panel_recursive(int64_t n, int64_t k, int kb, C *A, int64_t sA, int32_t *ipiv,
                     int64_t sP, int batch, hipStream_t st) {
    constexpr int LW = 8;
    for (int j0 = 0; j0 < kb; j0 += LW) {
        const int lw = std::min(LW, kb - j0);
        const int64_t kk = k + j0;
        launch_panel<C>(n, kk, lw, A, sA, ipiv, sP, batch, st);
        rowswap_t<C>(n, kk, lw, A, sA, ipiv, sP, k, k + kb, kk, kk + lw, batch, st);
        trsm_u12_t<C>(n, kk, lw, A, sA, k + kb, batch, st);
        update_t<C>(n, kk, lw, k + kb, A, sA, batch, st);
    }
}

template <class C>
void lu_factor_t(int64_t n, int batch, C *A, int64_t sA, int32_t *ipiv, int64_t sP,
                 hipStream_t st) {
    for (int64_t k = 0; k < n; k += NB) {
        const int kb = (int)std::min<int64_t>(NB, n - k);
        stage_begin(3, st);
        if (n - k > RECURSIVE_PANEL_ROWS)
            panel_recursive<C>(n, k, kb, A, sA, ipiv, sP, batch, st);
        else
            launch_panel<C>(n, k, kb, A, sA, ipiv, sP, batch, st);
        rowswap_t<C>(n, k, kb, A, sA, ipiv, sP, 0, n, k, k + kb, batch, st);
        trsm_u12_t<C>(n, k, kb, A, sA, n, batch, st);
        stage_end(3, st);
        const int64_t rest = n - k - kb;
        if (rest > 0) {
            // Row-major A22 -= L21 U12, as a column-major GEMM on the transposes:
            // A22^T -= U12^T L21^T, every view having leading dimension n.
            stage_begin(4, st);
            update_t<C>(n, k, kb, n, A, sA, batch, st);
            stage_end(4, st, 8.0 * (double)rest * (double)rest * kb * batch);
        }
    }
    stage_flops(7, 8.0 / 3.0 * (double)n * n * n * batch);
}

// Largest shared memory per workgroup on the current device (bytes).
size_t max_lds_bytes() {
    int dev = 0, v = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    return (size_t)v;
}

template <class C>
void lu_solve_t(int64_t n, int nrhs, int batch, const C *LU, int64_t sA, const int32_t *perm,
                int64_t sP, C *B, int64_t sB, C *scratch, hipStream_t st,
                const int32_t *sel = nullptr) {
    if (batch <= 0 || nrhs <= 0) return;
    constexpr int NR = 4;
    const size_t static_lds = sizeof(C) * (NB * (NB + 1) + NR * NB);
    const size_t xl = sizeof(C) * (size_t)n * NR;
    const bool useX = static_lds + xl <= max_lds_bytes();
    dim3 g((unsigned)batch, (unsigned)((nrhs + NR - 1) / NR)), b(1024);
    if constexpr (sizeof(C) == sizeof(cf)) {
        if (useX)
            hipLaunchKernelGGL((k_trsv_ll<C, NR, 2, 4, true>), g, b, xl, st, n, nrhs, LU, sA,
                               perm, sP, B, sB, scratch, sel);
        else
            hipLaunchKernelGGL((k_trsv_ll<C, NR, 2, 4, false>), g, b, 0, st, n, nrhs, LU, sA,
                               perm, sP, B, sB, scratch, sel);
    } else {
        if (useX)
            hipLaunchKernelGGL((k_trsv_ll<C, NR, 2, 4, true>), g, b, xl, st, n, nrhs, LU, sA,
                               perm, sP, B, sB, scratch, sel);
        else
            hipLaunchKernelGGL((k_trsv_ll<C, NR, 2, 2, false>), g, b, 0, st, n, nrhs, LU, sA,
                               perm, sP, B, sB, scratch, sel);
    }
    HIPCHECK(hipGetLastError());
}

}  // namespace

void zgemm_batched(int m, int n, int k, cd alpha, const cd *A, int64_t lda, int64_t sA,
                   const cd *B, int64_t ldb, int64_t sB, cd beta, cd *C, int64_t ldc,
                   int64_t sC, int batch, hipStream_t st) {
    gemm_t<cd>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, st);
}

void lu_factor_batched(int64_t n, int batch, cd *A, int64_t sA, int32_t *ipiv, int64_t sP,
                       hipStream_t st) {
    lu_factor_t<cd>(n, batch, A, sA, ipiv, sP, st);
}

void lu_solve_batched(int64_t n, int nrhs, int batch, const cd *LU, int64_t sA,
                      const int32_t *perm, int64_t sP, cd *B, int64_t sB, cd *scratch,
                      hipStream_t st) {
    lu_solve_t<cd>(n, nrhs, batch, LU, sA, perm, sP, B, sB, scratch, st);
}

void lu_make_perm(int64_t n, int batch, const int32_t *ipiv, int64_t sP, int32_t *perm,
                  hipStream_t st) {
    size_t lds = (size_t)n * sizeof(int32_t);
    if (lds > 160 * 1024) throw ArgError("lu_make_perm: n too large for the LDS permutation");
    hipLaunchKernelGGL(k_make_perm, dim3((unsigned)batch), dim3(256), lds, st, n, ipiv, sP, perm);
    HIPCHECK(hipGetLastError());
}

void lu_factor_batched_f32(int64_t n, int batch, cf *A, int64_t sA, int32_t *ipiv, int64_t sP,
                           hipStream_t st) {
    lu_factor_t<cf>(n, batch, A, sA, ipiv, sP, st);
}

void lu_solve_batched_f32(int64_t n, int nrhs, int batch, const cf *LU, int64_t sA,
                          const int32_t *perm, int64_t sP, cf *B, int64_t sB, cf *scratch,
                          hipStream_t st, const int32_t *sel) {
    lu_solve_t<cf>(n, nrhs, batch, LU, sA, perm, sP, B, sB, scratch, st, sel);
}

void convert_z_to_f32(const cd *Z, cf *Zf, int64_t count, hipStream_t st) {
    hipLaunchKernelGGL(k_cd_to_cf, dim3(grid1(count)), dim3(256), 0, st, Z, Zf, count);
    HIPCHECK(hipGetLastError());
}

void residual_batched(int64_t n, int nrhs, int batch, const cd *A, int64_t sA, const cd *X,
                      const cd *Bv, int64_t sB, cd *R, double *rnorm, double *xnorm,
                      hipStream_t st, const int32_t *sel, int nsel) {
    if (nrhs > 8) throw ArgError("residual_batched: at most 8 right-hand sides per call");
    launch_zero((cd *)rnorm, (batch * nrhs + 1) / 2, st);
    launch_zero((cd *)xnorm, (batch * nrhs + 1) / 2, st);
    const int ns = sel ? nsel : batch;
    if (ns <= 0) return;
    dim3 g((unsigned)((n + 3) / 4), (unsigned)ns);
    if (nrhs <= 4)
        hipLaunchKernelGGL(k_residual<4>, g, dim3(256), 0, st, n, nrhs, A, sA, X, Bv, sB, R,
                           (unsigned long long *)rnorm, (unsigned long long *)xnorm, sel);
    else
        hipLaunchKernelGGL(k_residual<8>, g, dim3(256), 0, st, n, nrhs, A, sA, X, Bv, sB, R,
                           (unsigned long long *)rnorm, (unsigned long long *)xnorm, sel);
    HIPCHECK(hipGetLastError());
}

void residual_to_f32(const cd *R, cf *Rf, int64_t count, hipStream_t st) {
    convert_z_to_f32(R, Rf, count, st);
}

void add_correction(cd *X, const cf *D, int64_t count, hipStream_t st) {
    hipLaunchKernelGGL(k_add_corr, dim3(grid1(count)), dim3(256), 0, st, X, D, count);
    HIPCHECK(hipGetLastError());
}

void add_correction_sel(cd *X, const cf *D, int64_t sB, const int32_t *sel, int nsel,
                        hipStream_t st) {
    if (nsel <= 0) return;
    int64_t bx = (sB + 255) / 256;
    dim3 g((unsigned)(bx > 64 ? 64 : bx), (unsigned)nsel);
    hipLaunchKernelGGL(k_add_corr_sel, g, dim3(256), 0, st, X, D, sB, sel);
    HIPCHECK(hipGetLastError());
}

void matrix_inf_norms(int64_t n, int batch, const cd *A, int64_t sA, double *out,
                      hipStream_t st) {
    launch_zero((cd *)out, (batch + 1) / 2, st);
    dim3 g((unsigned)((n + 3) / 4), (unsigned)batch);
    hipLaunchKernelGGL(k_rownorm, g, dim3(256), 0, st, n, A, sA, (unsigned long long *)out);
    HIPCHECK(hipGetLastError());
}
