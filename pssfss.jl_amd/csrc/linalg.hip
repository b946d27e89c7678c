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

constexpr int NB = 128;  // panel width (trailing updates are rank-128)
constexpr int SB = 64;   // sub-panel width: triangular blocks held in LDS

// ------------------------------------------------------------------ GEMM (NN)
// C = alpha*A*B + beta*C, column-major complex, batched by blockIdx.z, on the
// matrix cores.  A complex product is a real product with K doubled: with
// (re, im) interleaved along k, Re C = [Ar, -Ai] [Br; Bi] and
// Im C = [Ai, Ar] [Br; Bi], so B feeds the 16x16x4 MFMA in its memory layout
// and each 4-deep real step covers two complex k.  The matrix cores compute the
// transposed tile (B^T A^T), so the 16 lanes that share an output row of the
// instruction hold 16 consecutive rows of the column-major C: the epilogue's
// loads and stores are coalesced.  2 x WN wavefronts, each owning (16 WT) x
// (32 WT / WN) of the (32 WT) x (32 WT) workgroup tile; BK complex k per LDS
// stage, double-buffered: 8 in FP64, 16 in FP32.  WT = 2 (64 x 64) is what the library
// launches: the 128 x 128 tile (WT = 4) runs one wavefront per SIMD and measured
// slower on MI355X.  FP32 launches WN = 4 (eight wavefronts of 32 x 16, 60
// registers: eight per SIMD instead of four, which hides more of the operand and
// C loads; 0.850 -> 0.794 ms per two-layer trailing update); FP64 keeps WN = 2 (at
// WN = 4 it was no faster).  TA: A is given transposed (C = A^T B, A stored k x m), which
// is a row-major m x k operand such as Z.  sel (optional) lists the batch entries.
// FP32 uses three real products per complex k instead (Gauss): T1 = Ar Br,
// T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi), Re C = T1 - T2, Im C = T3 - T1 - T2, with
// four complex k per MFMA step: 3/4 of the matrix-core work and half the LDS
// operand reads.  Its error bound is normwise (Higham, "Stability of a method for
// multiplying complex matrices with three real matrix multiplications", 1992);
// the FP32 factors it builds are only used under FP64 refinement.
// TAG only names instantiations apart in kernel traces (1: the LU's trailing
// updates); the code is the same.
template <class C, int WT, bool TA, int TAG = 0, int WN = 2>
__global__ __launch_bounds__(128 * WN) void k_gemm_mfma(int m, int n, int k, C alpha,
                                                   const C *__restrict__ A, int64_t lda,
                                                   int64_t sA, const C *__restrict__ B,
                                                   int64_t ldb, int64_t sB, C beta, C *Cm,
                                                   int64_t ldc, int64_t sC, const int32_t *sel) {
    typedef typename CTraits<C>::R R;
    typedef MfmaT<R> MT;
    typedef typename MT::V V;
    // FP64: 8 complex k per LDS stage (33 KB double-buffered, three workgroups per
    // CU by registers) instead of 16 (66 KB, two workgroups per CU by LDS)
    constexpr int BM = 32 * WT, BN = 32 * WT, BK = sizeof(C) == sizeof(cd) ? 8 : 16,
                  WS = 16 * WT, NT = 128 * WN, WSN = BN / WN, TJ = WSN / 16;
    constexpr int NL = BM * BK / NT;   // elements of each operand per thread and stage
    constexpr bool G3 = sizeof(C) == sizeof(cf);   // three-product complex form
    constexpr int NACC = G3 ? 3 : 2;
    __shared__ C As[2][BK][BM + (TA ? 1 : 0)];
    __shared__ C Bs[2][BK][BN + 1];   // padded: the k-fastest stores spread over banks
    const int64_t bz = sel ? sel[blockIdx.z] : blockIdx.z;
    A += bz * sA;
    B += bz * sB;
    Cm += bz * sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
    const C zero = CTraits<C>::zero();
    V acc[WT][TJ][NACC];
#pragma unroll
    for (int a = 0; a < WT; ++a)
#pragma unroll
        for (int b = 0; b < TJ; ++b)
#pragma unroll
            for (int c = 0; c < NACC; ++c) acc[a][b][c] = V{0, 0, 0, 0};
    C ra[NL], rb[NL];
#define GEMM_LOAD(K0)                                                                          \
    _Pragma("unroll") for (int q = 0; q < NL; ++q) {                                          \
        const int t = tid + NT * q;                                                            \
        const int i = TA ? t / BK : t % BM, kk = TA ? t % BK : t / BM;                         \
        ra[q] = zero;                                                                          \
        if (m0 + i < m && (K0) + kk < k)                                                       \
            ra[q] = TA ? A[((K0) + kk) + lda * (m0 + i)] : A[(m0 + i) + lda * ((K0) + kk)];    \
        const int kb = t % BK, j = t / BK;                                                     \
        rb[q] = zero;                                                                          \
        if ((K0) + kb < k && n0 + j < n) rb[q] = B[((K0) + kb) + ldb * (n0 + j)];              \
    }
#define GEMM_STORE(BUF)                                                                        \
    _Pragma("unroll") for (int q = 0; q < NL; ++q) {                                          \
        const int t = tid + NT * q;                                                            \
        if (TA) As[BUF][t % BK][t / BK] = ra[q];                                               \
        else As[BUF][t / BM][t % BM] = ra[q];                                                  \
        Bs[BUF][t % BK][t / BK] = rb[q];                                                       \
    }
    const int nk = (k + BK - 1) / BK;
    GEMM_LOAD(0)
    GEMM_STORE(0)
    __syncthreads();
    const int c = (lane >> 4) & 1;        // re / im half of the doubled k
    const int kh = lane >> 5;             // which of the two complex k of a step
    // operands of two steps in flight: fewer registers, more waves per SIMD
    constexpr int UNR = 2;
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) { GEMM_LOAD((kt + 1) * BK) }
        if constexpr (G3) {
#pragma unroll UNR
        for (int s4 = 0; s4 < BK / 4; ++s4) {
            const int kc = 4 * s4 + (lane >> 4);
            R ar[WT], ai[WT], as[WT], br[TJ], bi[TJ], bs[TJ];
#pragma unroll
            for (int ti = 0; ti < WT; ++ti) {
                const C a = As[buf][kc][WS * wm + 16 * ti + (lane & 15)];
                ar[ti] = a.x;
                ai[ti] = a.y;
                as[ti] = a.x + a.y;
            }
#pragma unroll
            for (int tj = 0; tj < TJ; ++tj) {
                const C b = Bs[buf][kc][WSN * wn + 16 * tj + (lane & 15)];
                br[tj] = b.x;
                bi[tj] = b.y;
                bs[tj] = b.x + b.y;
            }
#pragma unroll
            for (int ti = 0; ti < WT; ++ti)
#pragma unroll
                for (int tj = 0; tj < TJ; ++tj) {
                    acc[ti][tj][0] = MT::mma(br[tj], ar[ti], acc[ti][tj][0]);
                    acc[ti][tj][1] = MT::mma(bi[tj], ai[ti], acc[ti][tj][1]);
                    acc[ti][tj][2] = MT::mma(bs[tj], as[ti], acc[ti][tj][2]);
                }
        }
        } else {
#pragma unroll UNR
        for (int s2 = 0; s2 < BK / 2; ++s2) {
            const int kc = 2 * s2 + kh;
            R are[WT], aim[WT], bop[TJ];
#pragma unroll
            for (int ti = 0; ti < WT; ++ti) {
                const C a = As[buf][kc][WS * wm + 16 * ti + (lane & 15)];
                are[ti] = c == 0 ? a.x : -a.y;
                aim[ti] = c == 0 ? a.y : a.x;
            }
#pragma unroll
            for (int tj = 0; tj < TJ; ++tj) {
                const C b = Bs[buf][kc][WSN * wn + 16 * tj + (lane & 15)];
                bop[tj] = c == 0 ? b.x : b.y;
            }
#pragma unroll
            for (int ti = 0; ti < WT; ++ti)
#pragma unroll
                for (int tj = 0; tj < TJ; ++tj) {
                    acc[ti][tj][0] = MT::mma(bop[tj], are[ti], acc[ti][tj][0]);
                    acc[ti][tj][1] = MT::mma(bop[tj], aim[ti], acc[ti][tj][1]);
                }
        }
        }
        if (kt + 1 < nk) { GEMM_STORE(buf ^ 1) }
        __syncthreads();
    }
#undef GEMM_LOAD
#undef GEMM_STORE
    const bool use_beta = !(beta.x == 0 && beta.y == 0);
#pragma unroll
    for (int ti = 0; ti < WT; ++ti)
#pragma unroll
        for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t i = m0 + WS * wm + 16 * ti + (lane & 15);
                const int64_t j = n0 + WSN * wn + 16 * tj + MT::row(lane, e);
                if (i < m && j < n) {
                    C v;
                    if constexpr (G3) {
                        const R t1 = acc[ti][tj][0][e], t2 = acc[ti][tj][1][e];
                        v.x = t1 - t2;
                        v.y = acc[ti][tj][2][e] - t1 - t2;
                    } else {
                        v.x = acc[ti][tj][0][e];
                        v.y = acc[ti][tj][1][e];
                    }
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

// The NB-column panel (rows [k, n)) is factored in a column-major scratch copy
// (m = n - k rows, leading dimension m, one per system), so the thread-per-row
// accesses of the leaf kernel are coalesced.  dir 0 copies the panel in, dir 1
// copies it back; 64-row x 64-column tiles are transposed through LDS.
template <class C>
__global__ __launch_bounds__(256) void k_panel_copy(int64_t n, int64_t k, int kb, C *A0,
                                                    int64_t sA, C *P0, int64_t sPt, int dir) {
    __shared__ C T[64][65];
    const int64_t m = n - k;
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    C *A = A0 + (int64_t)blockIdx.y * sA;
    C *P = P0 + (int64_t)blockIdx.y * sPt;
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    // Loads go to registers first, from an in-bounds address where masked (element
    // 0; those LDS entries are never read), so that all 16 issue before a wait.
    C v[16];
    for (int cb = 0; cb < kb; cb += 64) {   // 64-column blocks of the panel
        if (dir == 0) {
#pragma unroll
            for (int p = 0; p < 16; ++p) {
                const int rr = q + 4 * p;
                const bool ok = r0 + rr < m && cb + lane < kb;
                v[p] = A[ok ? (k + r0 + rr) * n + k + cb + lane : 0];
            }
#pragma unroll
            for (int p = 0; p < 16; ++p) T[q + 4 * p][lane] = v[p];
            __syncthreads();
            for (int p = 0; p < 16; ++p) {
                const int cc = q + 4 * p;
                if (r0 + lane < m && cb + cc < kb) P[(cb + cc) * m + r0 + lane] = T[lane][cc];
            }
        } else {
#pragma unroll
            for (int p = 0; p < 16; ++p) {
                const int cc = q + 4 * p;
                const bool ok = r0 + lane < m && cb + cc < kb;
                v[p] = P[ok ? (cb + cc) * m + r0 + lane : 0];
            }
#pragma unroll
            for (int p = 0; p < 16; ++p) T[lane][q + 4 * p] = v[p];
            __syncthreads();
            for (int p = 0; p < 16; ++p) {
                const int rr = q + 4 * p;
                if (r0 + rr < m && cb + lane < kb) A[(k + r0 + rr) * n + k + cb + lane] = T[rr][lane];
            }
        }
        __syncthreads();
    }
}

// Argmax over a wavefront (largest value, first index on ties: izamax's rule)
// with DPP lane moves inside each row of 16 lanes (xor 1, xor 2, half mirror,
// mirror), then the four row results read out by readlane; no LDS round trips
// and no branches.  Every lane must be active; every lane gets the result.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, 0xf, 0xf,
                                               false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float read_lane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <class Rl>
__device__ __forceinline__ void amax_pick(Rl &v, int &i, Rl o, int oi) {
    const bool t = (o > v) | ((o == v) & (oi < i));
    v = t ? o : v;
    i = t ? oi : i;
}
template <int CTRL, class Rl>
__device__ __forceinline__ void amax_dpp(Rl &v, int &i) {
    amax_pick(v, i, dpp_mov<CTRL>(v), __builtin_amdgcn_update_dpp(i, i, CTRL, 0xf, 0xf, false));
}
template <class Rl>
__device__ __forceinline__ void wave_argmax(Rl &v, int &i) {
    amax_dpp<0xB1>(v, i);    // quad_perm [1, 0, 3, 2]
    amax_dpp<0x4E>(v, i);    // quad_perm [2, 3, 0, 1]
    amax_dpp<0x141>(v, i);   // row_half_mirror
    amax_dpp<0x140>(v, i);   // row_mirror: every lane holds its row's best
    Rl b = read_lane(v, 0);
    int bi = __builtin_amdgcn_readlane(i, 0);
#pragma unroll
    for (int r = 1; r < 4; ++r)
        amax_pick(b, bi, read_lane(v, 16 * r), __builtin_amdgcn_readlane(i, 16 * r));
    v = b;
    i = bi;
}

// Factor leaf columns [c0, c0 + lw) of the scratch panel over rows [c0, m) with
// partial pivoting (largest |re| + |im|, first index on ties: izamax), the rows
// in registers, R per thread.  An interchange swaps two register rows, so the
// rows go back to memory in the order LAPACK's sequential interchanges give.
// Two barriers per column: after the wavefront winners of the pivot search, and
// after the pivot row (and, on an interchange, row j) is staged in LDS.  The
// pivots are kept in LDS and written once after the column loop: a global store
// inside it makes the next LDS staging of its wavefront wait for the store.
// ipiv receives absolute pivot rows (k is the panel's first row).
// When the panel (kb columns) is wider than the leaf, the same workgroup then applies
// the leaf's interchanges to the panel's other columns and forms the leaf's U rows
// for the columns to its right up to ce, the end of its sub-panel (unit-lower solve
// with the leaf's lw x lw block): one thread per panel column.  Thread 0 first
// resolves the lw interchanges on row indices (the rows they touch, at most 2 lw, and
// for each the row whose value ends up there), so that every column's thread issues
// all its loads before it stores any value.
template <class C, int R, int LW>
__global__ __launch_bounds__(1024) void k_leaf(int64_t m, int c0, int lw, int kb, int ce,
                                               C *P0, int64_t sPt, int32_t *ipiv0, int64_t sP,
                                               int64_t k) {
    typedef typename CTraits<C>::R Rl;
    constexpr int NT = 1024;
    C *P = P0 + (int64_t)blockIdx.x * sPt;
    int32_t *ipiv = ipiv0 + (int64_t)blockIdx.x * sP;
    const int mr = (int)(m - c0);
    __shared__ C xch[2][LW];
    __shared__ Rl rv[16];
    __shared__ int ri[16];
    __shared__ int pv[LW];
    __shared__ C L11[LW][LW];
    __shared__ int trow[2 * LW], tsrc[2 * LW];
    __shared__ int ntouch;
    const int tid = threadIdx.x, lane = tid & 63;
    const C zero = CTraits<C>::zero();
    C x[R][LW];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = tid + NT * r;
#pragma unroll
        for (int c = 0; c < LW; ++c) {
            x[r][c] = zero;
            if (i < mr && c < lw) x[r][c] = P[(int64_t)(c0 + c) * m + c0 + i];
        }
    }
#pragma unroll
    for (int j = 0; j < LW; ++j) {
        if (j < lw) {
        Rl best = (Rl)-1;
        int bi = 0x7fffffff;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tid + NT * r;
            if (i >= j && i < mr) {
                const Rl v = cabs1(x[r][j]);
                if (v > best) { best = v; bi = i; }
            }
        }
        wave_argmax(best, bi);
        if (lane == 0) { rv[tid >> 6] = best; ri[tid >> 6] = bi; }
        __syncthreads();
        // every wavefront reduces the 16 wavefront winners (no second barrier)
        int p;
        {
            Rl b = lane < NT / 64 ? rv[lane] : (Rl)-1;
            int bx = lane < NT / 64 ? ri[lane] : 0x7fffffff;
            wave_argmax(b, bx);
            if (bx >= mr) bx = j;   // no comparable entry (NaN column): keep the diagonal
            p = bx;
        }
        if (tid == 0) pv[j] = p;
        // the pivot row goes to xch[0] (it is also the row the update uses); on an
        // interchange row j goes to xch[1] for the pivot row's owner
        const int op = p % NT, sp = p / NT;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (tid == op && r == sp)
#pragma unroll
                for (int c = 0; c < LW; ++c) xch[0][c] = x[r][c];
        if (p != j && tid == j)
#pragma unroll
            for (int c = 0; c < LW; ++c) xch[1][c] = x[0][c];
        __syncthreads();
        if (p != j) {
            if (tid == j)
#pragma unroll
                for (int c = 0; c < LW; ++c) x[0][c] = xch[0][c];
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (tid == op && r == sp)
#pragma unroll
                    for (int c = 0; c < LW; ++c) x[r][c] = xch[1][c];
        }
        const C piv = xch[0][j];
        const bool nz = !(piv.x == 0 && piv.y == 0);
        const C rp = nz ? CTraits<C>::one() / piv : zero;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = tid + NT * r;
            if (i > j && i < mr) {
                const C l = nz ? x[r][j] * rp : x[r][j];
                x[r][j] = l;
                const C ml = -l;
                // at LW = 16 a constant trip count lets this loop unroll fully
                // once the column loop has (else x is demoted to scratch memory)
#pragma unroll
                for (int c = (LW >= 16 ? 0 : j + 1); c < LW; ++c)
                    if (c > j) cfma(x[r][c], ml, xch[0][c]);
            }
        }
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = tid + NT * r;
        if (i < mr)
#pragma unroll
            for (int c = 0; c < LW; ++c)
                if (c < lw) P[(int64_t)(c0 + c) * m + c0 + i] = x[r][c];
    }
    if (tid < lw) ipiv[k + c0 + tid] = (int32_t)(k + c0 + pv[tid]);
    if (kb <= lw) return;
    // rows 0..lw-1 of the leaf are register row 0 of threads 0..lw-1
    if (tid < lw)
#pragma unroll
        for (int c = 0; c < LW; ++c) L11[tid][c] = x[0][c];
    if (tid == 0) {
        int cnt = 0;
        for (int j = 0; j < lw; ++j) {
            const int r = c0 + j, p = c0 + pv[j];
            if (p == r) continue;
            int a = -1, b = -1;
            for (int t = 0; t < cnt; ++t) {
                if (trow[t] == r) a = t;
                if (trow[t] == p) b = t;
            }
            if (a < 0) { trow[cnt] = r; tsrc[cnt] = r; a = cnt++; }
            if (b < 0) { trow[cnt] = p; tsrc[cnt] = p; b = cnt++; }
            const int t2 = tsrc[a];
            tsrc[a] = tsrc[b];
            tsrc[b] = t2;
        }
        ntouch = cnt;
    }
    __syncthreads();
    const int c = tid;
    if (c >= kb || (c >= c0 && c < c0 + lw)) return;
    C *col = P + (int64_t)c * m;
    const int cnt = ntouch;
    Rl vr[2 * LW], vi[2 * LW];      // parts apart: a complex array went to scratch in FP64
#pragma unroll
    for (int t = 0; t < 2 * LW; ++t) {
        const C v = col[t < cnt ? tsrc[t] : c0];
        vr[t] = v.x;
        vi[t] = v.y;
    }
#pragma unroll
    for (int t = 0; t < 2 * LW; ++t)
        if (t < cnt) {
            C v;
            v.x = vr[t];
            v.y = vi[t];
            col[trow[t]] = v;
        }
    if (c >= c0 + lw && c < ce) {
        C u[LW];
#pragma unroll
        for (int j = 0; j < LW; ++j) {
            if (j < lw) {
                C w = col[c0 + j];
#pragma unroll
                for (int l = 0; l < j; ++l) cfma(w, -L11[j][l], u[l]);
                u[j] = w;
                col[c0 + j] = w;
            }
        }
    }
}

// U rows [0, SB) of the scratch panel's second sub-panel: L00^{-1} times columns
// [SB, kb) (L00 the unit-lower SB x SB block), one thread per column.
template <class C>
__global__ __launch_bounds__(64) void k_pt_trsm(int64_t m, int kb, C *P0, int64_t sPt) {
    C *P = P0 + (int64_t)blockIdx.x * sPt;
    __shared__ C L[SB][SB];
    for (int t = threadIdx.x; t < SB * SB; t += 64) {
        const int r = t % SB, l = t / SB;
        L[r][l] = P[(int64_t)l * m + r];
    }
    __syncthreads();
    const int c = SB + threadIdx.x;
    if (c >= kb) return;
    C *col = P + (int64_t)c * m;
    C x[SB];
#pragma unroll
    for (int r = 0; r < SB; ++r) {
        C v = col[r];
#pragma unroll
        for (int l = 0; l < r; ++l) cfma(v, -L[r][l], x[l]);
        x[r] = v;
        col[r] = v;
    }
}

// Apply the panel's interchanges to the rest of each row (columns outside the
// panel), in pivot order.
template <class C>
__global__ void k_rowswap(int64_t n, int64_t k, int kb, C *A0, int64_t sA,
                          const int32_t *ipiv0, int64_t sP, int64_t cbeg, int64_t cend,
                          int64_t sbeg, int64_t send) {
    // columns [cbeg, cend) without [sbeg, send), in pivot order of rows k..k+kb-1;
    // the pivots are staged in LDS first, so that the loop waits only on swaps
    __shared__ int32_t pv[NB];
    C *A = A0 + (int64_t)blockIdx.y * sA;
    const int32_t *ipiv = ipiv0 + (int64_t)blockIdx.y * sP;
    for (int t = threadIdx.x; t < kb; t += blockDim.x) pv[t] = ipiv[k + t];
    __syncthreads();
    const int64_t c0 = cbeg + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t skip = send - sbeg;
    if (c0 >= cend - skip) return;
    const int64_t c = c0 < sbeg ? c0 : c0 + skip;
    for (int j = 0; j < kb; ++j) {
        const int64_t r = k + j, p = pv[j];
        if (p != r) {
            C t = A[r * n + c];
            A[r * n + c] = A[p * n + c];
            A[p * n + c] = t;
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

// T = P B (row i of T is row perm[i] of B) for nrhs columns, the systems of the
// batch (or those listed in sel).
template <class C>
__global__ void k_perm_rows(int64_t n, int nrhs, const int32_t *perm0, int64_t sP, const C *B0,
                            C *T0, int64_t sB, const int32_t *sel) {
    const int64_t sys = sel ? sel[blockIdx.y] : blockIdx.y;
    const int32_t *perm = perm0 + sys * sP;
    const C *B = B0 + sys * sB;
    C *T = T0 + sys * sB;
    const int64_t tot = n * nrhs;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % n, c = t / n;
        T[t] = B[perm[i] + n * c];
    }
}

template <class C>
__global__ void k_copy_sys(int64_t cnt, const C *S0, C *D0, int64_t sB, const int32_t *sel) {
    const int64_t sys = sel ? sel[blockIdx.y] : blockIdx.y;
    const C *S = S0 + sys * sB;
    C *D = D0 + sys * sB;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cnt;
         t += (int64_t)gridDim.x * blockDim.x)
        D[t] = S[t];
}

// B[perm[i]] = T[i] (the inverse of k_perm_rows), for the transposed solves.
template <class C>
__global__ void k_unperm_rows(int64_t n, int nrhs, const int32_t *perm0, int64_t sP, const C *T0,
                              C *B0, int64_t sB, const int32_t *sel) {
    const int64_t sys = sel ? sel[blockIdx.y] : blockIdx.y;
    const int32_t *perm = perm0 + sys * sP;
    const C *T = T0 + sys * sB;
    C *B = B0 + sys * sB;
    const int64_t tot = n * nrhs;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % n, c = t / n;
        B[perm[i] + n * c] = T[t];
    }
}

// Inverse of one triangular diagonal block of a row-major factor, rows and columns
// [b0, b0 + bs) (bs <= 64): unit lower (upper false) or upper.  Called by all 64
// threads of a workgroup; thread c computes column c by solving L x = e_c (U x =
// e_c) with the block in LDS.  Writes the row-major 64 x 64 inverse (zero beyond
// bs) to out.
template <class C>
__device__ void tri_inverse_block(const C *LU, int64_t n, int64_t b0, int bs, bool upper,
                                  C *out) {
    __shared__ C D[64][65];
    __shared__ C X[64][65];
    const int c = threadIdx.x;
    const C zero = CTraits<C>::zero(), one = CTraits<C>::one();
    for (int r = 0; r < 64; ++r) {
        C v = zero;
        if (r < bs && c < bs) v = LU[(b0 + r) * n + b0 + c];
        D[r][c] = v;
        X[r][c] = zero;
    }
    __syncthreads();
    if (c < bs) {
        if (!upper) {
            X[c][c] = one;
            for (int r = c + 1; r < bs; ++r) {
                C v = zero;
                for (int l = c; l < r; ++l) cfma(v, D[r][l], X[l][c]);
                X[r][c] = -v;
            }
        } else {
            X[c][c] = one / D[c][c];
            for (int r = c - 1; r >= 0; --r) {
                C v = zero;
                for (int l = r + 1; l <= c; ++l) cfma(v, D[r][l], X[l][c]);
                X[r][c] = -(v / D[r][r]);
            }
        }
    }
    __syncthreads();
    for (int r = 0; r < 64; ++r) out[r * 64 + c] = X[r][c];
}

// Inverses of the 64 x 64 diagonal blocks of L (unit lower, which 0) and U (upper,
// which 1) of each factored system: Dinv[((system * 2 + which) * nblk + block) *
// 4096], row-major 64 x 64.
template <class C>
__global__ __launch_bounds__(64) void k_diag_inv(int64_t n, const C *LU0, int64_t sA, C *Dinv) {
    const int64_t nblk = (n + 63) / 64;
    const int64_t b = blockIdx.x, sys = blockIdx.y;
    const int which = blockIdx.z;
    const int64_t b0 = 64 * b;
    const int bs = n - b0 < 64 ? (int)(n - b0) : 64;
    tri_inverse_block<C>(LU0 + sys * sA, n, b0, bs, which == 1,
                         Dinv + ((sys * 2 + which) * nblk + b) * 4096);
}

// Inverses of the unit-lower diagonal blocks of panel [k, k + kb) (one or two
// blocks of SB rows): Linv[(system * 2 + h) * 4096].
template <class C>
__global__ __launch_bounds__(64) void k_panel_linv(int64_t n, const C *A0, int64_t sA, int64_t k,
                                                   int kb, C *Linv) {
    const int h = blockIdx.x;
    const int64_t sys = blockIdx.y;
    const int bs = kb - SB * h < SB ? kb - SB * h : SB;
    tri_inverse_block<C>(A0 + sys * sA, n, k + SB * h, bs, false, Linv + (sys * 2 + h) * 4096);
}

// T = rows [r0, r0 + nr) x columns [c0, n) of the row-major A (nr x (n - c0),
// leading dimension n - c0) per system.
template <class C>
__global__ void k_copy_rows(int64_t n, int64_t r0, int nr, int64_t c0, const C *A0, int64_t sA,
                            C *T0, int64_t sT) {
    const int64_t cols = n - c0, tot = nr * cols;
    const C *A = A0 + (int64_t)blockIdx.y * sA;
    C *T = T0 + (int64_t)blockIdx.y * sT;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t l = t / cols, j = t - l * cols;
        T[t] = A[(r0 + l) * n + c0 + j];
    }
}

// a where ok, else 0, selected per component: a select of the whole struct is
// lowered through a stack slot in scratch memory.
template <class C>
__device__ __forceinline__ C csel(bool ok, const C &a) {
    C r;
    r.x = ok ? a.x : 0;
    r.y = ok ? a.y : 0;
    return r;
}

// One step of the blocked triangular solves L y = b (unit lower, UP false) and
// U x = y (UP true) on a column-major X (n x nrhs per system), in place.  With TR
// the factors are used transposed, for Z^T x = b: U^T y = b (lower, UP false) and
// L^T z = y (unit upper, UP true); their tiles are rows of U and L read as stored.  The
// rows [s0, s0 + ss) are solved already; every other workgroup subtracts their
// contribution from a 64-row tile (forward: rows below, backward: rows above),
// and workgroup 0 owns the diagonal block [d0, d0 + ds), which it then solves
// (one wavefront per right-hand side, lane = row).  Grid: x = tile (0 = the
// diagonal block), y = group of NRG right-hand sides, z = system (sel lists
// them when given).  A pass is 2 + 2 ceil(n / 64) launches, each reading its
// tiles of the factors once.
template <class C, int NRG, bool UP, bool TR>
__global__ __launch_bounds__(256) void k_trsv_step(int64_t n, int nrhs, const C *LU0, int64_t sA,
                                                   C *X0, int64_t sB, const int32_t *sel,
                                                   int64_t s0, int ss, int64_t d0, int ds,
                                                   const C *Dinv0) {
    constexpr int NS = NRG / 4;
    const int64_t sys = sel ? sel[blockIdx.z] : blockIdx.z;
    const C *LU = LU0 + sys * sA;
    const int cg = (int)blockIdx.y * NRG;
    const int nc = nrhs - cg < NRG ? nrhs - cg : NRG;
    C *X = X0 + sys * sB + (int64_t)cg * n;
    __shared__ C Tt[64][65];
    __shared__ C XS[64][NRG];
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    const C zero = CTraits<C>::zero();
    int64_t r0;
    int rows;
    if (blockIdx.x == 0) {
        r0 = d0;
        rows = ds;
    } else if (!UP) {
        r0 = d0 + ds + 64 * (int64_t)(blockIdx.x - 1);
        rows = n - r0 < 64 ? (int)(n - r0) : 64;
    } else {
        r0 = d0 - 64 * (int64_t)blockIdx.x;
        rows = 64;
    }
    C y[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int c = q + 4 * s;
        y[s] = zero;
        if (lane < rows && c < nc) y[s] = X[r0 + lane + n * c];
    }
    if (ss > 0) {
        // factor tile rows [r0, r0 + rows) x columns [s0, s0 + ss) as Tt[column][row].
        // Every lane loads from an in-bounds address (element 0 where masked) so
        // that all loads issue before the first wait: a load under a branch is
        // waited for before its LDS store.
        C v[16];
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            const int rr = q + 4 * p;
            const bool ok = TR ? (rr < ss && lane < rows) : (rr < rows && lane < ss);
            const int64_t at = TR ? (s0 + rr) * n + r0 + lane : (r0 + rr) * n + s0 + lane;
            const C t = LU[ok ? at : 0];
            v[p] = csel(ok, t);
        }
        C xv[NRG / 4];
#pragma unroll
        for (int u = 0; u < NRG / 4; ++u) {
            const int t = tid + 256 * u, l = t % 64, c = t / 64;
            const bool ok = l < ss && c < nc;
            const C t2 = X[ok ? s0 + l + n * c : 0];
            xv[u] = csel(ok, t2);
        }
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            const int rr = q + 4 * p;
            if (TR) Tt[rr][lane] = v[p];
            else Tt[lane][rr] = v[p];
        }
#pragma unroll
        for (int u = 0; u < NRG / 4; ++u) {
            const int t = tid + 256 * u;
            XS[t % 64][t / 64] = xv[u];
        }
        __syncthreads();
        C acc[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[s] = zero;
#pragma unroll 8
        for (int l = 0; l < 64; ++l) {
            const C a = Tt[l][lane];
#pragma unroll
            for (int s = 0; s < NS; ++s) cfma(acc[s], a, XS[l][q + 4 * s]);
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) y[s] = y[s] - acc[s];
    }
    if (blockIdx.x != 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int c = q + 4 * s;
            if (lane < rows && c < nc) X[r0 + lane + n * c] = y[s];
        }
        return;
    }
    // the diagonal block's inverse (k_diag_inv, made for the FP32 factors only):
    // y <- D^{-1} y, a 64 x 64 product per right-hand side through LDS instead of
    // the sequential solve
    if constexpr (sizeof(C) == sizeof(cf)) if (Dinv0) {
        const int64_t nblk = (n + 63) / 64;
        const C *Di = Dinv0 + ((sys * 2 + (UP ? 1 : 0)) * nblk + d0 / 64) * 4096;
        C dv[16];
#pragma unroll
        for (int p = 0; p < 16; ++p) dv[p] = Di[(q + 4 * p) * 64 + lane];
        __syncthreads();
#pragma unroll
        for (int p = 0; p < 16; ++p)
            Tt[lane][q + 4 * p] = dv[p];      // Tt[col][row] = D^{-1}[row][col]
#pragma unroll
        for (int s = 0; s < NS; ++s) XS[lane][q + 4 * s] = y[s];
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NS; ++s) y[s] = zero;
        for (int r = 0; r < ds; ++r) {
            const C dl = Tt[r][lane];
#pragma unroll
            for (int s = 0; s < NS; ++s) cfma(y[s], dl, XS[r][q + 4 * s]);
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int c = q + 4 * s;
            if (lane < ds && c < nc) X[d0 + lane + n * c] = y[s];
        }
        return;
    }
    // the diagonal block as Tt[col][row]; non-unit diagonal for U and U^T
    constexpr bool DIV = UP != TR;
    __syncthreads();
    {
        C v[16];
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            const int rr = q + 4 * p;
            const bool ok = rr < ds && lane < ds;
            const C t = LU[ok ? (d0 + rr) * n + d0 + lane : 0];
            v[p] = csel(ok, t);
        }
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            const int rr = q + 4 * p;
            if (TR) Tt[rr][lane] = v[p];
            else Tt[lane][rr] = v[p];
        }
    }
    __syncthreads();
    if (!UP) {
        for (int r = 0; r < ds; ++r) {
            const C dl = Tt[r][lane];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (DIV && lane == r) y[s] = y[s] / dl;
                const C xr = shfl_c(y[s], r);
                if (lane > r) cfma(y[s], -dl, xr);
            }
        }
    } else {
        for (int r = ds - 1; r >= 0; --r) {
            const C dl = Tt[r][lane];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (DIV && lane == r) y[s] = y[s] / dl;
                const C xr = shfl_c(y[s], r);
                if (lane < r) cfma(y[s], -dl, xr);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int c = q + 4 * s;
        if (lane < ds && c < nc) X[d0 + lane + n * c] = y[s];
    }
}

// R = B - A X for row-major A, column-major X/B/R (n x nrhs <= NR), RW rows per
// wavefront (each X element loaded once serves RW rows), reading A once; also
// max |R| and max |X| per (system, rhs) as monotone bit patterns.  Each row's
// sum has the same order as with one row per wavefront: lane l takes the columns
// j = l mod 64 in turn, then a butterfly over the 64 lanes.  Used for few
// right-hand sides, where the GEMM's 64-column tile would be mostly idle.
template <int NR, int RW>
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
    const int64_t i0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RW;
    if (i0 >= n) return;
    cd acc[RW][NR];
#pragma unroll
    for (int r = 0; r < RW; ++r)
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[r][c] = cd{0.0, 0.0};
    // Every load of a step is issued before the first product: rows past n read
    // row n - 1 and unused right-hand sides read column 0 (their sums are never
    // written), so no load sits under a branch that would wait for it.
    const int64_t ilast = n - 1;
    for (int64_t j = lane; j < n; j += 64) {
        cd xv[NR], av[RW];
#pragma unroll
        for (int c = 0; c < NR; ++c) {
            const cd t = X[j + n * (c < nrhs ? c : 0)];
            xv[c].x = c < nrhs ? t.x : 0.0;
            xv[c].y = c < nrhs ? t.y : 0.0;
        }
#pragma unroll
        for (int r = 0; r < RW; ++r) av[r] = A[min(i0 + r, ilast) * n + j];
#pragma unroll
        for (int r = 0; r < RW; ++r)
#pragma unroll
            for (int c = 0; c < NR; ++c) cfma(acc[r][c], av[r], xv[c]);
    }
#pragma unroll
    for (int r = 0; r < RW; ++r)
#pragma unroll
        for (int c = 0; c < NR; ++c) {
            if (c < nrhs)
                for (int off = 32; off > 0; off >>= 1) {
                    acc[r][c].x += __shfl_xor(acc[r][c].x, off, 64);
                    acc[r][c].y += __shfl_xor(acc[r][c].y, off, 64);
                }
        }
    // lane l writes row i0 + l / NR, right-hand side l % NR
    const int rl = lane / NR, cl = lane % NR;
    if (lane < RW * NR && cl < nrhs && i0 + rl < n) {
        cd a = cd{0.0, 0.0};
#pragma unroll
        for (int r = 0; r < RW; ++r)
#pragma unroll
            for (int c = 0; c < NR; ++c)
                if (r == rl && c == cl) a = acc[r][c];
        const int64_t i = i0 + rl;
        const cd rv = Bv[i + n * cl] - a;
        R[i + n * cl] = rv;
        const double rm = fmax(fabs(rv.x), fabs(rv.y));
        const cd xv = X[i + n * cl];
        const double xm = fmax(fabs(xv.x), fabs(xv.y));
        atomicMax(rmax + b * nrhs + cl, (unsigned long long)__double_as_longlong(rm));
        atomicMax(xmax + b * nrhs + cl, (unsigned long long)__double_as_longlong(xm));
    }
}

// Largest |re| or |im| of each column of R and of X (n x nrhs per system), as
// monotone bit patterns through atomicMax.  Grid: x = row chunk, y = rhs,
// z = system (sel lists them when given).
__global__ __launch_bounds__(256) void k_colmax(int64_t n, int nrhs, const cd *R0, const cd *X0,
                                                int64_t sB, const int32_t *sel,
                                                unsigned long long *rmax,
                                                unsigned long long *xmax) {
    const int64_t b = sel ? sel[blockIdx.z] : blockIdx.z;
    const int c = blockIdx.y;
    const cd *R = R0 + b * sB + (int64_t)c * n;
    const cd *X = X0 + b * sB + (int64_t)c * n;
    double rm = 0, xm = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        rm = fmax(rm, fmax(fabs(R[i].x), fabs(R[i].y)));
        xm = fmax(xm, fmax(fabs(X[i].x), fabs(X[i].y)));
    }
    for (int off = 32; off > 0; off >>= 1) {
        rm = fmax(rm, __shfl_xor(rm, off, 64));
        xm = fmax(xm, __shfl_xor(xm, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(rmax + b * nrhs + c, (unsigned long long)__double_as_longlong(rm));
        atomicMax(xmax + b * nrhs + c, (unsigned long long)__double_as_longlong(xm));
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

// FP32 copy of each row-major n x n system and its infinity norm in one read of
// A: one wavefront per row, the row sums in k_rownorm's order.
__global__ __launch_bounds__(256) void k_cd_to_cf_rownorm(int64_t n, const cd *A0, int64_t sA,
                                                          cf *F0, unsigned long long *out) {
    const int64_t b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const cd *row = A0 + b * sA + i * n;
    cf *frow = F0 + b * sA + i * n;
    double s = 0;
    for (int64_t j = lane; j < n; j += 64) {
        const cd a = row[j];
        frow[j] = cf{(float)a.x, (float)a.y};
        s += cabs(a);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) atomicMax(out + b, (unsigned long long)__double_as_longlong(s));
}

// One norm (largest column sum of |a_ij|) of each row-major n x n system, one
// thread per column (consecutive threads read consecutive columns of a row).
__global__ __launch_bounds__(256) void k_colnorm(int64_t n, const cd *A0, int64_t sA,
                                                 unsigned long long *out) {
    const int64_t b = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const cd *A = A0 + b * sA;
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += cabs(A[i * n + j]);
    atomicMax(out + b, (unsigned long long)__double_as_longlong(s));
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

template <class C, bool TA = false, int TAG = 0>
void gemm_t(int m, int n, int k, C alpha, const C *A, int64_t lda, int64_t sA, const C *B,
            int64_t ldb, int64_t sB, C beta, C *Cm, int64_t ldc, int64_t sC, int batch,
            hipStream_t st, const int32_t *sel = nullptr) {
    if (m <= 0 || n <= 0 || batch <= 0) return;
    dim3 grid((unsigned)((m + 63) / 64), (unsigned)((n + 63) / 64), (unsigned)batch);
    constexpr int WN = sizeof(C) == sizeof(cf) ? 4 : 2;
    hipLaunchKernelGGL((k_gemm_mfma<C, 2, TA, TAG, WN>), grid, dim3(128 * WN), 0, st, m, n, k, alpha, A, lda, sA,
                       B, ldb, sB, beta, Cm, ldc, sC, sel);
    HIPCHECK(hipGetLastError());
}

// Rows held in registers by the leaf kernel: panels of up to this many rows are
// factored in the column-major scratch, taller ones by the streaming k_panel.
constexpr int64_t LEAF_MAX_ROWS = 12288;

// Leaf width for a panel of m rows: as wide as R x LW complex registers allow
// (R = ceil(m / 1024) rows per thread, at most 24 FP64 or 48 FP32 complex, and
// no more than 16 per row: larger register arrays are placed in scratch memory).
// Wider leaves mean fewer rank-LW updates of the sub-panel.
template <class C>
int leaf_width(int64_t m) {
    constexpr bool F = sizeof(C) == sizeof(cf);
    if (m <= 1024) return 16;
    if (m <= 3072) return F ? 16 : 8;
    if (m <= 6144) return F ? 8 : 4;
    return F ? 4 : 2;
}

template <class C>
void launch_leaf(int64_t m, int c0, int lw, int kb, int ce, C *P, int64_t sPt, int32_t *ipiv,
                 int64_t sP, int64_t k, int batch, hipStream_t st) {
    dim3 g((unsigned)batch), b(1024);
#define LF(RR, LL) hipLaunchKernelGGL((k_leaf<C, RR, LL>), g, b, 0, st, m, c0, lw, kb, ce, P, sPt, ipiv, sP, k)
    constexpr bool F = sizeof(C) == sizeof(cf);
    if constexpr (F) {
        if (m <= 1024) LF(1, 16);
        else if (m <= 2048) LF(2, 16);
        else if (m <= 3072) LF(3, 16);
        else if (m <= 5120) LF(5, 8);
        else if (m <= 6144) LF(6, 8);
        else LF(12, 4);
    } else {
        if (m <= 1024) LF(1, 16);
        else if (m <= 2048) LF(2, 8);
        else if (m <= 3072) LF(3, 8);
        else if (m <= 5120) LF(5, 4);
        else if (m <= 6144) LF(6, 4);
        else LF(12, 2);
    }
#undef LF
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

// U12 of panel [k, k + kb) for the columns [k + kb, n), per block of SB rows: the
// inverse of its unit-lower diagonal block times the block's rows (copied to T
// first, one GEMM on the transposed row-major views); between the blocks the
// rank-SB update of the second block's rows.  Linv holds 2 x 4096 and T SB x n
// entries per system.
template <class C>
void panel_linv_t(int64_t n, int64_t k, int kb, const C *A, int64_t sA, int batch, C *Linv,
                  hipStream_t st) {
    const int nh = (kb + SB - 1) / SB;
    hipLaunchKernelGGL(k_panel_linv<C>, dim3((unsigned)nh, (unsigned)batch), dim3(64), 0, st, n, A,
                       sA, k, kb, Linv);
    HIPCHECK(hipGetLastError());
}

// have_linv: Linv already holds the panel's diagonal-block inverses.
template <class C>
void trsm_u12_t(int64_t n, int64_t k, int kb, C *A, int64_t sA, int batch, C *Linv, C *T,
                hipStream_t st, bool have_linv = false) {
    const int64_t c0 = k + kb, cols = n - c0;
    if (cols <= 0) return;
    const int nh = (kb + SB - 1) / SB;
    if (!have_linv) panel_linv_t<C>(n, k, kb, A, sA, batch, Linv, st);
    C mone, one, zero;
    mone.x = -1; mone.y = 0;
    one.x = 1; one.y = 0;
    zero.x = 0; zero.y = 0;
    const int64_t sT = (int64_t)SB * n;
    for (int h = 0; h < nh; ++h) {
        const int64_t r0 = k + SB * h;
        const int hb = std::min(SB, kb - SB * h);
        if (h == 1)
            gemm_t<C>((int)cols, hb, SB, mone, A + k * n + c0, n, sA, A + (k + SB) * n + k, n, sA,
                      one, A + (k + SB) * n + c0, n, sA, batch, st);
        const int64_t tot = (int64_t)hb * cols;
        hipLaunchKernelGGL(k_copy_rows<C>,
                           dim3((unsigned)std::min<int64_t>((tot + 255) / 256, 1024),
                                (unsigned)batch),
                           dim3(256), 0, st, n, r0, hb, c0, A, sA, T, sT);
        HIPCHECK(hipGetLastError());
        gemm_t<C>((int)cols, hb, hb, one, T, cols, sT, Linv + 4096 * h, 64, 2 * 4096, zero,
                  A + r0 * n + c0, n, sA, batch, st);
    }
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
    gemm_t<C, false, 1>((int)ncols, (int)nrows, kw, mone, A + kk * n + r0, n, sA,
                        A + r0 * n + kk, n, sA, one, A + r0 * n + r0, n, sA, batch, st);
}

// A[r0:n, c0:c1] -= A[r0:n, kk:kk+kw] A[kk:kk+kw, c0:c1] (row-major, r0 = kk + kw), any
// column range to the right of the panel; TAG names the launches apart in traces.
template <class C, int TAG>
void update_cols_t(int64_t n, int64_t kk, int kw, int64_t c0, int64_t c1, C *A, int64_t sA,
                   int batch, hipStream_t st) {
    const int64_t r0 = kk + kw;
    const int64_t ncols = c1 - c0, nrows = n - r0;
    if (ncols <= 0 || nrows <= 0) return;
    C mone, one;
    mone.x = -1; mone.y = 0;
    one.x = 1; one.y = 0;
    gemm_t<C, false, TAG>((int)ncols, (int)nrows, kw, mone, A + kk * n + c0, n, sA,
                          A + r0 * n + kk, n, sA, one, A + r0 * n + c0, n, sA, batch, st);
}

// Panel [k, k + kb) of rows [k, n): into the scratch, then per sub-panel of SB
// columns leaves of leaf_width columns (each: factor, its interchanges on the
// other panel columns and U rows on the rest of its sub-panel, rank-lw GEMM
// update of those columns); between the sub-panels the second one's U rows and a
// rank-SB GEMM update; back into A.  Every launch covers the whole batch.
template <class C>
void panel_leaves(int64_t n, int64_t k, int kb, C *A, int64_t sA, int32_t *ipiv, int64_t sP,
                  C *Pt, int64_t sPt, int batch, hipStream_t st) {
    const int64_t m = n - k;
    const dim3 gc((unsigned)((m + 63) / 64), (unsigned)batch);
    hipLaunchKernelGGL(k_panel_copy<C>, gc, dim3(256), 0, st, n, k, kb, A, sA, Pt, sPt, 0);
    HIPCHECK(hipGetLastError());
    const int LW = leaf_width<C>(m);
    C mone, one;
    mone.x = -1; mone.y = 0;
    one.x = 1; one.y = 0;
    for (int sb0 = 0; sb0 < kb; sb0 += SB) {
        const int ce = std::min(kb, sb0 + SB);
        if (sb0 > 0) {
            hipLaunchKernelGGL(k_pt_trsm<C>, dim3((unsigned)batch), dim3(64), 0, st, m, kb, Pt,
                               sPt);
            HIPCHECK(hipGetLastError());
            gemm_t<C>((int)(m - SB), kb - SB, SB, mone, Pt + SB, m, sPt, Pt + (int64_t)SB * m, m,
                      sPt, one, Pt + (int64_t)SB * m + SB, m, sPt, batch, st);
        }
        for (int c0 = sb0; c0 < ce; c0 += LW) {
            const int lw = std::min(LW, ce - c0);
            launch_leaf<C>(m, c0, lw, kb, ce, Pt, sPt, ipiv, sP, k, batch, st);
            const int64_t rr = m - c0 - lw;
            const int rc = ce - c0 - lw;
            if (rr > 0 && rc > 0)
                gemm_t<C>((int)rr, rc, lw, mone, Pt + (int64_t)c0 * m + c0 + lw, m, sPt,
                          Pt + (int64_t)(c0 + lw) * m + c0, m, sPt, one,
                          Pt + (int64_t)(c0 + lw) * m + c0 + lw, m, sPt, batch, st);
        }
    }
    hipLaunchKernelGGL(k_panel_copy<C>, gc, dim3(256), 0, st, n, k, kb, A, sA, Pt, sPt, 1);
    HIPCHECK(hipGetLastError());
}

// Blocked LU of the batch with caller-provided scratch: Pt (n x NB per system, the
// column-major panel copy), T (SB x n per system, U12 block rows) and Linv
// (2 x 64 x 64 per system, the panel's diagonal-block inverses).
template <class C>
void lu_factor_ws_t(int64_t n, int batch, C *A, int64_t sA, int32_t *ipiv, int64_t sP, C *Pt,
                    C *T, C *Linv, hipStream_t st) {
    const int64_t sPt = n * NB;
    for (int64_t k = 0; k < n; k += NB) {
        const int kb = (int)std::min<int64_t>(NB, n - k);
        stage_begin(3, st);
        if (n - k <= LEAF_MAX_ROWS) {
            panel_leaves<C>(n, k, kb, A, sA, ipiv, sP, Pt, sPt, batch, st);
        } else {
            hipLaunchKernelGGL(k_panel<C>, dim3((unsigned)batch), dim3(1024), 0, st, n, k, kb, A,
                               sA, ipiv, sP);
            HIPCHECK(hipGetLastError());
        }
        rowswap_t<C>(n, k, kb, A, sA, ipiv, sP, 0, n, k, k + kb, batch, st);
        trsm_u12_t<C>(n, k, kb, A, sA, batch, Linv, T, st);
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

// The same factorisation with one panel of look-ahead: the panel stream ov.panel
// factors panel k + 1 while the update stream ov.update (kept off the first few
// CUs, where the panel's one-workgroup-per-system leaves run) applies panel k's
// rank-NB update to the columns beyond panel k + 1.  Per step k, on the update
// stream: panel k's interchanges on the columns to its right, U12 of panel k, the
// update of panel k + 1's columns; then on the panel stream: panel k + 1's
// factorisation and its diagonal-block inverses; on the update stream, at the same
// time, the update of the remaining columns.  The interchanges of each panel on
// the columns to its left (the L factor) are applied after the last step, in panel
// order, which leaves the factors LAPACK's getrf produces with the same pivots.
template <class C>
void lu_factor_la_t(int64_t n, int batch, C *A, int64_t sA, int32_t *ipiv, int64_t sP, C *Pt,
                    C *T, C *Linv, hipStream_t st, const LuOverlap &ov) {
    const int64_t sPt = n * NB;
    hipStream_t P = ov.panel, G = ov.update;
    auto record = [](hipEvent_t e, hipStream_t s) { HIPCHECK(hipEventRecord(e, s)); };
    auto wait = [](hipStream_t s, hipEvent_t e) { HIPCHECK(hipStreamWaitEvent(s, e, 0)); };
    auto factor_panel = [&](int64_t k) {
        const int kb = (int)std::min<int64_t>(NB, n - k);
        stage_begin(3, P);
        if (n - k <= LEAF_MAX_ROWS) {
            panel_leaves<C>(n, k, kb, A, sA, ipiv, sP, Pt, sPt, batch, P);
        } else {
            hipLaunchKernelGGL(k_panel<C>, dim3((unsigned)batch), dim3(1024), 0, P, n, k, kb, A,
                               sA, ipiv, sP);
            HIPCHECK(hipGetLastError());
        }
        if (k + kb < n) panel_linv_t<C>(n, k, kb, A, sA, batch, Linv, P);
        stage_end(3, P);
    };
    record(ov.e_start, st);
    wait(P, ov.e_start);
    wait(G, ov.e_start);
    factor_panel(0);
    record(ov.e_panel, P);
    for (int64_t k = 0; k < n; k += NB) {
        const int kb = (int)std::min<int64_t>(NB, n - k);
        const int64_t k1 = k + kb;
        wait(G, ov.e_panel);                         // panel k factored, its inverses ready
        if (k1 >= n) break;
        const int kb1 = (int)std::min<int64_t>(NB, n - k1);
        const int64_t rest = n - k1;
        stage_begin(3, G);
        rowswap_t<C>(n, k, kb, A, sA, ipiv, sP, k1, n, k1, k1, batch, G);
        trsm_u12_t<C>(n, k, kb, A, sA, batch, Linv, T, G, true);
        update_cols_t<C, 2>(n, k, kb, k1, k1 + kb1, A, sA, batch, G);   // panel k + 1's columns
        stage_end(3, G);
        stage_flops(3, 8.0 * (double)rest * kb1 * kb * batch);
        record(ov.e_u12, G);
        // the rest of the trailing matrix, enqueued before the panel's few dozen
        // launches so that the host's time issuing them does not hold it back
        if (k1 + kb1 < n) {
            stage_begin(4, G);
            update_cols_t<C, 1>(n, k, kb, k1 + kb1, n, A, sA, batch, G);
            stage_end(4, G, 8.0 * (double)rest * (double)(rest - kb1) * kb * batch);
        }
        wait(P, ov.e_u12);
        factor_panel(k1);
        record(ov.e_panel, P);
    }
    // every panel's interchanges on the L columns to its left
    for (int64_t k = NB; k < n; k += NB) {
        const int kb = (int)std::min<int64_t>(NB, n - k);
        rowswap_t<C>(n, k, kb, A, sA, ipiv, sP, 0, k, k, k, batch, G);
    }
    record(ov.e_done, G);
    wait(st, ov.e_done);
    stage_flops(7, 8.0 / 3.0 * (double)n * n * n * batch);
}

// ws (optional): batch * lu_factor_ws_elems(n) entries of scratch kept by the caller;
// without it the scratch is allocated here and released once the factorisation
// has run (the host then waits for it).
template <class C>
void lu_factor_t(int64_t n, int batch, C *A, int64_t sA, int32_t *ipiv, int64_t sP,
                 hipStream_t st, const LuOverlap *ov = nullptr, C *ws = nullptr) {
    DBuf<C> own(ws ? 0 : (size_t)batch * lu_factor_ws_elems(n));
    C *Pt = ws ? ws : own.p;
    C *T = Pt + n * NB * batch, *Linv = T + (int64_t)SB * n * batch;
    if (ov && ov->panel && batch <= ov->max_batch && n > NB)
        lu_factor_la_t<C>(n, batch, A, sA, ipiv, sP, Pt, T, Linv, st, *ov);
    else
        lu_factor_ws_t<C>(n, batch, A, sA, ipiv, sP, Pt, T, Linv, st);
    if (!ws) HIPCHECK(hipStreamSynchronize(st));
}

// Z x = b: P B, then L y = b and U x = y: 2 + 2 ceil(n / 64) launches of
// k_trsv_step on the permuted copy in scratch, which is copied back into B.
// TR (Z^T x = b): U^T y = b and L^T z = y on a copy, then x = P^T z into B.
template <class C, bool TR = false>
void lu_solve_t(int64_t n, int nrhs, int batch, const C *LU, int64_t sA, const int32_t *perm,
                int64_t sP, C *B, int64_t sB, C *scratch, hipStream_t st,
                const int32_t *sel = nullptr, const C *dinv = nullptr) {
    if (batch <= 0 || nrhs <= 0) return;
    const int64_t cnt = n * nrhs;
    const dim3 gp((unsigned)std::min<int64_t>((cnt + 255) / 256, 1024), (unsigned)batch);
    if (TR)
        hipLaunchKernelGGL(k_copy_sys<C>, gp, dim3(256), 0, st, cnt, B, scratch, sB, sel);
    else
        hipLaunchKernelGGL(k_perm_rows<C>, gp, dim3(256), 0, st, n, nrhs, perm, sP, B, scratch,
                           sB, sel);
    HIPCHECK(hipGetLastError());
    // right-hand sides per workgroup: each group of them reads the factors once
    const int nrg = nrhs <= 4 ? 4 : (nrhs <= 8 ? 8 : 16);
    const unsigned gy = (unsigned)((nrhs + nrg - 1) / nrg);
    const int64_t nblk = (n + 63) / 64;
    auto step = [&](bool up, int64_t s0, int ss, int64_t d0, int ds, int64_t ntile) {
        const dim3 g((unsigned)(1 + ntile), gy, (unsigned)batch);
#define TS(NRG, UP) hipLaunchKernelGGL((k_trsv_step<C, NRG, UP, TR>), g, dim3(256), 0, st, n, nrhs, LU, sA, scratch, sB, sel, s0, ss, d0, ds, TR ? nullptr : dinv)
        if (nrg == 4) { if (up) TS(4, true); else TS(4, false); }
        else if (nrg == 8) { if (up) TS(8, true); else TS(8, false); }
        else { if (up) TS(16, true); else TS(16, false); }
#undef TS
        HIPCHECK(hipGetLastError());
    };
    // forward: block 0, then each solved block b updates the rows below and
    // workgroup 0 solves block b + 1
    step(false, 0, 0, 0, (int)std::min<int64_t>(64, n), 0);
    for (int64_t b = 0; b + 1 < nblk; ++b) {
        const int64_t d0 = 64 * (b + 1);
        const int ds = (int)std::min<int64_t>(64, n - d0);
        step(false, 64 * b, 64, d0, ds, (n - d0 - ds + 63) / 64);
    }
    // backward: the last block, then each solved block b updates the rows above
    // and workgroup 0 solves block b - 1
    const int64_t dl = 64 * (nblk - 1);
    step(true, 0, 0, dl, (int)(n - dl), 0);
    for (int64_t b = nblk - 1; b >= 1; --b) {
        const int64_t s0 = 64 * b;
        const int64_t d0 = 64 * (b - 1);
        step(true, s0, (int)std::min<int64_t>(64, n - s0), d0, 64, d0 / 64);
    }
    if (TR)
        hipLaunchKernelGGL(k_unperm_rows<C>, gp, dim3(256), 0, st, n, nrhs, perm, sP, scratch, B,
                           sB, sel);
    else
        hipLaunchKernelGGL(k_copy_sys<C>, gp, dim3(256), 0, st, cnt, scratch, B, sB, sel);
    HIPCHECK(hipGetLastError());
}

}  // namespace

void zgemm_batched(int m, int n, int k, cd alpha, const cd *A, int64_t lda, int64_t sA,
                   const cd *B, int64_t ldb, int64_t sB, cd beta, cd *C, int64_t ldc,
                   int64_t sC, int batch, hipStream_t st) {
    gemm_t<cd>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, st);
}

void cgemm_batched(int m, int n, int k, cf alpha, const cf *A, int64_t lda, int64_t sA,
                   const cf *B, int64_t ldb, int64_t sB, cf beta, cf *C, int64_t ldc,
                   int64_t sC, int batch, hipStream_t st) {
    gemm_t<cf>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, st);
}

void lu_factor_batched(int64_t n, int batch, cd *A, int64_t sA, int32_t *ipiv, int64_t sP,
                       hipStream_t st, const LuOverlap *ov, cd *ws) {
    lu_factor_t<cd>(n, batch, A, sA, ipiv, sP, st, ov, ws);
}

int64_t lu_factor_ws_elems(int64_t n) { return n * NB + (int64_t)SB * n + 2 * 4096; }

void lu_factor_batched_ws(int64_t n, int batch, cd *A, int64_t sA, int32_t *ipiv, int64_t sP,
                          cd *ws, hipStream_t st) {
    cd *Pt = ws, *T = ws + n * NB * batch, *Linv = T + (int64_t)SB * n * batch;
    lu_factor_ws_t<cd>(n, batch, A, sA, ipiv, sP, Pt, T, Linv, st);
}

void lu_solve_batched(int64_t n, int nrhs, int batch, const cd *LU, int64_t sA,
                      const int32_t *perm, int64_t sP, cd *B, int64_t sB, cd *scratch,
                      hipStream_t st, const int32_t *sel, bool tr) {
    if (tr) lu_solve_t<cd, true>(n, nrhs, batch, LU, sA, perm, sP, B, sB, scratch, st, sel);
    else lu_solve_t<cd, false>(n, nrhs, batch, LU, sA, perm, sP, B, sB, scratch, st, sel);
}

void lu_make_perm(int64_t n, int batch, const int32_t *ipiv, int64_t sP, int32_t *perm,
                  hipStream_t st) {
    size_t lds = (size_t)n * sizeof(int32_t);
    if (lds > 160 * 1024) throw ArgError("lu_make_perm: n too large for the LDS permutation");
    hipLaunchKernelGGL(k_make_perm, dim3((unsigned)batch), dim3(256), lds, st, n, ipiv, sP, perm);
    HIPCHECK(hipGetLastError());
}

void lu_factor_batched_f32(int64_t n, int batch, cf *A, int64_t sA, int32_t *ipiv, int64_t sP,
                           hipStream_t st, const LuOverlap *ov, cf *ws) {
    lu_factor_t<cf>(n, batch, A, sA, ipiv, sP, st, ov, ws);
}

void lu_solve_batched_f32(int64_t n, int nrhs, int batch, const cf *LU, int64_t sA,
                          const int32_t *perm, int64_t sP, cf *B, int64_t sB, cf *scratch,
                          hipStream_t st, const int32_t *sel, bool tr, const cf *dinv) {
    if (tr) lu_solve_t<cf, true>(n, nrhs, batch, LU, sA, perm, sP, B, sB, scratch, st, sel);
    else lu_solve_t<cf, false>(n, nrhs, batch, LU, sA, perm, sP, B, sB, scratch, st, sel, dinv);
}

void lu_diag_inverses_f32(int64_t n, int batch, const cf *LU, int64_t sA, cf *dinv,
                          hipStream_t st) {
    const int64_t nblk = (n + 63) / 64;
    hipLaunchKernelGGL(k_diag_inv<cf>, dim3((unsigned)nblk, (unsigned)batch, 2), dim3(64), 0, st,
                       n, LU, sA, dinv);
    HIPCHECK(hipGetLastError());
}

void convert_z_to_f32(const cd *Z, cf *Zf, int64_t count, hipStream_t st) {
    hipLaunchKernelGGL(k_cd_to_cf, dim3(grid1(count)), dim3(256), 0, st, Z, Zf, count);
    HIPCHECK(hipGetLastError());
}

void residual_batched(int64_t n, int nrhs, int batch, const cd *A, int64_t sA, const cd *X,
                      const cd *Bv, int64_t sB, cd *R, double *rnorm, double *xnorm,
                      hipStream_t st, const int32_t *sel, int nsel, bool tr) {
    launch_zero((cd *)rnorm, (batch * nrhs + 1) / 2, st);
    launch_zero((cd *)xnorm, (batch * nrhs + 1) / 2, st);
    const int ns = sel ? nsel : batch;
    if (ns <= 0 || nrhs <= 0) return;
    if (tr) {
        // R = B - Z^T X: the row-major Z read as a column-major matrix is Z^T
        const int64_t cnt = n * nrhs;
        const dim3 gp((unsigned)std::min<int64_t>((cnt + 255) / 256, 1024), (unsigned)ns);
        hipLaunchKernelGGL(k_copy_sys<cd>, gp, dim3(256), 0, st, cnt, Bv, R, sB, sel);
        HIPCHECK(hipGetLastError());
        gemm_t<cd, false>((int)n, nrhs, (int)n, cd{-1.0, 0.0}, A, n, sA, X, n, sB, cd{1.0, 0.0},
                          R, n, sB, ns, st, sel);
        const dim3 gm((unsigned)std::min<int64_t>((n + 255) / 256, 16), (unsigned)nrhs,
                      (unsigned)ns);
        hipLaunchKernelGGL(k_colmax, gm, dim3(256), 0, st, n, nrhs, R, X, sB, sel,
                           (unsigned long long *)rnorm, (unsigned long long *)xnorm);
        HIPCHECK(hipGetLastError());
        return;
    }
    if (nrhs <= 8) {
        // four wavefronts per workgroup: 16 rows (nrhs <= 4) or 8 rows per workgroup
        if (nrhs <= 4)
            hipLaunchKernelGGL((k_residual<4, 4>), dim3((unsigned)((n + 15) / 16), (unsigned)ns),
                               dim3(256), 0, st, n, nrhs, A, sA, X, Bv, sB, R,
                               (unsigned long long *)rnorm, (unsigned long long *)xnorm, sel);
        else
            hipLaunchKernelGGL((k_residual<8, 2>), dim3((unsigned)((n + 7) / 8), (unsigned)ns),
                               dim3(256), 0, st, n, nrhs, A, sA, X, Bv, sB, R,
                               (unsigned long long *)rnorm, (unsigned long long *)xnorm, sel);
        HIPCHECK(hipGetLastError());
        return;
    }
    // R = B, then R -= Z X as one GEMM per system (Z row-major: the transposed operand)
    const int64_t cnt = n * nrhs;
    const dim3 gp((unsigned)std::min<int64_t>((cnt + 255) / 256, 1024), (unsigned)ns);
    hipLaunchKernelGGL(k_copy_sys<cd>, gp, dim3(256), 0, st, cnt, Bv, R, sB, sel);
    HIPCHECK(hipGetLastError());
    gemm_t<cd, true>((int)n, nrhs, (int)n, cd{-1.0, 0.0}, A, n, sA, X, n, sB, cd{1.0, 0.0}, R, n,
                     sB, ns, st, sel);
    const dim3 gm((unsigned)std::min<int64_t>((n + 255) / 256, 16), (unsigned)nrhs, (unsigned)ns);
    hipLaunchKernelGGL(k_colmax, gm, dim3(256), 0, st, n, nrhs, R, X, sB, sel,
                       (unsigned long long *)rnorm, (unsigned long long *)xnorm);
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

void matrix_one_norms(int64_t n, int batch, const cd *A, int64_t sA, double *out,
                      hipStream_t st) {
    launch_zero((cd *)out, (batch + 1) / 2, st);
    dim3 g((unsigned)((n + 255) / 256), (unsigned)batch);
    hipLaunchKernelGGL(k_colnorm, g, dim3(256), 0, st, n, A, sA, (unsigned long long *)out);
    HIPCHECK(hipGetLastError());
}

void convert_z_to_f32_inf_norms(int64_t n, int batch, const cd *Z, int64_t sA, cf *Zf,
                                double *out, hipStream_t st) {
    launch_zero((cd *)out, (batch + 1) / 2, st);
    dim3 g((unsigned)((n + 3) / 4), (unsigned)batch);
    hipLaunchKernelGGL(k_cd_to_cf_rownorm, g, dim3(256), 0, st, n, Z, sA, Zf,
                       (unsigned long long *)out);
    HIPCHECK(hipGetLastError());
}

void matrix_inf_norms(int64_t n, int batch, const cd *A, int64_t sA, double *out,
                      hipStream_t st) {
    launch_zero((cd *)out, (batch + 1) / 2, st);
    dim3 g((unsigned)((n + 3) / 4), (unsigned)batch);
    hipLaunchKernelGGL(k_rownorm, g, dim3(256), 0, st, n, A, sA, (unsigned long long *)out);
    HIPCHECK(hipGetLastError());
}
