// Internal declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdexcept>
#include <string>
#include <vector>

#include "cplx.hpp"
#include "dual.hpp"

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHECK(x)                                                                    \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));           \
    } while (0)

// Device buffer with RAII ownership.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        if (hipMalloc((void **)&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw HipError("hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void ensure(size_t count) { if (n < count) alloc(count); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T *h, size_t count, hipStream_t st) {
        ensure(count);
        HIPCHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
    }
};

constexpr int MAXL = 16;        // layers per Gblock region
constexpr int JK_RT = 10;       // rings whose phases are tabulated in LDS
constexpr int JK_RINGMAX = 65;  // src/PGF.jl:15
constexpr int JK_MFMA_MIN = 24;  // phasings from which launch_jk uses the MFMA kernels

// Geometry of a sheet on the device (structure of arrays, 0-based indices).
struct SheetDev {
    int64_t nface, nbf, nufp;
    int identity;               // 1: class i is face pair i
    const double *fxs;          // nface*6: vertex (x,y) x 3, sheet units
    const double *fxm;          // nface*6: same in meters
    const int32_t *fbf;         // nface*3: basis function of local edge, -1 none
    const int32_t *fsg;         // nface*3: +1 plus face, -1 minus face
    const int32_t *feci;        // nface*3: edge cell code 0..4
    const int32_t *fedge;       // nface*3: global edge index
    const double *fr;           // nface
    const int64_t *uoff;        // CSR offsets (nufp+1) or null
    const int64_t *umem;        // CSR members (0-based face pairs)
};

// Faces (match ifm, source ifs) of a face-pair class.  With every face pair its
// own class, classes are numbered row-major (ifm * nface + ifs) so that a
// wavefront's classes share the match face and scatter into one row of Z; with
// translation classes the representative is the first member (ufp2fp order).
__device__ __forceinline__ void class_faces(const SheetDev &sh, int64_t iu, int64_t &ifm,
                                            int64_t &ifs) {
    if (sh.identity) {
        ifm = iu / sh.nface;
        ifs = iu - ifm * sh.nface;
    } else {
        const int64_t fp = sh.umem[sh.uoff[iu]];
        ifm = fp % sh.nface;
        ifs = fp / sh.nface;
    }
}

// Per-point spectral-table parameters.
struct TabParams {
    double k0sq, u, b2min;
    double b00[2], b1[2], b2[2];
    int nl, s, magnetic, pad;
    cd eps[MAXL], mu[MAXL];
    double width[MAXL];
    cd mut, epsb, c3, d3, f1, f3, p1, p3;
    cd dc3, dd3;                // d/dk0 of c3, d3 (electric) or of f3, p3 (magnetic)
};

// Per-point fill parameters.
struct FillParams {
    double k0, u, omega, psi1, psi2;
    double b1[2], b2[2];        // 1/m
    cd mut, epsb, c3, d3;
    cd I1fact, KFfact, I2fact, KPfact;
    cd psf_base;                // electric: im/epsb;  magnetic: 2im/(pi*omega*mu0*mut)
    cd ff[5];
    cd shift9[9];               // cis(-(a psi1 + b psi2)), entry 3 (a + 1) + (b + 1), |a|, |b| <= 1
    int M, pad;
    const cd *tab1, *tab2;      // (M+2)^2 each
    const cd *jk;               // 6*nufp: J, Jxi, Jeta, K, Kxi, Keta (SoA)
    cd *Z;                      // nbf x nbf, row-major with leading dimension ldz
    int64_t ldz;                // 0: nbf
    const cd *dtab1, *dtab2;    // d/dk0 of the tables (k0-derivative fill only)
    cd dc3, dd3;                // d/dk0 of c3, d3 (electric) or KFfact, KPfact (magnetic)
};

// ---- pgf.hip
void launch_summands(const TabParams *d_tp, int npts, int64_t cap_entries, int r0, int r1,
                     cd *t1, cd *t2, hipStream_t st);
// d/dk0 of the summands at fixed phasing and u, rings r0..r1.
void launch_summands_dk0(const TabParams *d_tp, int npts, int64_t cap_entries, int r0, int r1,
                         cd *t1, cd *t2, hipStream_t st);
void launch_ringsums(const TabParams *d_tp, int npts, int64_t cap_entries, int r0, int r1,
                     int rcap, const cd *t1, const cd *t2, double convtest, int *flags,
                     hipStream_t st);
void launch_gather_square(int npts, const int *d_M, int Mmax, int64_t cap_entries,
                          const cd *t1, const cd *t2, cd *S, int64_t s_stride,
                          hipStream_t st);
void launch_dft_matrix(int M, cd *W, hipStream_t st);
void launch_phase_ghost(int npts, const int *d_M, const double *d_psi, const int *d_mag,
                        const double *d_area, const cd *F, int64_t f_stride, cd *tab,
                        int64_t tab_stride, int Mmax, hipStream_t st);
// total (optional, device): adds the lattice points the reference's per-pair
// ring rule sums (tiled / MFMA paths only; untouched for fewer phasings).
void launch_jk(const SheetDev &sh, int64_t nufp, double ulocal, const double *us1,
               const double *us2, const double *d_psi, int nf, cd *jk, int64_t jk_stride,
               int mixed, hipStream_t st, unsigned long long *total = nullptr);
void launch_singular(const SheetDev &sh, int64_t nufp, double ulocal, double *rho_r,
                     double *rinv, hipStream_t st);

// ---- fill.hip
void launch_zero(cd *p, int64_t n, hipStream_t st);
// Phase 1: G[p] holds 9 planes of nufp class terms (plane 3*iml + isl).
void launch_fill_classes(const SheetDev &sh, const FillParams *d_fp, int npts,
                         const double *rho_r, const double *rinv, int magnetic, cd *G,
                         int64_t gstride, hipStream_t st);
// d/dk0 of the class terms of one point (P.dtab1, P.dtab2, P.dc3, P.dd3 set).
void launch_fill_classes_dk0(const SheetDev &sh, const FillParams *d_fp, const double *rho_r,
                             const double *rinv, int magnetic, cd *G, hipStream_t st);
// Rows and columns [n, ld) of npts row-major ld x ld blocks (stride sZ) set to the
// identity's, so that a system of n unknowns padded to ld factors as itself.
void launch_pad_identity(cd *Z, int64_t n, int64_t ld, int64_t sZ, int npts, hipStream_t st);
// Phase 2: Z of each point from G (fpcls: class of face pair ifm + nface*ifs, CSR only).
void launch_fill_gather(const SheetDev &sh, const FillParams *d_fp, int npts,
                        const int32_t *bf_face, const int32_t *bf_loc, const int32_t *fpcls,
                        const cd *G, int64_t gstride, int has_fr, int magnetic, hipStream_t st);

// ---- linalg.hip
void zgemm_batched(int m, int n, int k, cd alpha, const cd *A, int64_t lda, int64_t sA,
                   const cd *B, int64_t ldb, int64_t sB, cd beta, cd *C, int64_t ldc,
                   int64_t sC, int batch, hipStream_t st);
void cgemm_batched(int m, int n, int k, cf alpha, const cf *A, int64_t lda, int64_t sA,
                   const cf *B, int64_t ldb, int64_t sB, cf beta, cf *C, int64_t ldc,
                   int64_t sC, int batch, hipStream_t st);
// Streams and events of the look-ahead factorisation (owned by the caller):
// `update` is kept off a few CUs, where the panel's one-workgroup-per-system leaves
// on `panel` find room; batches above max_batch (more leaves than those CUs) use
// the single-stream factorisation.
struct LuOverlap {
    hipStream_t panel = nullptr, update = nullptr;
    hipEvent_t e_start = nullptr, e_u12 = nullptr, e_panel = nullptr, e_done = nullptr;
    int max_batch = 0;
};
// ws (optional): batch * lu_factor_ws_elems(n) entries of caller-kept scratch (the
// call then neither allocates nor waits); without it the scratch is allocated and
// released inside, after a wait for the factorisation.
void lu_factor_batched(int64_t n, int batch, cd *A, int64_t sA, int32_t *ipiv, int64_t sP,
                       hipStream_t st, const LuOverlap *ov = nullptr, cd *ws = nullptr);
// Scratch entries per system of lu_factor_batched_ws, and the same factorisation
// (single stream) with that scratch passed in (ws holds batch * lu_factor_ws_elems(n)
// entries; nothing is allocated).
int64_t lu_factor_ws_elems(int64_t n);
void lu_factor_batched_ws(int64_t n, int batch, cd *A, int64_t sA, int32_t *ipiv, int64_t sP,
                          cd *ws, hipStream_t st);
// perm from lu_make_perm; scratch holds batch * sB entries.  sel (optional): the
// batch lists system indices, one per entry.  tr: solve Z^T x = b instead.
void lu_solve_batched(int64_t n, int nrhs, int batch, const cd *LU, int64_t sA,
                      const int32_t *perm, int64_t sP, cd *B, int64_t sB, cd *scratch,
                      hipStream_t st, const int32_t *sel = nullptr, bool tr = false);
void lu_make_perm(int64_t n, int batch, const int32_t *ipiv, int64_t sP, int32_t *perm,
                  hipStream_t st);
void lu_factor_batched_f32(int64_t n, int batch, cf *A, int64_t sA, int32_t *ipiv,
                           int64_t sP, hipStream_t st, const LuOverlap *ov = nullptr,
                           cf *ws = nullptr);
// sel (optional): the batch lists system indices, one per entry.  tr: Z^T x = b.
// dinv (optional, not with tr): the diagonal-block inverses of lu_diag_inverses_f32,
// used in place of the sequential diagonal solves.
void lu_solve_batched_f32(int64_t n, int nrhs, int batch, const cf *LU, int64_t sA,
                          const int32_t *perm, int64_t sP, cf *B, int64_t sB, cf *scratch,
                          hipStream_t st, const int32_t *sel = nullptr, bool tr = false,
                          const cf *dinv = nullptr);
// Inverses of the 64 x 64 diagonal blocks of L and U for each of the batch systems:
// dinv holds batch * 2 * ceil(n / 64) * 4096 entries.
void lu_diag_inverses_f32(int64_t n, int batch, const cf *LU, int64_t sA, cf *dinv,
                          hipStream_t st);
void convert_z_to_f32(const cd *Z, cf *Zf, int64_t count, hipStream_t st);
// Zf = Z in FP32 and out[b] = infinity norm of system b (as matrix_inf_norms), in
// one read of Z; Zf has Z's layout (row-major n x n, stride sA).
void convert_z_to_f32_inf_norms(int64_t n, int batch, const cd *Z, int64_t sA, cf *Zf,
                                double *out, hipStream_t st);
// r = b - A x (FP64; a row-per-wavefront kernel for up to 8 right-hand sides, else
// one transposed-A GEMM per system), batched; then the largest
// |re| or |im| of each column of r and of x, at rnorm/xnorm[system * nrhs + rhs].
// With sel, only the nsel listed systems are computed (norms of the others are 0).
// tr: r = b - A^T x (one GEMM per system).
void residual_batched(int64_t n, int nrhs, int batch, const cd *A, int64_t sA, const cd *X,
                      const cd *Bv, int64_t sB, cd *R, double *rnorm, double *xnorm,
                      hipStream_t st, const int32_t *sel = nullptr, int nsel = 0,
                      bool tr = false);
void residual_to_f32(const cd *R, cf *Rf, int64_t count, hipStream_t st);
// out[b] = one norm of system b (out holds 2 * ceil(batch / 2) doubles).
void matrix_one_norms(int64_t n, int batch, const cd *A, int64_t sA, double *out,
                      hipStream_t st);
// out[b] = infinity norm of system b (out holds 2 * ceil(batch / 2) doubles).
void matrix_inf_norms(int64_t n, int batch, const cd *A, int64_t sA, double *out,
                      hipStream_t st);
void add_correction(cd *X, const cf *D, int64_t count, hipStream_t st);
// X[sys] += D[sys] for the nsel systems listed in sel, each of sB entries.
void add_correction_sel(cd *X, const cf *D, int64_t sB, const int32_t *sel, int nsel,
                        hipStream_t st);

// ---- gsmk.hip
struct BfftJob {                // one Fourier transform: wavevector and phases
    double k[2];
    cd ff[5];
};
void launch_bfft(const SheetDev &sh, const int32_t *bf_face, const int32_t *bf_loc,
                 const BfftJob *d_jobs, int njobs, cd *ft, hipStream_t st);
void launch_rhs(int64_t nbf, int ncol, const cd *ft, const int32_t *d_col_job,
                const cd *d_col_src, cd *rhs, int64_t ld, hipStream_t st);
void launch_project(int64_t nbf, int ncol, int njobs, const cd *X, int64_t ldx, const cd *ft,
                    cd *out, hipStream_t st);
void cascade_batched_dev(int nb, int n1, int n2, int n3, const cd *a, const cd *b, cd *c,
                         hipStream_t st);

// ---- stage timing (HIP events on the launching stream)
constexpr int NSTAGE = 8;
struct StageRecorder {
    struct Rec {
        int stage;
        hipEvent_t a, b;
        double flops;
    };
    std::vector<Rec> open, pending;
    std::vector<hipEvent_t> pool;
    double ms[NSTAGE] = {0}, flops[NSTAGE] = {0};
    int64_t launches[NSTAGE] = {0};
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHECK(hipEventCreate(&e));
        return e;
    }
    void begin(int stage, hipStream_t st) {
        Rec r{stage, get(), get(), 0.0};
        HIPCHECK(hipEventRecord(r.a, st));
        open.push_back(r);
    }
    void end(int stage, hipStream_t st, double fl) {
        for (size_t i = open.size(); i-- > 0;) {
            if (open[i].stage == stage) {
                Rec r = open[i];
                open.erase(open.begin() + i);
                r.flops = fl;
                HIPCHECK(hipEventRecord(r.b, st));
                pending.push_back(r);
                return;
            }
        }
    }
    // wait = false: only the records whose end event has completed (the host does
    // not wait for work it left running, e.g. spatial sums started ahead of a call)
    void flush(bool wait = true) {
        std::vector<Rec> keep;
        for (Rec &r : pending) {
            if (!wait) {
                const hipError_t q = hipEventQuery(r.b);
                if (q == hipErrorNotReady) {
                    keep.push_back(r);
                    continue;
                }
                HIPCHECK(q);
            }
            HIPCHECK(hipEventSynchronize(r.b));
            float t = 0;
            HIPCHECK(hipEventElapsedTime(&t, r.a, r.b));
            ms[r.stage] += t;
            flops[r.stage] += r.flops;
            launches[r.stage] += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.swap(keep);
    }
    void reset() {
        for (int i = 0; i < NSTAGE; ++i) { ms[i] = 0; flops[i] = 0; launches[i] = 0; }
    }
    ~StageRecorder() {
        for (auto &r : pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        for (auto &r : open) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};
extern thread_local StageRecorder *g_stage_rec;
inline void stage_begin(int s, hipStream_t st) {
    if (g_stage_rec) g_stage_rec->begin(s, st);
}
inline void stage_end(int s, hipStream_t st, double flops = 0.0) {
    if (g_stage_rec) g_stage_rec->end(s, st, flops);
}
inline void stage_flops(int s, double flops) {
    if (g_stage_rec) g_stage_rec->flops[s] += flops;
}

// ---- adjoint.hip (forward-mode directions through the fill; transposed solves)
// ftan_s / ftan_m: 6 per face, derivative of the face's vertex coordinates along
// the direction (sheet units / metres).
// Tangent spatial sums (J, J_xi, J_eta, K, K_xi, K_eta as duals) along nd <= 4
// directions in one pass: ftan and out hold nd blocks.
void launch_jk_tan_multi(const SheetDev &sh, double ul, const double *us1, const double *us2,
                         double psi1, double psi2, const double *ftan_s, int64_t ftan_stride,
                         int nd, cdd *out, int64_t out_stride, hipStream_t st);
void launch_singular_tan(const SheetDev &sh, double ul, const double *ftan_s, dd *rho_r,
                         dd *rinv, hipStream_t st);
// out[o] = sum_{m,s} dZ_ms sum_c A[o*sAo + m + nbf*c] X[s + nbf*c]; part: one slot per
// (256-class block, objective).  mag selects the 'M' class terms; has_fr adds the
// surface-resistance self terms ('J').  Identity face-pair classes only.
void launch_class_tan(const SheetDev &sh, const FillParams *d_fp, const cdd *jkd,
                      const dd *rho_r, const dd *rinv, const double *ftan_m, const cd *A,
                      int64_t sAo, const cd *X, int ncol, int nobj, int mag, int has_fr,
                      cd *part, cd *out, hipStream_t st);
void launch_bfft_tan(const SheetDev &sh, const int32_t *bf_face, const int32_t *bf_loc,
                     const BfftJob *d_jobs, int njobs, const double *ftan_m, cd *dft,
                     hipStream_t st);
void launch_obs(int64_t nbf, int ncol, const cd *ft, const int32_t *d_col_job,
                const double *d_col_t, cd *g, int64_t ld, hipStream_t st);

// ---- ufp.hip
// Face-pair classes of a mesh (rho: 2*nnode, fv: 3*nface 1-based, device arrays).
// Writes ufpm (nface^2, 1-based class of pair m + nface*n), members (nface^2,
// ascending within classes) and offsets (nufp + 1, capacity nface^2 + 1); returns nufp.
int64_t face_pair_classes_dev(int64_t nnode, const double *d_rho, int64_t nface,
                              const int64_t *d_fv, double radius, int64_t *d_ufpm,
                              int64_t *d_members, uint64_t *d_offsets_u64, hipStream_t st);
