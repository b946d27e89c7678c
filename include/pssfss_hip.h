/*
 * libpssfss_hip — MI355X (gfx950) implementation of PSSFSS.jl's per-frequency
 * method-of-moments path behind a plain C ABI.
 *
 * Every entry point returns 0 on success or a negative PSSFSS_E* status and never
 * throws; pssfss_last_error() describes the last failure of a context.  All
 * arrays are host memory in the reference's own layout: complex values are
 * interleaved (re, im) doubles (Julia ComplexF64), matrices are column-major,
 * index arrays are 1-based int64 (Julia Int).  Calls are synchronous.  A context
 * is not thread-safe: use one context per calling thread.
 *
 * Reference: PSSFSS.jl v0.1.4 (src/ paths below).
 */
#ifndef PSSFSS_HIP_H
#define PSSFSS_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PSSFSS_OK = 0,
    PSSFSS_EINVAL = -1,   /* bad argument or inconsistent mesh data */
    PSSFSS_EHIP = -2,     /* HIP runtime failure */
    PSSFSS_ENOMEM = -3,   /* device allocation failed */
    PSSFSS_ECONV = -4     /* a convergence test failed (see info flags) */
};

enum { PSSFSS_FP64 = 0, PSSFSS_MIXED = 1 };

typedef struct pssfss_ctx pssfss_ctx;
typedef struct pssfss_sheet pssfss_sheet;

/* Triangulated sheet plus its RWG maps: the RWGSheet fields of src/Sheets.jl:19-58
 * and the RWGData fields of src/RWG.jl:27-76 that the fill reads. */
typedef struct {
    int64_t nnode, nface, nedge, nbf;
    const double *rho;          /* 2*nnode: node (x,y), sheet length units           */
    const int64_t *fv;          /* 3*nface column-major: vertices of each face       */
    const int64_t *fe;          /* 3*nface: edges of each face, fe[i] opposite fv[i] */
    const double *fr;           /* nface: sheet resistance (Ohm/sq); NULL = PEC      */
    const int64_t *bff;         /* 2*nbf: plus / minus face of each basis function  */
    const int64_t *ebf;         /* nedge: basis function of each edge, 0 = none     */
    const int64_t *eci;         /* nedge: unit-cell boundary code 0..4              */
    int64_t nufp;               /* number of unique-face-pair classes                */
    const int64_t *ufp_offsets; /* nufp+1 0-based CSR offsets; NULL: each face pair
                                   is its own class (nufp == nface^2)                */
    const int64_t *ufp_members; /* face-pair indices ifm + nface*(ifs-1), each class
                                   ascending, representative first (ufp2fp)          */
    double s1[2], s2[2];        /* direct lattice vectors (sheet units)              */
    double beta1[2], beta2[2];  /* reciprocal lattice vectors (1/sheet unit)         */
    double units_per_meter;     /* sheet units per meter                             */
    char sheet_class;           /* 'J' (electric currents) or 'M' (magnetic)        */
} pssfss_sheet_desc;

/* One analysis point of one sheet: the arguments calculate_jtype_gsm /
 * calculate_mtype_gsm receive or compute before the fill (src/PSSFSS.jl:362-395,
 * 486-519), i.e. after gsm_electric_gblock / gsm_magnetic_gblock. */
typedef struct {
    double k0;                  /* free-space wavenumber (rad/m)                     */
    double u;                   /* Green's function smoothing parameter (1/m)        */
    double psi1, psi2;          /* unit-cell incremental phase shifts (rad)          */
    int32_t nl;                 /* layers of the Gblock region                        */
    int32_t s;                  /* 1-based interface of the sheet within the region  */
    const double *epsr;         /* 2*nl complex relative permittivities               */
    const double *mur;          /* 2*nl complex relative permeabilities               */
    const double *width;        /* nl layer widths (m)                                */
    int32_t nmodes[2];          /* modes in the first and last region layer           */
    const double *beta[2];      /* 2*nmodes: modal transverse wavevectors (1/m)      */
    const double *tvec[2];      /* 2*nmodes: modal unit field vectors                 */
    const double *c[2];         /* 2*nmodes complex: modal normalisation constants    */
    const double *Y[2];         /* 2*nmodes complex: modal admittances ('M' sheets)   */
    int32_t nmax;               /* leading dimension of vinc / tlgf                   */
    const double *vinc;         /* (nmax,2) complex: vincs ('J') or iincs ('M')       */
    const double *tlgf;         /* (nmax,2) complex: tlgfvi ('J') or tlgfiv ('M')     */
    double acf[2];              /* area correction factors of the two end regions     */
    double area;                /* unit cell area (m^2)                               */
} pssfss_point;

/* GSM blocks of the Gblock (src/GSMs.jl:16-29), column-major complex, holding the
 * incident-field partial GSM on entry; the sheet's scattered part is added. */
typedef struct {
    double *s11, *s12, *s21, *s22;
} pssfss_gsm;

typedef struct {
    int32_t mmax;               /* spectral table size used                           */
    int32_t table_converged;    /* 0: table convergence loop hit mmax = 2048          */
    int32_t jk_rings_max;       /* largest spatial-sum ring count over the sheet      */
    int32_t refine_iters;       /* iterative-refinement sweeps (PSSFSS_MIXED)         */
    int32_t fallback_fp64;      /* 1: refinement stalled, point re-solved in FP64     */
    int32_t pad;
    double refine_resid;        /* best normwise backward error of the mixed solves,
                                   max_c |r_c| / (sqrt(n) |x_c| |A| eps); converged
                                   below 1000 for the sheet GSM calls (their Z holds
                                   FP32-summed spatial sums), below 1 for
                                   pssfss_lu_solve_batch                               */
} pssfss_point_info;

/* ---- context ---------------------------------------------------------------- */
int pssfss_ctx_create(int device, pssfss_ctx **out);
/* Releases the context's stream and workspaces and every sheet still uploaded to it. */
int pssfss_ctx_destroy(pssfss_ctx *ctx);
const char *pssfss_last_error(const pssfss_ctx *ctx);
/* Device name, CU count and HBM size, for logs. */
int pssfss_device_info(pssfss_ctx *ctx, char *name, int name_len, int *num_cu,
                       double *hbm_gib);

/* Chunking of the batched sheet calls (no counterpart in the reference): at most
 * max_points points per chunk (0: 128), and a chunk's workspaces within
 * mem_fraction of the free HBM (0: 0.45).  Several contexts driven concurrently
 * on one GPU (uneven problems side by side) each take a share. */
int pssfss_ctx_set_chunk(pssfss_ctx *ctx, int max_points, double mem_fraction);

/* ---- sheets ----------------------------------------------------------------- */
/* Upload a sheet; replaces setup_rwg's in-memory RWGData for the device
 * (src/RWG.jl:90-324 builds the maps; src/PSSFSS.jl:220 calls it). */
int pssfss_sheet_upload(pssfss_ctx *ctx, const pssfss_sheet_desc *desc, pssfss_sheet **out);
/* Frees a sheet of this context (its device mesh and J/K cache); PSSFSS_EINVAL for a
 * sheet that is not live in the context.  NULL is a no-op. */
int pssfss_sheet_free(pssfss_ctx *ctx, pssfss_sheet *sheet);

/* Unique-face-pair classes of setup_rwg (src/RWG.jl:270-322) on the device, for
 * the sheet upload's ufp_offsets / ufp_members.  Pair index m + nface*(n-1)
 * (1-based faces m, n as the reference), stored 0-based: mn = m0 + nface*n0.
 * rho: 2*nnode, fv: 3*nface (1-based, column-major as RWGSheet.fv).  On return
 * *nufp classes; offsets (capacity nface^2 + 1) gets nufp + 1 CSR offsets,
 * members (nface^2) the pairs of each class ascending, ufpm (nface^2, may be
 * NULL) the 1-based class of each pair (RWGData.ufpm).  Classes are numbered by
 * their smallest member, as the reference's loop numbers them. */
int pssfss_face_pair_classes(pssfss_ctx *ctx, int64_t nnode, const double *rho,
                             int64_t nface, const int64_t *fv, int64_t *nufp, int64_t *offsets,
                             int64_t *members, int64_t *ufpm);

/* ---- the hot path ------------------------------------------------------------ */
/* Sheet GSM at one point: fill -> LU -> RHS -> solve -> extract.
 * Replaces calculate_jtype_gsm (src/PSSFSS.jl:362-460) and calculate_mtype_gsm
 * (src/PSSFSS.jl:486-587) from the fillz/filly call on; the sheet keeps the
 * frequency-independent J/K integrals keyed by (psi1, psi2, u) exactly as
 * src/FillZY.jl:97-108 does.  info may be NULL. */
int pssfss_sheet_gsm(pssfss_ctx *ctx, pssfss_sheet *sheet, const pssfss_point *pt,
                     pssfss_gsm *gsm, int precision, pssfss_point_info *info);

/* Batched form: npts points of one sheet (any frequencies / phasings) filled,
 * factored and solved together; the sweep loop of src/PSSFSS.jl:230-305 calls it
 * once per sheet instead of calculate_*_gsm once per point. */
int pssfss_sheet_gsm_batch(pssfss_ctx *ctx, pssfss_sheet *sheet, int npts,
                           const pssfss_point *pts, pssfss_gsm *gsms, int precision,
                           pssfss_point_info *infos);

/* Starts, without waiting, the spatial sums and singular integrals that a following
 * pssfss_sheet_gsm_batch call over points with these keys (psi1, psi2 in radians, u
 * in 1/m, as in pssfss_point) will read, so that the host can prepare those points
 * meanwhile.  Keys follow the reference's cache rule (src/FillZY.jl:97-108: phasings
 * equal, u within 1e-4 relative).  When every point reuses the first point's key
 * (e.g. normal incidence) the sums go to the sheet's J/K cache.  Otherwise, when the
 * points fit one chunk of the GSM call, the integrals of every new key go to the
 * context's slots, and the GSM call over the same points (same sheet, precision and
 * cached key) uses them instead of summing again; a later prefetch or GSM call
 * replaces them.  Results are identical either way.  No counterpart in the
 * reference, which sums them inside fillz at the first point of a key. */
int pssfss_sheet_prefetch_jk(pssfss_ctx *ctx, pssfss_sheet *sheet, int npts, const double *psi1,
                             const double *psi2, const double *u, int precision);

/* Points of several sheets (independent problems, or several sheets of one stack)
 * in shared batches: sheet g contributes npts[g] consecutive entries of pts / gsms /
 * infos.  Every point must have the same end-region mode counts.  Each batch is
 * filled per sheet, then factored and solved as one set of systems padded to the
 * batch's largest unknown count (identity block below and right of each sheet's
 * matrix, zero right-hand-side rows), which leaves each sheet's solution what its
 * own batch would give.  Replaces the reference's problem-by-problem, point-by-point
 * loop (src/PSSFSS.jl:230-305) when many structures of uneven size are analysed.
 * infos may be NULL. */
int pssfss_sheets_gsm_batch(pssfss_ctx *ctx, int nsheet, pssfss_sheet *const *sheets,
                            const int *npts, const pssfss_point *pts, pssfss_gsm *gsms,
                            int precision, pssfss_point_info *infos);

/* ---- sensitivities (no counterpart in the reference) ---------------------------
 * Derivatives of the sheet GSM with respect to mesh geometry, for gradient-based
 * design through a ChainRulesCore.rrule on the S-parameter output.  For each of
 * npts points (as pssfss_sheet_gsm_batch) and nobj weight sets, returns for every
 * direction k
 *     grad[p][o][k] = sum_e wbar[p][o][e] * d S_e / d t_k   (complex)
 * where e runs over the entries of the Gblock GSM packed as in pssfss_gsm
 * (s11, s12, s21, s22 blocks, column-major complex, (n1+n2)^2 entries) and the
 * mesh nodes move as rho(t) = rho + t_k vtan[k] (vtan: nparam blocks of nnode
 * (x, y) pairs, sheet units).  An objective f(S) enters through wbar = df/dS (holomorphic
 * derivative); Re(grad) is then the derivative of a real objective.  The forward
 * solve is repeated in `precision` (PSSFSS_FP64, or PSSFSS_MIXED: FP32 LU with FP64
 * refinement); the adjoint systems Z^T lambda = g use the same factors (refined in
 * FP64 against Z in the mixed form); the fill is recomputed with derivatives and
 * contracted, Z's derivative is never stored.  'J' sheets (surface resistance
 * included) and 'M' sheets uploaded without face-pair classes (every pair its own
 * class; a sheet with ufp classes is refused).  grad: npts x nobj x nparam
 * complex. */
int pssfss_sheet_adjoint(pssfss_ctx *ctx, pssfss_sheet *sheet, int npts,
                         const pssfss_point *pts, int nparam, const double *vtan, int nobj,
                         const double *wbar, double *grad, int precision);

/* ---- single stages, each a parity boundary ---------------------------------- */
/* fillz (src/FillZY.jl:50-245) / filly (src/FillZY.jl:273-457): the nbf x nbf
 * impedance (admittance) matrix, written to zmat (complex, column-major). */
int pssfss_fill(pssfss_ctx *ctx, pssfss_sheet *sheet, const pssfss_point *pt, double *zmat,
                pssfss_point_info *info);

/* d/dk0 of the fill (no counterpart in the reference): Z as pssfss_fill and its
 * derivative with respect to the free-space wavenumber pt->k0 (1/m) at fixed
 * phasing (pt->psi1, pt->psi2) and fixed u, both nbf x nbf complex column-major.
 * The spectral tables are differentiated summand by summand (same table size M);
 * the spatial sums and singular integrals do not depend on k0.  Both sheet
 * classes ('J': dZ/dk0 of the impedance matrix, 'M': of the admittance matrix). */
int pssfss_fill_dk0(pssfss_ctx *ctx, pssfss_sheet *sheet, const pssfss_point *pt,
                    double *zmat, double *dzmat, pssfss_point_info *info);

/* filljk! (src/Zint.jl:54-179): jk = (6, nufp) complex (J, J_xi, J_eta, K, K_xi,
 * K_eta per class), rho_r = (2, nufp), rinv = nufp.  u_local in 1/sheet unit. */
int pssfss_filljk(pssfss_ctx *ctx, pssfss_sheet *sheet, double u_local, double psi1,
                  double psi2, double *jk, double *rho_r, double *rinv);

/* electric_modal_sum_funcs (src/PGF.jl:196-365, magnetic = 0) and
 * magnetic_modal_sum_funcs (src/PGF.jl:478-629, magnetic = 1): the interpolation
 * tables over -1..mmax in both indices, (mmax+2)^2 complex each, column-major.
 * beta1/beta2/beta00 in 1/m.  tab1/tab2 must hold 2050^2 complex values. */
int pssfss_modal_tables(pssfss_ctx *ctx, int magnetic, const pssfss_point *pt,
                        const double *beta1, const double *beta2, const double *beta00,
                        double convtest, int *mmax, int *converged, double *tab1,
                        double *tab2);

/* rwgbfft! (src/RWG.jl:428-512): ft = (2, nbf) complex at wavevector k (1/m). */
int pssfss_rwgbfft(pssfss_ctx *ctx, pssfss_sheet *sheet, const double *k, double psi1,
                   double psi2, double *ft);

/* lu! + ldiv! (src/PSSFSS.jl:401,439): solves A X = B for nbatch independent
 * n x n systems.  A: nbatch column-major complex matrices (not modified);
 * B: nbatch (n, nrhs) right-hand sides, overwritten by X.  PSSFSS_MIXED factors
 * in FP32 and refines in FP64.  iters (nbatch ints) may be NULL. */
int pssfss_lu_solve_batch(pssfss_ctx *ctx, int64_t n, int nrhs, int nbatch, const double *A,
                          double *B, int precision, int *iters, double *resid);

/* FP64 lu! + ldiv! for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops): no counterpart in the reference; the
 * allocation-free variant of pssfss_lu_solve_batch.  A and B are DEVICE pointers:
 * A holds nbatch column-major n x n complex matrices and is overwritten by LU
 * factors; B holds nbatch (n, nrhs) right-hand sides and is overwritten by X.
 * Pivots, permutations and all other working memory come from one fixed 64 MiB
 * device array per GPU, kept for the life of the process: a call needs about
 * (3080 + 16 nrhs) n + 131072 bytes per system (nbatch x n up to about 20000 at
 * nrhs = 8); larger calls fail with PSSFSS_EINVAL.  Nothing in that array needs to
 * start at zero.  Every call on one GPU shares the array: calls that overlap in
 * time on the same GPU (two threads, two contexts) race and must be serialised
 * by the caller. */
int pssfss_lu_solve_batch_noalloc(pssfss_ctx *ctx, int64_t n, int nrhs, int nbatch, double *A,
                                  double *B);

/* cascade (src/GSMs.jl:125-141) for nbatch independent GSM pairs: a has port
 * sizes (n1, n2), b has (n2, n3).  Each GSM is packed as s11, s12, s21, s22
 * consecutively, column-major complex.  c receives (n1, n3) GSMs. */
int pssfss_cascade_batch(pssfss_ctx *ctx, int nbatch, int n1, int n2, int n3,
                         const double *a, const double *b, double *c);

/* ---- measurement (no counterpart in the reference) ---------------------------
 * Stage timing with HIP events recorded on the context's stream around every
 * launch of a stage.  Stages: 0 spatial sums, 1 spectral tables, 2 fill,
 * 3 LU panels + interchanges + U12 solves, 4 LU trailing GEMM updates,
 * 5 triangular solves and refinement, 6 transforms, RHS and projection.
 * flops[] holds algorithmic flop counts: stage 0, 8 per lattice point that the
 * reference's per-(class, quadrature point, phasing) ring rule sums (counted by
 * the many-phasing kernels, 24 or more phasings per launch); stage 4, 8 m n k
 * per complex GEMM update; flops[7], 8/3 n^3 per LU (whole factorisation). */
typedef struct {
    double ms[8];
    double flops[8];
    int64_t launches[8];
} pssfss_stats;

int pssfss_stats_enable(pssfss_ctx *ctx, int on);
int pssfss_stats_read(pssfss_ctx *ctx, pssfss_stats *out, int reset);

#ifdef __cplusplus
}
#endif
#endif /* PSSFSS_HIP_H */
