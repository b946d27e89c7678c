// C ABI of libpssfss_hip: contexts, sheets and the per-point MoM pipeline
//   J/K spatial integrals (cached by (psi1, psi2, u))  -> spectral tables ->
//   fill -> LU (FP64, or FP32 + FP64 refinement) -> RHS -> solve -> projection
// for one or many points of a sheet.  Host code here only orchestrates and does
// the O(n_modes^2) GSM bookkeeping of src/PSSFSS.jl:431-456 / 557-584.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <set>
#include <vector>

#include "../../include/pssfss_hip.h"
#include "internal.hpp"

namespace {

constexpr double TWOPI = 6.283185307179586;
constexpr double PI_ = 3.141592653589793;
constexpr double TOL_U = 1e-4;       // src/Constants.jl:19 (J/K cache key on u)
constexpr int CONVREPEAT = 40;       // src/PGF.jl:217
constexpr int MMAX_LAST = 2048;      // src/PGF.jl:18
constexpr double CONVTEST_FILL = 1e-7;   // src/FillZY.jl:69,293

// c3 / d3 wide-band coefficients, eqs. (4.30)/(4.33), src/PGF.jl:138-162.
cd c3_calc(double k0, double u, cd mu1, cd eps1, cd mu2, cd eps2) {
    cd w1 = ((k0 * k0) * mu1) * eps1 + u * u;
    cd w2 = ((k0 * k0) * mu2) * eps2 + u * u;
    return (mu1 * w2 + mu2 * w1) / (2.0 * (mu1 + mu2));
}
cd d3_calc(double k0, double u, cd mu1, cd eps1, cd mu2, cd eps2) {
    cd w1 = ((k0 * k0) * mu1) * eps1 + u * u;
    cd w2 = ((k0 * k0) * mu2) * eps2 + u * u;
    cd num = mu1 * (w2 * (2.0 * eps1 + eps2) - w1 * eps1) +
             mu2 * (w1 * (2.0 * eps2 + eps1) - w2 * eps2);
    return num / ((2.0 * (mu1 + mu2)) * (eps1 + eps2));
}

// d/dk0 of c3_calc / d3_calc: both are linear in w_i = k0^2 mu_i eps_i + u^2, so
// the derivative is the same formula with dw_i/dk0 = 2 k0 mu_i eps_i in place of w_i.
cd c3_dk0(double k0, cd mu1, cd eps1, cd mu2, cd eps2) {
    cd w1 = ((2.0 * k0) * mu1) * eps1;
    cd w2 = ((2.0 * k0) * mu2) * eps2;
    return (mu1 * w2 + mu2 * w1) / (2.0 * (mu1 + mu2));
}
cd d3_dk0(double k0, cd mu1, cd eps1, cd mu2, cd eps2) {
    cd w1 = ((2.0 * k0) * mu1) * eps1;
    cd w2 = ((2.0 * k0) * mu2) * eps2;
    cd num = mu1 * (w2 * (2.0 * eps1 + eps2) - w1 * eps1) +
             mu2 * (w1 * (2.0 * eps2 + eps1) - w2 * eps2);
    return num / ((2.0 * (mu1 + mu2)) * (eps1 + eps2));
}

cd hcis(double t) { return cd{cos(t), sin(t)}; }

// The mmax sequence of the table convergence loop (src/PGF.jl:223-233):
// 0, then successive even 5-smooth numbers, ending at 2048.
const std::vector<int> &mmax_sequence() {
    // built once, under the language's thread-safe initialisation of local statics
    static const std::vector<int> seq = [] {
        std::vector<int> q;
        auto next235 = [](int n) {
            for (int v = n;; ++v) {
                int x = v;
                while (x % 2 == 0) x /= 2;
                while (x % 3 == 0) x /= 3;
                while (x % 5 == 0) x /= 5;
                if (x == 1) return v;
            }
        };
        int M = 0;
        q.push_back(0);
        while (M < MMAX_LAST) {
            M = next235(1 + M);
            while (M % 2 != 0) M = next235(1 + M);
            q.push_back(M);
        }
        return q;
    }();
    return seq;
}

bool approx2(const double *a, const double *b) {
    double d = hypot(a[0] - b[0], a[1] - b[1]);
    double na = hypot(a[0], a[1]), nb = hypot(b[0], b[1]);
    return d <= 1.4901161193847656e-08 * std::max(na, nb);
}

}  // namespace

thread_local StageRecorder *g_stage_rec = nullptr;

// Device memory of pssfss_lu_solve_batch_noalloc (pivots, row permutations, the LU's
// panel scratch and the solve's staging), one copy per GPU for the whole process.
// Host code never uses the name `scratch` as an address: each device's copy is
// reached through the pointer hipGetSymbolAddress returns with that device current.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_addr;   // device ordinal -> its copy of `scratch`

// Address of `scratch` on `device`, looked up once per device; the device must be
// current.  nullptr when the lookup fails.
void *scratch_address(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) return it->second;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_scratch_addr[device] = ptr;
    return ptr;
}
}  // namespace

struct pssfss_ctx {
    int device = 0;
    void *scratch_dev = nullptr;     // this device's copy of `scratch`
    bool stats_on = false;
    StageRecorder rec;
    hipStream_t st = nullptr;
    std::string err;
    std::map<int, DBuf<cd>> dft;     // DFT matrices by size
    // reusable workspaces
    DBuf<cd> wZ, wB, wX, wR, wft, wout;
    DBuf<cd> wjk;                // spatial sums of the phasing keys of one fill call
    DBuf<cf> wZf, wRf;
    DBuf<int32_t> wpiv, wperm, wsel;
    DBuf<cd> wT;
    DBuf<cf> wTf;
    DBuf<cf> wDinv;              // inverses of the FP32 factors' diagonal blocks
    DBuf<double> wnorm;
    DBuf<cd> wG;                 // fill class terms
    DBuf<double> wanorm;         // infinity norms of the systems (refinement)
    DBuf<cd> wlu;                // LU panel scratch (FP64), kept between calls
    DBuf<cf> wluf;               // the same for the FP32 factorisation
    // Spectral tables are built on st_aux while the spatial sums run on st; their
    // buffers are kept between calls, so that building them frees nothing (hipFree
    // waits for the whole device, which would end the overlap).
    hipStream_t st_aux = nullptr;
    DBuf<TabParams> wtp;
    DBuf<cd> wt1, wt2, wtS, wtA, wtF;
    DBuf<int> wtflags, wtM, wtmag;
    DBuf<double> wtpsi, wtarea;
    DBuf<double> wjkpsi;                 // phasings of the spatial-sum launches
    DBuf<unsigned long long> wjkcnt;     // their lattice-point count (stats on)
    bool jk_cnt_pending = false;         // wjkcnt not read back yet
    // spatial sums of a sweep's keys started ahead in wjk (pssfss_sheet_prefetch_jk):
    // the sheet, precision, plan and the sheet's cached key they were planned from
    struct Prefetched {
        bool valid = false;
        const pssfss_sheet *sheet = nullptr;
        int mixed = 0;
        std::vector<double> keys;
        bool uses_old = false;
        bool c_valid = false, c_mixed = false;
        double c_psi1 = 0, c_psi2 = 0, c_u = 0;
    } pf;
    std::set<pssfss_sheet *> sheets;   // sheets uploaded and not yet freed
    LuOverlap lu_ov;             // look-ahead LU streams (panel / update CU subsets)
    int chunk_max = 128;         // points per chunk of a batched sheet call
    double mem_frac = 0.45;      // share of free HBM a chunk's workspaces may take
    // built on stream s on first use; callers on another stream order after s
    const cd *dft_matrix(int M, hipStream_t s) {
        auto it = dft.find(M);
        if (it != dft.end()) return it->second.p;
        DBuf<cd> &w = dft[M];
        w.alloc((size_t)M * M);
        launch_dft_matrix(M, w.p, s);
        return w.p;
    }
};

struct pssfss_sheet {
    int64_t nnode = 0, nface = 0, nedge = 0, nbf = 0, nufp = 0;
    bool identity = true;
    char cls = 'J';
    double upm = 1.0;
    double s1[2] = {0, 0}, s2[2] = {0, 0}, b1[2] = {0, 0}, b2[2] = {0, 0};
    DBuf<double> fxs, fxm, fr;
    DBuf<int32_t> fbf, fsg, feci, fedge, bf_face, bf_loc;
    DBuf<int32_t> fpcls;        // class of face pair ifm + nface*ifs (translation classes)
    std::vector<int32_t> hfv;   // 3*nface 0-based face vertices (adjoint directions)
    bool has_fr = false;        // any face with nonzero sheet resistance
    DBuf<int64_t> uoff, umem;
    // J/K cache with the key semantics of src/FillZY.jl:97-108, plus the precision
    // the integrals were summed in: FP32-summed rings (jk_mixed) serve only mixed
    // consumers, FP64 integrals serve both
    bool jk_valid = false;
    bool jk_mixed = false;
    double jk_psi1 = 0, jk_psi2 = 0, jk_u = 0;
    DBuf<cd> jk;
    // singular integrals (depend on u only)
    bool sing_valid = false;
    double sing_u = 0;
    DBuf<double> rho_r, rinv;

    SheetDev dev() const {
        SheetDev d;
        d.nface = nface; d.nbf = nbf; d.nufp = nufp;
        d.identity = identity ? 1 : 0;
        d.fxs = fxs.p; d.fxm = fxm.p;
        d.fbf = fbf.p; d.fsg = fsg.p; d.feci = feci.p; d.fedge = fedge.p;
        d.fr = fr.p;
        d.uoff = uoff.p; d.umem = umem.p;
        return d;
    }
};

namespace {

template <class F>
int guarded(pssfss_ctx *ctx, F &&f) {
    if (!ctx) return PSSFSS_EINVAL;
    struct RecScope {
        RecScope(pssfss_ctx *c) { g_stage_rec = c->stats_on ? &c->rec : nullptr; }
        ~RecScope() { g_stage_rec = nullptr; }
    } scope(ctx);
    try {
        f();
        if (ctx->stats_on) ctx->rec.flush(false);
        ctx->err.clear();
        return PSSFSS_OK;
    } catch (const ArgError &e) {
        ctx->err = e.what();
        return PSSFSS_EINVAL;
    } catch (const HipError &e) {
        ctx->err = e.what();
        return strstr(e.what(), "hipMalloc") ? PSSFSS_ENOMEM : PSSFSS_EHIP;
    } catch (const std::bad_alloc &) {
        ctx->err = "host allocation failed";
        return PSSFSS_ENOMEM;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return PSSFSS_EINVAL;
    }
}

void sync(pssfss_ctx *ctx) { HIPCHECK(hipStreamSynchronize(ctx->st)); }

template <class T>
void d2h(pssfss_ctx *ctx, T *h, const T *d, size_t n) {
    if (n == 0) return;
    HIPCHECK(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, ctx->st));
}

cd lay(const double *a, int i) { return cd{a[2 * i], a[2 * i + 1]}; }

void check_point(const pssfss_sheet *sh, const pssfss_point *pt, bool need_modes) {
    if (!pt) throw ArgError("null point");
    if (pt->nl < 2 || pt->nl > MAXL) throw ArgError("point: nl must be in 2..16");
    if (pt->s < 1 || pt->s >= pt->nl) throw ArgError("point: s must be in 1..nl-1");
    if (!pt->epsr || !pt->mur || !pt->width) throw ArgError("point: null layer arrays");
    if (!(pt->k0 > 0) || !(pt->u > 0)) throw ArgError("point: k0 and u must be positive");
    if (need_modes) {
        for (int r = 0; r < 2; ++r) {
            if (pt->nmodes[r] < 0 || pt->nmodes[r] > pt->nmax)
                throw ArgError("point: nmodes out of range");
            if (pt->nmodes[r] > 0 && (!pt->beta[r] || !pt->tvec[r] || !pt->c[r]))
                throw ArgError("point: null mode arrays");
            if (sh->cls == 'M' && pt->nmodes[r] > 0 && !pt->Y[r])
                throw ArgError("point: class 'M' needs modal admittances");
        }
        if (!pt->vinc || !pt->tlgf) throw ArgError("point: null vinc/tlgf");
    }
}

// ------------------------------------------------------------ spectral tables
struct TabInput {
    TabParams tp;
    double psi1, psi2, area;
};

TabInput make_tab_input(const pssfss_point *pt, bool magnetic, const double *b1,
                        const double *b2, const double *b00) {
    TabInput in;
    TabParams &c = in.tp;
    memset(&c, 0, sizeof(c));
    c.k0sq = pt->k0 * pt->k0;
    c.u = pt->u;
    c.b2min = 1e-10 * c.k0sq;
    c.b00[0] = b00[0]; c.b00[1] = b00[1];
    c.b1[0] = b1[0]; c.b1[1] = b1[1]; c.b2[0] = b2[0]; c.b2[1] = b2[1];
    c.nl = pt->nl;
    c.s = pt->s;
    c.magnetic = magnetic ? 1 : 0;
    for (int i = 0; i < pt->nl; ++i) {
        c.eps[i] = lay(pt->epsr, i);
        c.mu[i] = lay(pt->mur, i);
        c.width[i] = pt->width[i];
    }
    const int s = pt->s;
    cd ms = c.mu[s - 1], ms1 = c.mu[s], es = c.eps[s - 1], es1 = c.eps[s];
    c.mut = ((2.0 * ms) * ms1) / (ms + ms1);
    c.epsb = (es + es1) / 2.0;
    const double k0 = pt->k0, u = pt->u;
    if (!magnetic) {
        c.c3 = c3_calc(k0, u, ms, es, ms1, es1);
        c.d3 = d3_calc(k0, u, ms, es, ms1, es1);
        c.dc3 = c3_dk0(k0, ms, es, ms1, es1);
        c.dd3 = d3_dk0(k0, ms, es, ms1, es1);
    } else {
        cd c3s = c3_calc(k0, u, es, ms, es, ms), c3sp1 = c3_calc(k0, u, es1, ms1, es1, ms1);
        cd d3s = d3_calc(k0, u, es, ms, es, ms), d3sp1 = d3_calc(k0, u, es1, ms1, es1, ms1);
        c.f1 = 2.0 * c.epsb;
        c.f3 = es * c3s + es1 * c3sp1;
        c.p1 = 2.0 / c.mut;
        c.p3 = d3s / ms + d3sp1 / ms1;
        c.dc3 = es * c3_dk0(k0, es, ms, es, ms) + es1 * c3_dk0(k0, es1, ms1, es1, ms1);
        c.dd3 = d3_dk0(k0, es, ms, es, ms) / ms + d3_dk0(k0, es1, ms1, es1, ms1) / ms1;
    }
    in.psi1 = pt->psi1;
    in.psi2 = pt->psi2;
    in.area = (4.0 * (PI_ * PI_)) / fabs(b1[0] * b2[1] - b1[1] * b2[0]);
    return in;
}

TabInput sheet_tab_input(const pssfss_sheet *sh, const pssfss_point *pt) {
    const double upm = sh->upm;
    double b1[2] = {sh->b1[0] * upm, sh->b1[1] * upm}, b2[2] = {sh->b2[0] * upm, sh->b2[1] * upm};
    double b00[2] = {(pt->psi1 * b1[0] + pt->psi2 * b2[0]) / TWOPI,
                     (pt->psi1 * b1[1] + pt->psi2 * b2[1]) / TWOPI};
    return make_tab_input(pt, sh->cls == 'M', b1, b2, b00);
}

struct TableSet {
    int npts = 0, Mmax = 0;
    std::vector<int> M, converged;
    DBuf<cd> tabs;          // per point: two tables of (M+2)^2
    int64_t tab_stride = 0; // complex entries per point
    const cd *tab1(int p) const { return tabs.p + p * tab_stride; }
    const cd *tab2(int p) const {
        int64_t E = (int64_t)M[p] + 2;
        return tabs.p + p * tab_stride + E * E;
    }
};

void finish_tables(pssfss_ctx *ctx, const std::vector<TabInput> &in, int64_t cap,
                   const cd *t1, const cd *t2, TableSet &out, hipStream_t st);

// Tables of several points: summands ring by ring until the convergence test of
// src/PGF.jl:291-298 passes at an mmax of the loop's sequence (rings below 40
// can never pass it, so the first test is at mmax = 80), then the 2-D DFT as two
// GEMMs, the phase of eq. (5.32) and the ghost rows.  Everything runs on stream st
// and the host waits for st alone; the buffers are the context's (ctx->wt*).
void build_tables(pssfss_ctx *ctx, const std::vector<TabInput> &in, double convtest,
                  TableSet &out, hipStream_t st) {
    const int np = (int)in.size();
    const auto &seq = mmax_sequence();
    std::vector<TabParams> tp(np);
    for (int p = 0; p < np; ++p) tp[p] = in[p].tp;
    DBuf<TabParams> &d_tp = ctx->wtp;
    d_tp.upload(tp.data(), np, st);
    int rcap = 64;
    int64_t cap = (int64_t)(2 * rcap + 1) * (2 * rcap + 1);
    DBuf<cd> &t1 = ctx->wt1, &t2 = ctx->wt2;
    t1.ensure((size_t)cap * np);
    t2.ensure((size_t)cap * np);
    const int rstride = MMAX_LAST / 2 + 1;
    DBuf<int> &dflags = ctx->wtflags;
    dflags.ensure((size_t)rstride * np);
    std::vector<int> hflags((size_t)rstride * np, 0);
    std::vector<size_t> cand(np);
    std::vector<int> decided(np, 0);
    out.M.assign(np, 0);
    out.converged.assign(np, 0);
    size_t first = 0;
    while (seq[first] < 2 * CONVREPEAT) ++first;
    for (int p = 0; p < np; ++p) cand[p] = first;
    int computed = -1;
    for (;;) {
        int need = computed;
        for (int p = 0; p < np; ++p)
            if (!decided[p]) need = std::max(need, seq[cand[p]] / 2);
        if (need > computed) {
            if (need > rcap) {
                int nr = rcap;
                while (nr < need) nr *= 2;
                nr = std::min(nr, MMAX_LAST / 2);
                int64_t ncap = (int64_t)(2 * nr + 1) * (2 * nr + 1);
                DBuf<cd> n1((size_t)ncap * np), n2((size_t)ncap * np);
                if (computed >= 0) {
                    int64_t used = (int64_t)(2 * computed + 1) * (2 * computed + 1);
                    for (int p = 0; p < np; ++p) {
                        HIPCHECK(hipMemcpyAsync(n1.p + p * ncap, t1.p + p * cap,
                                                used * sizeof(cd), hipMemcpyDeviceToDevice, st));
                        HIPCHECK(hipMemcpyAsync(n2.p + p * ncap, t2.p + p * cap,
                                                used * sizeof(cd), hipMemcpyDeviceToDevice, st));
                    }
                }
                HIPCHECK(hipStreamSynchronize(st));
                t1 = std::move(n1);
                t2 = std::move(n2);
                rcap = nr;
                cap = ncap;
            }
            stage_begin(1, st);
            launch_summands(d_tp.p, np, cap, computed + 1, need, t1.p, t2.p, st);
            launch_ringsums(d_tp.p, np, cap, computed + 1, need, rstride, t1.p, t2.p, convtest,
                            dflags.p, st);
            stage_end(1, st);
            HIPCHECK(hipMemcpyAsync(hflags.data(), dflags.p, hflags.size() * sizeof(int),
                                    hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            computed = need;
        }
        bool all = true;
        for (int p = 0; p < np; ++p) {
            while (!decided[p] && seq[cand[p]] / 2 <= computed) {
                int M = seq[cand[p]], mo2 = M / 2;
                bool ok = true;
                for (int r = mo2 - CONVREPEAT + 1; r <= mo2 && ok; ++r)
                    ok = hflags[(size_t)p * rstride + r] != 0;
                if (ok || M >= MMAX_LAST) {
                    decided[p] = 1;
                    out.M[p] = M;
                    out.converged[p] = ok ? 1 : 0;
                } else {
                    ++cand[p];
                }
            }
            if (!decided[p]) all = false;
        }
        if (all) break;
    }
    finish_tables(ctx, in, cap, t1.p, t2.p, out, st);
}

// Tables from ring-major summands (rings 0..M/2 of each point; out.M set): the
// square gather, the 2-D DFT as two GEMMs, the phase of eq. (5.32) and the ghost rows.
void finish_tables(pssfss_ctx *ctx, const std::vector<TabInput> &in, int64_t cap,
                   const cd *t1, const cd *t2, TableSet &out, hipStream_t st) {
    const int np = (int)in.size();
    int Mmax = 0;
    for (int p = 0; p < np; ++p) Mmax = std::max(Mmax, out.M[p]);
    out.npts = np;
    out.Mmax = Mmax;
    const int64_t sq = (int64_t)Mmax * 2 * Mmax;
    DBuf<cd> &S = ctx->wtS, &A = ctx->wtA, &F = ctx->wtF;
    S.ensure((size_t)sq * np);
    A.ensure((size_t)sq * np);
    F.ensure((size_t)sq * np);
    DBuf<int> &dM = ctx->wtM, &dmag = ctx->wtmag;
    std::vector<int> hmag(np);
    std::vector<double> hpsi(2 * np), harea(np);
    for (int p = 0; p < np; ++p) {
        hmag[p] = in[p].tp.magnetic;
        hpsi[2 * p] = in[p].psi1;
        hpsi[2 * p + 1] = in[p].psi2;
        harea[p] = in[p].area;
    }
    DBuf<double> &dpsi = ctx->wtpsi, &darea = ctx->wtarea;
    dM.upload(out.M.data(), np, st);
    dmag.upload(hmag.data(), np, st);
    dpsi.upload(hpsi.data(), 2 * np, st);
    darea.upload(harea.data(), np, st);
    stage_begin(1, st);
    launch_gather_square(np, dM.p, Mmax, cap, t1, t2, S.p, sq, st);
    const cd one = {1.0, 0.0}, zero = {0.0, 0.0};
    for (int p = 0; p < np; ++p) {
        const int M = out.M[p];
        const cd *W = ctx->dft_matrix(M, st);
        zgemm_batched(M, 2 * M, M, one, W, M, 0, S.p + p * sq, M, 0, zero, A.p + p * sq, M, 0, 1,
                      st);
        zgemm_batched(M, M, M, one, A.p + p * sq, M, (int64_t)M * M, W, M, 0, zero, F.p + p * sq,
                      M, (int64_t)M * M, 2, st);
    }
    out.tab_stride = 2 * (int64_t)(Mmax + 2) * (Mmax + 2);
    out.tabs.alloc((size_t)out.tab_stride * np);
    launch_phase_ghost(np, dM.p, dpsi.p, dmag.p, darea.p, F.p, sq, out.tabs.p, out.tab_stride,
                       Mmax, st);
    stage_end(1, st);
    HIPCHECK(hipStreamSynchronize(st));   // the tables are complete for any stream
}

// d/dk0 of the tables `val` of the points `in` at fixed phasing and u:
// the derivative summands over the rings the tables use, then the same linear
// steps as the values, with the same M per point.
void build_table_dk0(pssfss_ctx *ctx, const std::vector<TabInput> &in, const TableSet &val,
                     TableSet &out) {
    hipStream_t st = ctx->st;
    const int np = (int)in.size();
    std::vector<TabParams> tp(np);
    int Mmax = 0;
    for (int p = 0; p < np; ++p) {
        tp[p] = in[p].tp;
        Mmax = std::max(Mmax, val.M[p]);
    }
    DBuf<TabParams> d_tp(np);
    d_tp.upload(tp.data(), np, st);
    const int R = Mmax / 2;
    const int64_t cap = (int64_t)(2 * R + 1) * (2 * R + 1);
    DBuf<cd> t1((size_t)cap * np), t2((size_t)cap * np);
    launch_summands_dk0(d_tp.p, np, cap, 0, R, t1.p, t2.p, st);
    out.M = val.M;
    out.converged = val.converged;
    finish_tables(ctx, in, cap, t1.p, t2.p, out, st);
}

// --------------------------------------------------------- J/K spatial sums
void ensure_singular(pssfss_ctx *ctx, pssfss_sheet *sh, double ul) {
    if (sh->sing_valid && sh->sing_u == ul) return;
    sh->rho_r.ensure(2 * sh->nufp);
    sh->rinv.ensure(sh->nufp);
    launch_singular(sh->dev(), sh->nufp, ul, sh->rho_r.p, sh->rinv.p, ctx->st);
    sh->sing_valid = true;
    sh->sing_u = ul;
}

// J/K integrals of nk phasings (all with the same u_local) into dst + i*6*nufp.
void compute_jk(pssfss_ctx *ctx, pssfss_sheet *sh, double ul, const std::vector<double> &psis,
                cd *dst, int mixed) {
    const int nk = (int)psis.size() / 2;
    double us1[2] = {ul * sh->s1[0], ul * sh->s1[1]};
    double us2[2] = {ul * sh->s2[0], ul * sh->s2[1]};
    // the host does not wait: the phasings sit in a context buffer (the next upload
    // is ordered after this launch on the same stream)
    DBuf<double> &dpsi = ctx->wjkpsi;
    dpsi.upload(psis.data(), psis.size(), ctx->st);
    const int64_t stride = 6 * sh->nufp;
    const int per = 1024;      // phasings per launch (grid y covers them in tiles)
    // with stats on: lattice points of the reference's sums, 8 flops each (J and K,
    // real term times complex phase), accumulated until flush_jk_count reads them
    unsigned long long *cnt = nullptr;
    if (g_stage_rec) {
        ctx->wjkcnt.ensure(2);                               // 16 bytes: cleared as one cd
        if (!ctx->jk_cnt_pending) launch_zero((cd *)ctx->wjkcnt.p, 1, ctx->st);
        ctx->jk_cnt_pending = true;
        cnt = ctx->wjkcnt.p;
    }
    for (int g = 0; g < nk; g += per) {
        int nf = std::min(per, nk - g);
        stage_begin(0, ctx->st);
        launch_jk(sh->dev(), sh->nufp, ul, us1, us2, dpsi.p + 2 * g, nf, dst + g * stride, stride,
                  mixed, ctx->st, cnt);
        stage_end(0, ctx->st);
    }
}

// The lattice-point count of the spatial sums launched since the last call (stats
// on), read back after the context's stream has drained.
void flush_jk_count(pssfss_ctx *ctx) {
    if (!ctx->jk_cnt_pending) return;
    ctx->jk_cnt_pending = false;
    unsigned long long npts = 0;
    HIPCHECK(hipMemcpyAsync(&npts, ctx->wjkcnt.p, sizeof(npts), hipMemcpyDeviceToHost, ctx->st));
    HIPCHECK(hipStreamSynchronize(ctx->st));
    stage_flops(0, 8.0 * (double)npts);
}

bool key_matches(const pssfss_sheet *sh, bool valid, double k1, double k2, double ku,
                 double psi1, double psi2, double ul) {
    (void)sh;
    if (!valid) return false;
    if (psi1 != k1 || psi2 != k2 || ku == 0) return false;
    return !(fabs((ul - ku) / ku) > TOL_U);
}

// The spatial-sum keys of a chunk's points under the sequential cache of
// src/FillZY.jl:97-108: a point reuses the most recent integrals when its (psi, u)
// pass the test of src/FillZY.jl:97-98, otherwise its key is new.  Starts from the
// sheet's cached key (an FP64 consumer treats FP32-summed integrals as a miss).
struct JkPlan {
    std::vector<int> slot;       // per point: new-key index, -1 = the sheet's cached integrals
    std::vector<double> keys;    // psi1, psi2, u per new key
    bool uses_old = false;
};

JkPlan plan_jk(const pssfss_sheet *sh, int np, const double *psi1, const double *psi2,
               const double *ul, int mixed) {
    JkPlan pl;
    pl.slot.resize(np);
    bool valid = sh->jk_valid && (mixed || !sh->jk_mixed);
    double k1 = sh->jk_psi1, k2 = sh->jk_psi2, ku = sh->jk_u;
    int cur = -1;
    for (int p = 0; p < np; ++p) {
        if (key_matches(sh, valid, k1, k2, ku, psi1[p], psi2[p], ul[p])) {
            pl.slot[p] = cur;
            if (cur < 0) pl.uses_old = true;
        } else {
            pl.keys.push_back(psi1[p]);
            pl.keys.push_back(psi2[p]);
            pl.keys.push_back(ul[p]);
            cur = (int)pl.keys.size() / 3 - 1;
            valid = true;
            k1 = psi1[p]; k2 = psi2[p]; ku = ul[p];
            pl.slot[p] = cur;
        }
    }
    return pl;
}

// The integrals of a plan's new keys into slots (key i at i * 6 nufp), the sheet's
// cached integrals copied behind them when some points reuse those.  Nothing waits.
void compute_plan(pssfss_ctx *ctx, pssfss_sheet *sh, const JkPlan &pl, int mixed,
                  DBuf<cd> &slots) {
    const int nnew = (int)pl.keys.size() / 3;
    const int64_t stride = 6 * sh->nufp;
    slots.ensure((size_t)stride * (nnew + (pl.uses_old ? 1 : 0)));
    if (pl.uses_old)   // the cache is overwritten afterwards; keep the old integrals
        HIPCHECK(hipMemcpyAsync(slots.p + nnew * stride, sh->jk.p, stride * sizeof(cd),
                                hipMemcpyDeviceToDevice, ctx->st));
    int g0 = 0;     // consecutive keys with equal u share a launch
    while (g0 < nnew) {
        int g1 = g0;
        std::vector<double> ps;
        while (g1 < nnew && pl.keys[3 * g1 + 2] == pl.keys[3 * g0 + 2]) {
            ps.push_back(pl.keys[3 * g1]);
            ps.push_back(pl.keys[3 * g1 + 1]);
            ++g1;
        }
        compute_jk(ctx, sh, pl.keys[3 * g0 + 2], ps, slots.p + g0 * stride, mixed);
        g0 = g1;
    }
}

bool prefetched_matches(const pssfss_ctx *ctx, const pssfss_sheet *sh, const JkPlan &pl,
                        int mixed) {
    const auto &f = ctx->pf;
    return f.valid && f.sheet == sh && f.mixed == mixed && f.keys == pl.keys &&
           f.uses_old == pl.uses_old && f.c_valid == sh->jk_valid && f.c_mixed == sh->jk_mixed &&
           f.c_psi1 == sh->jk_psi1 && f.c_psi2 == sh->jk_psi2 && f.c_u == sh->jk_u;
}

// Resolve the J/K data every point of a chunk uses (plan_jk), computing the new
// keys' integrals into slots unless pssfss_sheet_prefetch_jk already started that
// very plan there.  Returns per-point device pointers; the sheet's cache ends
// holding the last key.
std::vector<const cd *> resolve_jk(pssfss_ctx *ctx, pssfss_sheet *sh, const pssfss_point *pts,
                                   int np, int mixed, DBuf<cd> &slots, double &key_u) {
    std::vector<double> p1(np), p2(np), ul(np);
    for (int p = 0; p < np; ++p) {
        p1[p] = pts[p].psi1;
        p2[p] = pts[p].psi2;
        ul[p] = pts[p].u / sh->upm;
    }
    const JkPlan pl = plan_jk(sh, np, p1.data(), p2.data(), ul.data(), mixed);
    key_u = pl.slot[0] < 0 ? sh->jk_u : pl.keys[3 * pl.slot[0] + 2];
    const int nnew = (int)pl.keys.size() / 3;
    const int64_t stride = 6 * sh->nufp;
    std::vector<const cd *> out(np);
    if (nnew == 0) {
        for (int p = 0; p < np; ++p) out[p] = sh->jk.p;
        return out;
    }
    const bool pre = &slots == &ctx->wjk && prefetched_matches(ctx, sh, pl, mixed);
    ctx->pf.valid = false;   // slots are (re)written or consumed here
    if (!pre) compute_plan(ctx, sh, pl, mixed, slots);
    for (int p = 0; p < np; ++p)
        out[p] = pl.slot[p] < 0 ? slots.p + nnew * stride : slots.p + pl.slot[p] * stride;
    sh->jk.ensure(stride);
    HIPCHECK(hipMemcpyAsync(sh->jk.p, slots.p + (nnew - 1) * stride, stride * sizeof(cd),
                            hipMemcpyDeviceToDevice, ctx->st));
    sh->jk_valid = true;
    sh->jk_mixed = mixed != 0;
    sh->jk_psi1 = pl.keys[3 * (nnew - 1)];
    sh->jk_psi2 = pl.keys[3 * (nnew - 1) + 1];
    sh->jk_u = pl.keys[3 * (nnew - 1) + 2];
    return out;
}

// ---------------------------------------------------------------- the fill
FillParams make_fill_params(const pssfss_sheet *sh, const pssfss_point *pt) {
    FillParams F;
    memset(&F, 0, sizeof(F));
    const bool mag = sh->cls == 'M';
    const int s = pt->s;
    cd e1 = lay(pt->epsr, s - 1), m1 = lay(pt->mur, s - 1);
    cd e2 = lay(pt->epsr, s), m2 = lay(pt->mur, s);
    const double k0 = pt->k0, u = pt->u;
    const double MU0 = 1.25663706212e-6, C0 = 299792458.0;
    F.k0 = k0;
    F.u = u;
    F.omega = k0 * C0;
    F.psi1 = pt->psi1;
    F.psi2 = pt->psi2;
    F.b1[0] = sh->b1[0] * sh->upm; F.b1[1] = sh->b1[1] * sh->upm;
    F.b2[0] = sh->b2[0] * sh->upm; F.b2[1] = sh->b2[1] * sh->upm;
    F.mut = ((2.0 * m1) * m2) / (m1 + m2);
    if (!mag) {
        F.epsb = (e1 + e2) / 2.0;
        F.c3 = c3_calc(k0, u, m1, e1, m2, e2);
        F.d3 = d3_calc(k0, u, m1, e1, m2, e2);
        F.psf_base = cd{0.0, 1.0} / F.epsb;
    } else {
        F.epsb = 0.5 * (e1 + e2);
        cd c3s = c3_calc(k0, u, m1, e1, m1, e1), c3sp1 = c3_calc(k0, u, m2, e2, m2, e2);
        cd d3s = d3_calc(k0, u, m1, e1, m1, e1), d3sp1 = d3_calc(k0, u, m2, e2, m2, e2);
        F.I1fact = PI_ / F.epsb;
        F.KFfact = (c3s * e1 + c3sp1 * e2) / ((2.0 * F.epsb) * u);
        F.I2fact = PI_ * F.mut;
        F.KPfact = (F.mut / (2.0 * u)) * (d3s / m1 + d3sp1 / m2);
        F.psf_base = cd{0.0, 2.0} / (((PI_ * F.omega) * MU0) * F.mut);
    }
    F.ff[0] = cd{1, 0};
    F.ff[1] = cd{1, 0};
    F.ff[2] = hcis(-pt->psi1);
    F.ff[3] = cd{1, 0};
    F.ff[4] = hcis(-pt->psi2);
    for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
            F.shift9[3 * (a + 1) + (b + 1)] =
                hcis(-((double)a * pt->psi1 + (double)b * pt->psi2));
    return F;
}

struct ChunkResult {
    std::vector<int> M, conv;
};

// Class terms of the fill are staged for at most this many bytes of points at once.
constexpr int64_t FILL_G_BYTES = 4ll << 30;

// Fill np points into ctx->wZ (row-major, one nbf^2 block per point), or, with zbase,
// into ld x ld blocks (ld >= nbf, leading dimension ld) from zbase on.
void fill_chunk(pssfss_ctx *ctx, pssfss_sheet *sh, const pssfss_point *pts, int np, int mixed,
                ChunkResult &res, TableSet *keep_tabs = nullptr,
                std::vector<FillParams> *keep_fp = nullptr, cd *zbase = nullptr,
                int64_t ld = 0) {
    hipStream_t st = ctx->st;
    const double ul = pts[0].u / sh->upm;
    for (int p = 1; p < np; ++p)
        if (pts[p].u != pts[0].u) throw ArgError("batched points must share u");
    // the spatial sums of every new phasing key go to the context's workspace,
    // which keeps its size between calls (about 96 B x classes x phasings, 40 GB
    // for a 102-point ring sweep: allocating and freeing it per call stalls)
    double key_u = ul;
    std::vector<const cd *> jkp = resolve_jk(ctx, sh, pts, np, mixed, ctx->wjk, key_u);
    ensure_singular(ctx, sh, key_u);
    std::vector<TabInput> tin(np);
    for (int p = 0; p < np; ++p) tin[p] = sheet_tab_input(sh, &pts[p]);
    // the tables are built on the auxiliary stream while the spatial sums run
    TableSet tabs;
    build_tables(ctx, tin, CONVTEST_FILL, tabs, ctx->st_aux);
    res.M = tabs.M;
    res.conv = tabs.converged;
    if (!zbase) {
        ld = sh->nbf;
        ctx->wZ.ensure((size_t)ld * ld * np);
        zbase = ctx->wZ.p;
    }
    std::vector<FillParams> fp(np);
    for (int p = 0; p < np; ++p) {
        fp[p] = make_fill_params(sh, &pts[p]);
        fp[p].M = tabs.M[p];
        fp[p].tab1 = tabs.tab1(p);
        fp[p].tab2 = tabs.tab2(p);
        fp[p].jk = jkp[p];
        fp[p].Z = zbase + p * ld * ld;
        fp[p].ldz = ld;
    }
    DBuf<FillParams> dfp(np);
    dfp.upload(fp.data(), np, st);
    // class terms of up to npg points at a time (9 complex per class and point)
    const int64_t gstride = 9 * sh->nufp;
    const int64_t gbytes = gstride * (int64_t)sizeof(cd);
    int npg = (int)std::max<int64_t>(1, std::min<int64_t>(np, FILL_G_BYTES / gbytes));
    ctx->wG.ensure((size_t)gstride * npg);
    const int mag = sh->cls == 'M' ? 1 : 0;
    stage_begin(2, st);
    for (int p0 = 0; p0 < np; p0 += npg) {
        const int n = std::min(npg, np - p0);
        launch_fill_classes(sh->dev(), dfp.p + p0, n, sh->rho_r.p, sh->rinv.p, mag, ctx->wG.p,
                            gstride, st);
        launch_fill_gather(sh->dev(), dfp.p + p0, n, sh->bf_face.p, sh->bf_loc.p, sh->fpcls.p,
                           ctx->wG.p, gstride, sh->has_fr ? 1 : 0, mag, st);
    }
    stage_end(2, st);
    sync(ctx);
    flush_jk_count(ctx);
    if (keep_fp) *keep_fp = fp;
    if (keep_tabs) *keep_tabs = std::move(tabs);
}

// Refinement sweeps before a system that has not converged is re-solved in FP64
// (LAPACK's zcgesv allows 30; stagnation ends it earlier).
constexpr int IR_MAXIT = 30;

// Mixed-precision refinement of Z X = B (tr false) or Z^T X = B (tr true) for the
// listed systems, from the FP32 factors in ctx->wZf / ctx->wperm and the FP64 Z in
// ctx->wZ.  B (device, n x ncol per system) is read; X goes to ctx->wX.  The
// first sweep solves with the FP32 factors from X = 0 (diagonal blocks applied
// through their inverses in ctx->wDinv, except for tr); each sweep then forms the
// FP64 residual and stops with LAPACK zcgesv's rule (max_c ||r_c|| < target sqrt(n)
// ||x_c|| ||op(Z)|| eps, target 1 being zcgesv's own), at IR_MAXIT sweeps, or when
// a sweep does not halve the backward error.  Systems not listed keep whatever ctx->wX holds.  With
// anorm_ready, ctx->wanorm already holds ||op(Z)||_inf of every system.
struct Refined {
    std::vector<int> iters, conv;
    std::vector<double> best;      // backward error in units of sqrt(n) eps
};

Refined refine_mixed(pssfss_ctx *ctx, int64_t n, int ncol, int np, bool tr, const cd *B,
                     const std::vector<int> &systems, bool anorm_ready = false,
                     double target = 1.0) {
    hipStream_t st = ctx->st;
    const int64_t nb2 = n * n, sB = n * ncol;
    ctx->wX.ensure((size_t)sB * np);
    ctx->wR.ensure((size_t)sB * np);
    ctx->wRf.ensure((size_t)sB * np);
    ctx->wTf.ensure((size_t)sB * np);
    ctx->wnorm.ensure((size_t)4 * np * ncol + 8);
    ctx->wsel.ensure((size_t)np);
    ctx->wanorm.ensure((size_t)np + 2);
    Refined out;
    out.iters.assign(np, 0);
    out.conv.assign(np, 0);
    out.best.assign(np, 1e300);
    if (systems.empty()) return out;
    double *rn = ctx->wnorm.p, *xn = ctx->wnorm.p + np * ncol;
    std::vector<double> hr(np * ncol), hx(np * ncol);
    std::vector<int32_t> act(systems.begin(), systems.end());
    auto upload_act = [&] {
        HIPCHECK(hipMemcpyAsync(ctx->wsel.p, act.data(), act.size() * sizeof(int32_t),
                                hipMemcpyHostToDevice, st));
    };
    upload_act();
    // ||op(Z)||_inf of each system for the stopping rule
    if (!anorm_ready) {
        if (tr) matrix_one_norms(n, np, ctx->wZ.p, nb2, ctx->wanorm.p, st);
        else matrix_inf_norms(n, np, ctx->wZ.p, nb2, ctx->wanorm.p, st);
    }
    std::vector<double> anorm(np + 1);
    d2h(ctx, anorm.data(), ctx->wanorm.p, np);
    launch_zero(ctx->wX.p, sB * np, st);
    convert_z_to_f32(B, ctx->wRf.p, sB * np, st);
    const cf *dinv = tr ? nullptr : ctx->wDinv.p;
    lu_solve_batched_f32(n, ncol, (int)act.size(), ctx->wZf.p, nb2, ctx->wperm.p, n, ctx->wRf.p,
                         sB, ctx->wTf.p, st, ctx->wsel.p, tr, dinv);
    add_correction_sel(ctx->wX.p, ctx->wRf.p, sB, ctx->wsel.p, (int)act.size(), st);
    std::vector<double> prev(np, 1e300);
    std::vector<int> done(np, 0);
    const double eps = 2.220446049250313e-16, sqn = std::sqrt((double)n);
    for (int it = 0; it < IR_MAXIT; ++it) {
        residual_batched(n, ncol, np, ctx->wZ.p, nb2, ctx->wX.p, B, sB, ctx->wR.p, rn, xn, st,
                         ctx->wsel.p, (int)act.size(), tr);
        d2h(ctx, hr.data(), rn, np * ncol);
        d2h(ctx, hx.data(), xn, np * ncol);
        sync(ctx);
        for (int p : act) {
            // normwise backward error of the worst column, in units of sqrt(n) eps
            double r = 0;
            for (int c = 0; c < ncol; ++c) {
                const double den = sqn * hx[p * ncol + c] * anorm[p] * eps;
                r = std::max(r, den > 0 ? hr[p * ncol + c] / den : (hr[p * ncol + c] > 0 ? 1e300 : 0));
            }
            out.iters[p] = it + 1;
            if (r < target) out.conv[p] = done[p] = 1;
            else if (r >= 0.5 * prev[p]) done[p] = 1;       // stagnated
            out.best[p] = std::min(out.best[p], r);
            prev[p] = r;
        }
        std::vector<int32_t> next;
        for (int p : act)
            if (!done[p]) next.push_back(p);
        if (next.empty()) break;
        act.swap(next);
        upload_act();
        // only the listed systems are solved and corrected
        residual_to_f32(ctx->wR.p, ctx->wRf.p, sB * np, st);
        lu_solve_batched_f32(n, ncol, (int)act.size(), ctx->wZf.p, nb2, ctx->wperm.p, n,
                             ctx->wRf.p, sB, ctx->wTf.p, st, ctx->wsel.p, tr, dinv);
        add_correction_sel(ctx->wX.p, ctx->wRf.p, sB, ctx->wsel.p, (int)act.size(), st);
    }
    return out;
}

// Backward-error target of the sheet GSM pipeline's mixed solves, in units of
// sqrt(n) eps (about 1.5e-11 normwise at 4485 unknowns).  Z of that pipeline holds
// spatial sums summed in FP32, which perturb it by about 1e-8 relative: a solve
// backward error three orders below that leaves S as the FP64-level solve does
// (max |dS| against the FP64 path unchanged to two digits on the two-layer patch,
// the oblique ring, the cross problems and the meander polarizer) and saves one
// or two residual sweeps per chunk.
constexpr double GSM_MIXED_IR_TARGET = 1000.0;

// Solve the np systems in ctx->wZ for the right-hand sides in ctx->wB (nbf x ncol
// column-major per point), leaving X in ctx->wB.  ir_target: refine_mixed's target
// (1: zcgesv's FP64-level rule).
void solve_chunk(pssfss_ctx *ctx, int64_t n, int ncol, int np, int precision,
                 pssfss_point_info *infos, double ir_target = 1.0) {
    hipStream_t st = ctx->st;
    const int64_t nb2 = n * n, sB = n * ncol;
    ctx->wpiv.ensure((size_t)n * np);
    ctx->wperm.ensure((size_t)n * np);
    if (precision == PSSFSS_FP64) {
        ctx->wT.ensure((size_t)sB * np);
        ctx->wlu.ensure((size_t)np * lu_factor_ws_elems(n));
        lu_factor_batched(n, np, ctx->wZ.p, nb2, ctx->wpiv.p, n, st, &ctx->lu_ov, ctx->wlu.p);
        stage_begin(5, st);
        lu_make_perm(n, np, ctx->wpiv.p, n, ctx->wperm.p, st);
        lu_solve_batched(n, ncol, np, ctx->wZ.p, nb2, ctx->wperm.p, n, ctx->wB.p, sB, ctx->wT.p, st);
        stage_end(5, st);
        sync(ctx);
        if (infos)
            for (int p = 0; p < np; ++p) { infos[p].refine_iters = 0; infos[p].refine_resid = 0; }
        return;
    }
    // FP32 LU with FP64 iterative refinement.
    ctx->wZf.ensure((size_t)nb2 * np);
    stage_begin(3, st);
    ctx->wanorm.ensure((size_t)np + 2);
    convert_z_to_f32_inf_norms(n, np, ctx->wZ.p, nb2, ctx->wZf.p, ctx->wanorm.p, st);
    stage_end(3, st);
    ctx->wluf.ensure((size_t)np * lu_factor_ws_elems(n));
    lu_factor_batched_f32(n, np, ctx->wZf.p, nb2, ctx->wpiv.p, n, st, &ctx->lu_ov, ctx->wluf.p);
    stage_begin(5, st);
    lu_make_perm(n, np, ctx->wpiv.p, n, ctx->wperm.p, st);
    ctx->wDinv.ensure((size_t)np * 2 * ((n + 63) / 64) * 4096);
    lu_diag_inverses_f32(n, np, ctx->wZf.p, nb2, ctx->wDinv.p, st);
    std::vector<int> all(np);
    for (int p = 0; p < np; ++p) all[p] = p;
    Refined rf = refine_mixed(ctx, n, ncol, np, false, ctx->wB.p, all, true, ir_target);
    stage_end(5, st);
    // Points whose refinement did not reach the target backward error: one
    // batched FP64 LU and solve of those systems.  After the refinement wZ (the FP64
    // matrices), wR (residuals) and the FP32 pivots are needed for these systems
    // only, so their Z and right-hand sides move to the front slots of wZ and wR:
    // slot i <= fb[i] and the moves go in increasing order, so every source is read
    // before its slot is overwritten.  The solutions go back to their points' slots
    // of wX.
    std::vector<int> fb;
    for (int p = 0; p < np; ++p)
        if (!rf.conv[p]) fb.push_back(p);
    if (!fb.empty()) {
        const int nf = (int)fb.size();
        ctx->wT.ensure((size_t)sB * nf);
        for (int i = 0; i < nf; ++i) {
            const int64_t p = fb[i];
            if (p != i)
                HIPCHECK(hipMemcpyAsync(ctx->wZ.p + i * nb2, ctx->wZ.p + p * nb2,
                                        (size_t)nb2 * sizeof(cd), hipMemcpyDeviceToDevice, st));
            HIPCHECK(hipMemcpyAsync(ctx->wR.p + i * sB, ctx->wB.p + p * sB, sB * sizeof(cd),
                                    hipMemcpyDeviceToDevice, st));
        }
        lu_factor_batched(n, nf, ctx->wZ.p, nb2, ctx->wpiv.p, n, st, &ctx->lu_ov);
        lu_make_perm(n, nf, ctx->wpiv.p, n, ctx->wperm.p, st);
        lu_solve_batched(n, ncol, nf, ctx->wZ.p, nb2, ctx->wperm.p, n, ctx->wR.p, sB, ctx->wT.p,
                         st);
        for (int i = 0; i < nf; ++i)
            HIPCHECK(hipMemcpyAsync(ctx->wX.p + (int64_t)fb[i] * sB, ctx->wR.p + i * sB,
                                    sB * sizeof(cd), hipMemcpyDeviceToDevice, st));
    }
    HIPCHECK(hipMemcpyAsync(ctx->wB.p, ctx->wX.p, (size_t)sB * np * sizeof(cd),
                            hipMemcpyDeviceToDevice, st));
    sync(ctx);
    if (infos)
        for (int p = 0; p < np; ++p) {
            infos[p].refine_iters = rf.iters[p];
            infos[p].refine_resid = rf.best[p];
            infos[p].fallback_fp64 = 0;
        }
    if (infos)
        for (int p : fb) infos[p].fallback_fp64 = 1;
}

// One chunk of the sheet GSM pipeline.
// Fourier-transform jobs of the points' end-region modes with the reuse rule of
// src/PSSFSS.jl:407-423; jobof[p][col] is the job of column col (region-1 modes
// first, then region-N modes).
void make_jobs(const pssfss_point *pts, int np, int n1, int n2, std::vector<BfftJob> &jobs,
               std::vector<std::vector<int>> &jobof) {
    const int ncol = n1 + n2;
    jobs.clear();
    jobof.assign(np, std::vector<int>(ncol));
    for (int p = 0; p < np; ++p) {
        const pssfss_point &pt = pts[p];
        const int nm[2] = {n1, n2};
        for (int sr = 0; sr < 2; ++sr)
            for (int qp = 0; qp < nm[sr]; ++qp) {
                const double *k = pt.beta[sr] + 2 * qp;
                int col = sr == 0 ? qp : n1 + qp;
                if (qp > 0 && approx2(k, pt.beta[sr] + 2 * (qp - 1))) {
                    jobof[p][col] = jobof[p][col - 1];
                    continue;
                }
                if (sr == 1 && n1 > qp && approx2(k, pt.beta[0] + 2 * qp)) {
                    jobof[p][col] = jobof[p][qp];
                    continue;
                }
                BfftJob J;
                J.k[0] = k[0];
                J.k[1] = k[1];
                J.ff[0] = cd{1, 0}; J.ff[1] = cd{1, 0}; J.ff[2] = hcis(-pt.psi1);
                J.ff[3] = cd{1, 0}; J.ff[4] = hcis(-pt.psi2);
                jobof[p][col] = (int)jobs.size();
                jobs.push_back(J);
            }
    }
}

// Source vector (7.39) of right-hand-side column col ('J' sheets: v c acf t;
// 'M' sheets: i c Y (z x t)).
void col_source(const pssfss_point &pt, int n1, int col, bool mag, cd &s0, cd &s1) {
    int sr = col < n1 ? 0 : 1, qp = col < n1 ? col : col - n1;
    cd vinc = lay(pt.vinc, qp + pt.nmax * sr);
    cd c = lay(pt.c[sr], qp);
    double tx = pt.tvec[sr][2 * qp], ty = pt.tvec[sr][2 * qp + 1];
    if (!mag) {
        cd f = (vinc * c) * pt.acf[sr];
        s0 = f * tx;
        s1 = f * ty;
    } else {
        cd f = (vinc * c) * lay(pt.Y[sr], qp);
        s0 = f * (-ty);
        s1 = f * tx;
    }
}

// The points of one sheet inside a chunk.
struct GsmGroup {
    pssfss_sheet *sh;
    const pssfss_point *pts;
    pssfss_gsm *gsms;
    pssfss_point_info *infos;   // optional
    int np;
};

// One chunk of the sheet GSM pipeline for the points of one or more sheets that
// share the end regions' mode counts.  Every system is held as an ld x ld block, ld
// the largest unknown count of the chunk; a sheet with fewer unknowns n fills the
// leading n x n block and the rest is the identity's (launch_pad_identity), its
// right-hand sides are zero below row n.  P [Z 0; 0 I] = [L 0; 0 I][U 0; 0 I] with
// the same pivots as Z alone (no interchange can reach a padded row: its entries in
// the first n columns are zero), so one batched factorisation and solve serve
// sheets of uneven size and give each sheet's own solutions.
void gsm_chunk_groups(pssfss_ctx *ctx, const std::vector<GsmGroup> &groups, int precision) {
    hipStream_t st = ctx->st;
    if (groups.empty()) return;
    const int n1 = groups[0].pts[0].nmodes[0], n2 = groups[0].pts[0].nmodes[1];
    int64_t ld = 0;
    int npt = 0;
    std::vector<int> off;
    for (const GsmGroup &g : groups) {
        for (int p = 0; p < g.np; ++p)
            if (g.pts[p].nmodes[0] != n1 || g.pts[p].nmodes[1] != n2)
                throw ArgError("batched points must have equal mode counts");
        ld = std::max<int64_t>(ld, g.sh->nbf);
        off.push_back(npt);
        npt += g.np;
    }
    const int ncol = n1 + n2;
    const int ng = (int)groups.size();
    // fill every group into its slots of one ld x ld block per point
    ctx->wZ.ensure((size_t)ld * ld * npt);
    std::vector<ChunkResult> cr(ng);
    bool padded = false;
    for (int g = 0; g < ng; ++g) {
        const GsmGroup &G = groups[g];
        cd *zb = ctx->wZ.p + (int64_t)off[g] * ld * ld;
        fill_chunk(ctx, G.sh, G.pts, G.np, precision == PSSFSS_MIXED, cr[g], nullptr, nullptr, zb,
                   ld);
        if (G.sh->nbf < ld) {
            launch_pad_identity(zb, G.sh->nbf, ld, ld * ld, G.np, st);
            padded = true;
        }
    }
    // transforms of the end-region modes and right-hand sides, per sheet
    ctx->wB.ensure((size_t)ld * ncol * npt);
    if (padded) launch_zero(ctx->wB.p, ld * ncol * npt, st);
    std::vector<std::vector<BfftJob>> jobs(ng);
    std::vector<std::vector<std::vector<int>>> jobof(ng);
    std::vector<DBuf<cd>> ft(ng);
    std::vector<DBuf<BfftJob>> djobs(ng);
    std::vector<DBuf<int32_t>> dcj(ng);
    std::vector<DBuf<cd>> dcs(ng);
    stage_begin(6, st);
    for (int g = 0; g < ng; ++g) {
        const GsmGroup &G = groups[g];
        const int64_t n = G.sh->nbf;
        const bool mag = G.sh->cls == 'M';
        make_jobs(G.pts, G.np, n1, n2, jobs[g], jobof[g]);
        const int nj = (int)jobs[g].size();
        djobs[g].upload(jobs[g].data(), nj, st);
        ft[g].ensure((size_t)nj * n * 2);
        launch_bfft(G.sh->dev(), G.sh->bf_face.p, G.sh->bf_loc.p, djobs[g].p, nj, ft[g].p, st);
        // Right-hand sides, eq. (7.39) / (7.64).
        std::vector<int32_t> coljob(G.np * ncol);
        std::vector<cd> colsrc(2 * G.np * ncol);
        for (int p = 0; p < G.np; ++p)
            for (int col = 0; col < ncol; ++col) {
                cd s0, s1;
                col_source(G.pts[p], n1, col, mag, s0, s1);
                coljob[p * ncol + col] = jobof[g][p][col];
                colsrc[2 * (p * ncol + col)] = s0;
                colsrc[2 * (p * ncol + col) + 1] = s1;
            }
        dcj[g].upload(coljob.data(), coljob.size(), st);
        dcs[g].upload(colsrc.data(), colsrc.size(), st);
        launch_rhs(n, G.np * ncol, ft[g].p, dcj[g].p, dcs[g].p,
                   ctx->wB.p + (int64_t)off[g] * ld * ncol, ld, st);
    }
    stage_end(6, st);
    std::vector<pssfss_point_info> infos(npt);
    memset(infos.data(), 0, sizeof(pssfss_point_info) * npt);
    solve_chunk(ctx, ld, ncol, npt, precision, infos.data(), GSM_MIXED_IR_TARGET);
    // Projection of the solutions onto the modes' transforms: one launch per point,
    // its columns against all jobs of its sheet.
    std::vector<size_t> oout(ng + 1, 0);
    for (int g = 0; g < ng; ++g)
        oout[g + 1] = oout[g] + (size_t)groups[g].np * ncol * jobs[g].size() * 2;
    ctx->wout.ensure(oout[ng]);
    std::vector<cd> hout(oout[ng]);
    stage_begin(6, st);
    for (int g = 0; g < ng; ++g) {
        const GsmGroup &G = groups[g];
        const int nj = (int)jobs[g].size();
        for (int p = 0; p < G.np; ++p)
            launch_project(G.sh->nbf, ncol, nj, ctx->wB.p + (int64_t)(off[g] + p) * ld * ncol, ld,
                           ft[g].p, ctx->wout.p + oout[g] + (size_t)p * ncol * nj * 2, st);
    }
    stage_end(6, st);
    d2h(ctx, hout.data(), ctx->wout.p, hout.size());
    sync(ctx);
    // Scattered-field GSM entries, eqs. (6.18) / (6.37).
    for (int g = 0; g < ng; ++g) {
        const GsmGroup &G = groups[g];
        const pssfss_sheet *sh = G.sh;
        const bool mag = sh->cls == 'M';
        const int njobs = (int)jobs[g].size();
        const cd *hg = hout.data() + oout[g];
        const double area = sh->s1[0] * sh->s2[1] - sh->s1[1] * sh->s2[0];
        const double area_m2 = fabs(area) / (sh->upm * sh->upm);
        for (int p = 0; p < G.np; ++p) {
            const pssfss_point &pt = G.pts[p];
            pssfss_gsm &gs = G.gsms[p];
            const int nm[2] = {n1, n2};
            double *blk[2][2] = {{gs.s11, gs.s12}, {gs.s21, gs.s22}};
            for (int col = 0; col < ncol; ++col) {
                int sr = col < n1 ? 0 : 1, qp = col < n1 ? col : col - n1;
                double sigma = sr == 0 ? 1.0 : -1.0;
                for (int orr = 0; orr < 2; ++orr) {
                    double *sm = blk[orr][sr];
                    if (!sm) continue;
                    for (int q = 0; q < nm[orr]; ++q) {
                        int ocol = orr == 0 ? q : n1 + q;
                        int job = jobof[g][p][ocol];
                        const cd *FT = &hg[(((size_t)p * ncol + col) * njobs + job) * 2];
                        double tx = pt.tvec[orr][2 * q], ty = pt.tvec[orr][2 * q + 1];
                        cd tl = lay(pt.tlgf, q + pt.nmax * orr);
                        cd cq = lay(pt.c[orr], q);
                        cd *e = (cd *)(sm + 2 * (q + (int64_t)nm[orr] * qp));
                        if (!mag) {
                            cd dot = tx * FT[0] + ty * FT[1];
                            *e = *e - dot * (tl / ((cq * pt.acf[orr]) * area_m2));
                        } else {
                            cd dot = (-ty) * FT[0] + tx * FT[1];
                            *e = *e + dot * ((sigma * tl) * cq);
                        }
                    }
                }
            }
            if (G.infos) {
                G.infos[p] = infos[off[g] + p];
                G.infos[p].mmax = cr[g].M[p];
                G.infos[p].table_converged = cr[g].conv[p];
                G.infos[p].jk_rings_max = 0;
            }
        }
    }
}

void gsm_chunk(pssfss_ctx *ctx, pssfss_sheet *sh, const pssfss_point *pts, pssfss_gsm *gsms,
               int np, int precision, pssfss_point_info *infos) {
    gsm_chunk_groups(ctx, {GsmGroup{sh, pts, gsms, infos, np}}, precision);
}

// Sensitivities of the sheet GSM along nparam geometry directions, FP64, 'J' or 'M'
// sheets with every face pair its own class.  With W the weight of each GSM entry
// (ocol = observed mode, col = source mode), S_scat[q,qp] = -kappa_{q,qp} g_q^T I_qp
// (eqs. 6.18 / 6.37: 'J' g_q = t_q . FT, kappa = tl_q / (c_q acf_q area); 'M'
// g_q = (z x t_q) . FT, kappa = -sigma_qp tl_q c_q, sigma = +1 for a source in
// region 1, -1 in region 2), Z^T lambda_q = g_q and
//   sum W dS = sum W (-kappa) (dg_q^T I_qp + lambda_q^T dV_qp)
//            + sum_{m,s} dZ_ms sum_qp A[m,qp] I[s,qp],  A = Lambda (W kappa).
// The first line uses the transforms' derivatives (host sums), the second the
// class terms' derivatives contracted on the device (dZ is never formed).
void adjoint_chunk(pssfss_ctx *ctx, pssfss_sheet *sh, const pssfss_point *pts, int np,
                   int nparam, const double *vtan, int nobj, const double *wbar,
                   double *grad, int precision) {
    hipStream_t st = ctx->st;
    const int n1 = pts[0].nmodes[0], n2 = pts[0].nmodes[1];
    for (int p = 1; p < np; ++p)
        if (pts[p].nmodes[0] != n1 || pts[p].nmodes[1] != n2)
            throw ArgError("batched points must have equal mode counts");
    const int ncol = n1 + n2;
    const int64_t n = sh->nbf, nf = sh->nface;
    const double ul = pts[0].u / sh->upm;
    const bool mag = sh->cls == 'M';
    // forward: fill (keeping tables), transforms, RHS, LU + solves in `precision`
    ChunkResult cr;
    TableSet tabs;
    std::vector<FillParams> fp;
    fill_chunk(ctx, sh, pts, np, 0, cr, &tabs, &fp);
    std::vector<BfftJob> jobs;
    std::vector<std::vector<int>> jobof;
    make_jobs(pts, np, n1, n2, jobs, jobof);
    const int njobs = (int)jobs.size();
    DBuf<BfftJob> djobs(njobs);
    djobs.upload(jobs.data(), njobs, st);
    ctx->wft.ensure((size_t)njobs * n * 2);
    launch_bfft(sh->dev(), sh->bf_face.p, sh->bf_loc.p, djobs.p, njobs, ctx->wft.p, st);
    std::vector<int32_t> coljob(np * ncol);
    std::vector<cd> colsrc(2 * np * ncol);
    std::vector<double> colt(2 * np * ncol);
    for (int p = 0; p < np; ++p)
        for (int col = 0; col < ncol; ++col) {
            col_source(pts[p], n1, col, mag, colsrc[2 * (p * ncol + col)],
                       colsrc[2 * (p * ncol + col) + 1]);
            coljob[p * ncol + col] = jobof[p][col];
            int orr = col < n1 ? 0 : 1, q = col < n1 ? col : col - n1;
            const double tx = pts[p].tvec[orr][2 * q], ty = pts[p].tvec[orr][2 * q + 1];
            colt[2 * (p * ncol + col)] = mag ? -ty : tx;
            colt[2 * (p * ncol + col) + 1] = mag ? tx : ty;
        }
    DBuf<int32_t> dcj(coljob.size());
    DBuf<cd> dcs(colsrc.size());
    DBuf<double> dct(colt.size());
    dcj.upload(coljob.data(), coljob.size(), st);
    dcs.upload(colsrc.data(), colsrc.size(), st);
    dct.upload(colt.data(), colt.size(), st);
    ctx->wB.ensure((size_t)n * ncol * np);
    launch_rhs(n, np * ncol, ctx->wft.p, dcj.p, dcs.p, ctx->wB.p, n, st);
    std::vector<pssfss_point_info> sinfo(np);
    solve_chunk(ctx, n, ncol, np, precision, sinfo.data());
    // adjoint: observation vectors, Z^T lambda = g with the same factors
    const int64_t sB = n * ncol, nb2 = n * n;
    DBuf<cd> lam((size_t)sB * np), tmp((size_t)sB * np);
    launch_obs(n, np * ncol, ctx->wft.p, dcj.p, dct.p, lam.p, n, st);
    if (precision == PSSFSS_FP64) {
        lu_solve_batched(n, ncol, np, ctx->wZ.p, nb2, ctx->wperm.p, n, lam.p, sB, tmp.p, st,
                         nullptr, true);
    } else {
        // systems refined in mixed precision still hold Z (FP64) and FP32 factors:
        // transposed refinement; systems that fell back hold FP64 factors in wZ
        std::vector<int> ref, fac;
        for (int p = 0; p < np; ++p) (sinfo[p].fallback_fp64 ? fac : ref).push_back(p);
        Refined rt = refine_mixed(ctx, n, ncol, np, true, lam.p, ref);
        for (int p : ref) {
            if (rt.conv[p]) {
                HIPCHECK(hipMemcpyAsync(lam.p + p * sB, ctx->wX.p + p * sB, sB * sizeof(cd),
                                        hipMemcpyDeviceToDevice, st));
            } else {
                lu_factor_batched(n, 1, ctx->wZ.p + p * nb2, nb2, ctx->wpiv.p + p * n, n, st);
                lu_make_perm(n, 1, ctx->wpiv.p + p * n, n, ctx->wperm.p + p * n, st);
                fac.push_back(p);
            }
        }
        for (int p : fac)
            lu_solve_batched(n, ncol, 1, ctx->wZ.p + p * nb2, nb2, ctx->wperm.p + p * n, n,
                             lam.p + p * sB, sB, tmp.p + p * sB, st, nullptr, true);
    }
    std::vector<cd> hX((size_t)n * ncol * np), hL((size_t)n * ncol * np);
    d2h(ctx, hX.data(), ctx->wB.p, hX.size());
    d2h(ctx, hL.data(), lam.p, hL.size());
    sync(ctx);
    // weights W kappa and A = Lambda (W kappa) per point and objective
    const double area = sh->s1[0] * sh->s2[1] - sh->s1[1] * sh->s2[0];
    const double area_m2 = fabs(area) / (sh->upm * sh->upm);
    const int nm[2] = {n1, n2};
    std::vector<cd> Wk((size_t)np * nobj * ncol * ncol), hA((size_t)np * nobj * n * ncol);
    std::vector<cd> kap((size_t)np * ncol);
    const int64_t wsz = (int64_t)ncol * ncol;
    for (int p = 0; p < np; ++p) {
        const pssfss_point &pt = pts[p];
        // kappa_{oc,c} = kap[oc] ('J'), kap[oc] * sigma_c ('M')
        for (int oc = 0; oc < ncol; ++oc) {
            int orr = oc < n1 ? 0 : 1, q = oc < n1 ? oc : oc - n1;
            cd tl = lay(pt.tlgf, q + pt.nmax * orr);
            cd cq = lay(pt.c[orr], q);
            kap[p * ncol + oc] = mag ? -(tl * cq) : tl / ((cq * pt.acf[orr]) * area_m2);
        }
        for (int o = 0; o < nobj; ++o) {
            const double *wb = wbar + 2 * ((int64_t)p * nobj + o) * wsz;
            // blocks s11, s12, s21, s22, column-major, as pssfss_gsm
            int64_t off = 0;
            for (int orr = 0; orr < 2; ++orr)
                for (int sr = 0; sr < 2; ++sr) {
                    for (int qp = 0; qp < nm[sr]; ++qp)
                        for (int q = 0; q < nm[orr]; ++q) {
                            int oc = orr == 0 ? q : n1 + q, c = sr == 0 ? qp : n1 + qp;
                            cd w = lay(wb, (int)(off + q + (int64_t)nm[orr] * qp));
                            const double sigma = mag && sr == 1 ? -1.0 : 1.0;
                            Wk[((size_t)p * nobj + o) * wsz + oc * ncol + c] =
                                w * (sigma * kap[p * ncol + oc]);
                        }
                    off += (int64_t)nm[orr] * nm[sr];
                }
            cd *Ao = &hA[((size_t)p * nobj + o) * n * ncol];
            const cd *L = &hL[(size_t)p * n * ncol];
            const cd *Wo = &Wk[((size_t)p * nobj + o) * wsz];
            for (int c = 0; c < ncol; ++c)
                for (int64_t m = 0; m < n; ++m) {
                    cd a = {0, 0};
                    for (int oc = 0; oc < ncol; ++oc) a += L[m + n * oc] * Wo[oc * ncol + c];
                    Ao[m + n * c] = a;
                }
        }
    }
    DBuf<cd> dA(hA.size());
    dA.upload(hA.data(), hA.size(), st);
    DBuf<FillParams> dfp(np);
    dfp.upload(fp.data(), np, st);
    // per group of up to ADJ_NDIR directions: tangent spatial sums (all directions
    // of the group in one pass per phasing key) and singular terms, class
    // contraction, transform derivatives
    const int64_t nu = sh->nufp;
    constexpr int ADJ_NDIR = 4;
    const int64_t sft = 6 * nf, sjk = 6 * nu, sdft = (int64_t)njobs * n * 2;
    DBuf<double> fts((size_t)ADJ_NDIR * sft), ftm((size_t)ADJ_NDIR * sft);
    DBuf<cdd> jkd((size_t)ADJ_NDIR * sjk);
    DBuf<dd> rrd((size_t)ADJ_NDIR * 2 * nu), rid((size_t)ADJ_NDIR * nu);
    DBuf<cd> part((size_t)((nu + 255) / 256) * nobj), zout((size_t)ADJ_NDIR * nobj * np);
    DBuf<cd> dft((size_t)ADJ_NDIR * sdft);
    std::vector<double> hts((size_t)ADJ_NDIR * sft), htm((size_t)ADJ_NDIR * sft);
    std::vector<cd> hz((size_t)ADJ_NDIR * nobj * np), hdft((size_t)ADJ_NDIR * sdft);
    const double us1[2] = {ul * sh->s1[0], ul * sh->s1[1]};
    const double us2[2] = {ul * sh->s2[0], ul * sh->s2[1]};
    for (int k0 = 0; k0 < nparam; k0 += ADJ_NDIR) {
        const int nd = std::min(ADJ_NDIR, nparam - k0);
        for (int d = 0; d < nd; ++d) {
            const double *vt = vtan + (size_t)(k0 + d) * 2 * sh->nnode;
            double *ts = &hts[d * sft], *tm = &htm[d * sft];
            for (int64_t f = 0; f < nf; ++f)
                for (int a = 0; a < 3; ++a) {
                    const int v = sh->hfv[3 * f + a];
                    ts[6 * f + 2 * a] = vt[2 * v];
                    ts[6 * f + 2 * a + 1] = vt[2 * v + 1];
                    tm[6 * f + 2 * a] = vt[2 * v] / sh->upm;
                    tm[6 * f + 2 * a + 1] = vt[2 * v + 1] / sh->upm;
                }
        }
        fts.upload(hts.data(), (size_t)nd * sft, st);
        ftm.upload(htm.data(), (size_t)nd * sft, st);
        for (int d = 0; d < nd; ++d)
            launch_singular_tan(sh->dev(), ul, fts.p + d * sft, rrd.p + d * 2 * nu,
                                rid.p + d * nu, st);
        double kpsi1 = NAN, kpsi2 = NAN;
        for (int p = 0; p < np; ++p) {
            if (!(pts[p].psi1 == kpsi1 && pts[p].psi2 == kpsi2)) {
                kpsi1 = pts[p].psi1;
                kpsi2 = pts[p].psi2;
                launch_jk_tan_multi(sh->dev(), ul, us1, us2, kpsi1, kpsi2, fts.p, sft, nd,
                                    jkd.p, sjk, st);
            }
            for (int d = 0; d < nd; ++d)
                launch_class_tan(sh->dev(), dfp.p + p, jkd.p + d * sjk, rrd.p + d * 2 * nu,
                                 rid.p + d * nu, ftm.p + d * sft,
                                 dA.p + (size_t)p * nobj * n * ncol, n * ncol,
                                 ctx->wB.p + (size_t)p * n * ncol, ncol, nobj, mag ? 1 : 0,
                                 sh->has_fr ? 1 : 0, part.p,
                                 zout.p + ((size_t)d * np + p) * nobj, st);
        }
        for (int d = 0; d < nd; ++d)
            launch_bfft_tan(sh->dev(), sh->bf_face.p, sh->bf_loc.p, djobs.p, njobs,
                            ftm.p + d * sft, dft.p + d * sdft, st);
        d2h(ctx, hz.data(), zout.p, (size_t)nd * nobj * np);
        d2h(ctx, hdft.data(), dft.p, (size_t)nd * sdft);
        sync(ctx);
        for (int d = 0; d < nd; ++d) {
            const int k = k0 + d;
            const cd *hzd = &hz[(size_t)d * nobj * np];
            const cd *hdftd = &hdft[(size_t)d * sdft];
            for (int p = 0; p < np; ++p) {
                const cd *X = &hX[(size_t)p * n * ncol];
                const cd *L = &hL[(size_t)p * n * ncol];
                // explicit terms per (observed, source) pair
                std::vector<cd> ex((size_t)wsz);
                for (int oc = 0; oc < ncol; ++oc) {
                    const cd *dg = &hdftd[(size_t)jobof[p][oc] * n * 2];
                    const double tx = colt[2 * (p * ncol + oc)];
                    const double ty = colt[2 * (p * ncol + oc) + 1];
                    for (int c = 0; c < ncol; ++c) {
                        const cd *dv = &hdftd[(size_t)jobof[p][c] * n * 2];
                        const cd s0 = colsrc[2 * (p * ncol + c)];
                        const cd s1 = colsrc[2 * (p * ncol + c) + 1];
                        cd t = {0, 0};
                        for (int64_t m = 0; m < n; ++m) {
                            cd g = tx * dg[2 * m] + ty * dg[2 * m + 1];
                            cd v = conj(dv[2 * m]) * s0 + conj(dv[2 * m + 1]) * s1;
                            t += g * X[m + n * c] + L[m + n * oc] * v;
                        }
                        ex[oc * ncol + c] = t;
                    }
                }
                for (int o = 0; o < nobj; ++o) {
                    const cd *Wo = &Wk[((size_t)p * nobj + o) * wsz];
                    cd gsum = hzd[(size_t)p * nobj + o];
                    for (int64_t e = 0; e < wsz; ++e) gsum = gsum - Wo[e] * ex[e];
                    double *gp = grad + 2 * (((size_t)p * nobj + o) * nparam + k);
                    gp[0] = gsum.x;
                    gp[1] = gsum.y;
                }
            }
        }
    }
}

int chunk_size(pssfss_ctx *ctx, const pssfss_sheet *sh, int npts, int precision) {
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) {
        (void)hipGetLastError();
        freeb = (size_t)8 << 30;
    }
    double per = 16.0 * sh->nbf * sh->nbf * (precision == PSSFSS_MIXED ? 1.5 : 1.0) +
                 96.0 * sh->nufp + 4.0e6;
    double budget = ctx->mem_frac * (double)freeb;
    int c = (int)std::max(1.0, std::floor(budget / per));
    return std::min(std::min(c, ctx->chunk_max), npts);
}

void upload_sheet(pssfss_ctx *ctx, const pssfss_sheet_desc *d, pssfss_sheet *sh) {
    if (!d) throw ArgError("null sheet descriptor");
    if (d->nface <= 0 || d->nnode <= 0 || d->nedge <= 0 || d->nbf <= 0)
        throw ArgError("sheet: sizes must be positive");
    if (!d->rho || !d->fv || !d->fe || !d->bff || !d->ebf || !d->eci)
        throw ArgError("sheet: null arrays");
    if (d->sheet_class != 'J' && d->sheet_class != 'M')
        throw ArgError("sheet: class must be 'J' or 'M'");
    if (!(d->units_per_meter > 0)) throw ArgError("sheet: units_per_meter must be positive");
    const int64_t nf = d->nface, ne = d->nedge, nb = d->nbf, nn = d->nnode;
    if (nb > 200000 || nf > 2000000) throw ArgError("sheet: too large");
    sh->nnode = nn; sh->nface = nf; sh->nedge = ne; sh->nbf = nb;
    sh->cls = d->sheet_class;
    sh->upm = d->units_per_meter;
    for (int i = 0; i < 2; ++i) {
        sh->s1[i] = d->s1[i]; sh->s2[i] = d->s2[i];
        sh->b1[i] = d->beta1[i]; sh->b2[i] = d->beta2[i];
    }
    std::vector<double> fxs(6 * nf), fxm(6 * nf), fr(nf, 0.0);
    std::vector<int32_t> fbf(3 * nf), fsg(3 * nf), feci(3 * nf), fedge(3 * nf);
    sh->hfv.assign(3 * nf, 0);
    for (int64_t f = 0; f < nf; ++f) {
        for (int i = 0; i < 3; ++i) {
            int64_t v = d->fv[3 * f + i] - 1;
            if (v < 0 || v >= nn) throw ArgError("sheet: fv index out of range");
            sh->hfv[3 * f + i] = (int32_t)v;
            fxs[6 * f + 2 * i] = d->rho[2 * v];
            fxs[6 * f + 2 * i + 1] = d->rho[2 * v + 1];
            fxm[6 * f + 2 * i] = d->rho[2 * v] / sh->upm;
            fxm[6 * f + 2 * i + 1] = d->rho[2 * v + 1] / sh->upm;
            int64_t e = d->fe[3 * f + i] - 1;
            if (e < 0 || e >= ne) throw ArgError("sheet: fe index out of range");
            int64_t b = d->ebf[e] - 1;
            if (b < -1 || b >= nb) throw ArgError("sheet: ebf index out of range");
            int64_t ec = d->eci[e];
            if (ec < 0 || ec > 4) throw ArgError("sheet: eci must be 0..4");
            fbf[3 * f + i] = (int32_t)b;
            feci[3 * f + i] = (int32_t)ec;
            fedge[3 * f + i] = (int32_t)e;
            if (b >= 0) {
                int64_t pf = d->bff[2 * b] - 1, mf = d->bff[2 * b + 1] - 1;
                if (pf == f) fsg[3 * f + i] = 1;
                else if (mf == f) fsg[3 * f + i] = -1;
                else throw ArgError("sheet: basis function does not touch its face");
            } else {
                fsg[3 * f + i] = 0;
            }
        }
        if (d->fr) fr[f] = d->fr[f];
    }
    // the two half-basis functions of every basis function
    std::vector<int32_t> bface(2 * nb, -1), bloc(2 * nb, -1);
    std::vector<int> cnt(nb, 0);
    for (int64_t f = 0; f < nf; ++f)
        for (int i = 0; i < 3; ++i) {
            int b = fbf[3 * f + i];
            if (b < 0) continue;
            if (cnt[b] >= 2) throw ArgError("sheet: basis function with more than two halves");
            bface[2 * b + cnt[b]] = (int32_t)f;
            bloc[2 * b + cnt[b]] = i;
            ++cnt[b];
        }
    for (int64_t b = 0; b < nb; ++b)
        if (cnt[b] != 2) throw ArgError("sheet: basis function without two halves");
    if (d->ufp_offsets) {
        if (!d->ufp_members) throw ArgError("sheet: ufp_members missing");
        sh->identity = false;
        sh->nufp = d->nufp;
        std::vector<int64_t> off(d->ufp_offsets, d->ufp_offsets + d->nufp + 1);
        int64_t nm = off[d->nufp];
        if (off[0] != 0 || nm != nf * nf) throw ArgError("sheet: ufp classes must cover nface^2");
        std::vector<int64_t> mem(nm);
        for (int64_t i = 0; i < nm; ++i) {
            int64_t v = d->ufp_members[i] - 1;
            if (v < 0 || v >= nf * nf) throw ArgError("sheet: ufp member out of range");
            mem[i] = v;
        }
        for (int64_t i = 0; i < d->nufp; ++i)
            if (off[i + 1] <= off[i]) throw ArgError("sheet: empty ufp class");
        std::vector<int32_t> fpc(nm);
        for (int64_t i = 0; i < d->nufp; ++i)
            for (int64_t j = off[i]; j < off[i + 1]; ++j) fpc[mem[j]] = (int32_t)i;
        sh->uoff.upload(off.data(), off.size(), ctx->st);
        sh->umem.upload(mem.data(), mem.size(), ctx->st);
        sh->fpcls.upload(fpc.data(), fpc.size(), ctx->st);
    } else {
        if (d->nufp != nf * nf) throw ArgError("sheet: nufp must be nface^2 without classes");
        sh->identity = true;
        sh->nufp = nf * nf;
    }
    sh->has_fr = false;
    for (double v : fr) if (v != 0) sh->has_fr = true;
    hipStream_t st = ctx->st;
    sh->fxs.upload(fxs.data(), fxs.size(), st);
    sh->fxm.upload(fxm.data(), fxm.size(), st);
    sh->fr.upload(fr.data(), fr.size(), st);
    sh->fbf.upload(fbf.data(), fbf.size(), st);
    sh->fsg.upload(fsg.data(), fsg.size(), st);
    sh->feci.upload(feci.data(), feci.size(), st);
    sh->fedge.upload(fedge.data(), fedge.size(), st);
    sh->bf_face.upload(bface.data(), bface.size(), st);
    sh->bf_loc.upload(bloc.data(), bloc.size(), st);
    sync(ctx);
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

namespace {
// CUs given to the look-ahead LU's panel stream; the update stream gets the rest.
// Batches of more systems than this keep the single-stream factorisation (each
// system's leaf takes a CU of its own).
constexpr int LU_PANEL_CUS = 32;

void destroy_overlap(LuOverlap &ov) {
    for (hipEvent_t *e : {&ov.e_start, &ov.e_u12, &ov.e_panel, &ov.e_done})
        if (*e) (void)hipEventDestroy(*e), *e = nullptr;
    if (ov.panel) (void)hipStreamDestroy(ov.panel), ov.panel = nullptr;
    if (ov.update) (void)hipStreamDestroy(ov.update), ov.update = nullptr;
    ov.max_batch = 0;
}

// Two streams for the look-ahead LU (default priority): the update stream is kept
// off the first LU_PANEL_CUS CUs; on any failure the context keeps the
// single-stream factorisation.
void init_overlap(pssfss_ctx *c) {
    LuOverlap &ov = c->lu_ov;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, c->device) != hipSuccess) { (void)hipGetLastError(); return; }
    const int ncu = p.multiProcessorCount;
    if (ncu < 4 * LU_PANEL_CUS) return;
    const int words = (ncu + 31) / 32;
    std::vector<uint32_t> mu(words, 0);
    for (int cu = LU_PANEL_CUS; cu < ncu; ++cu) mu[cu / 32] |= 1u << (cu % 32);
    // the panel stream may use every CU; the update stream leaves the first
    // LU_PANEL_CUS to the panel's leaves
    bool ok = hipStreamCreate(&ov.panel) == hipSuccess &&
              hipExtStreamCreateWithCUMask(&ov.update, words, mu.data()) == hipSuccess;
    for (hipEvent_t *e : {&ov.e_start, &ov.e_u12, &ov.e_panel, &ov.e_done})
        ok = ok && hipEventCreateWithFlags(e, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        destroy_overlap(ov);
        return;
    }
    ov.max_batch = LU_PANEL_CUS;
}
}  // namespace

int pssfss_ctx_create(int device, pssfss_ctx **out) {
    if (!out) return PSSFSS_EINVAL;
    *out = nullptr;
    pssfss_ctx *c = new (std::nothrow) pssfss_ctx();
    if (!c) return PSSFSS_ENOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return PSSFSS_EHIP;
    }
    if (hipStreamCreate(&c->st_aux) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamDestroy(c->st);
        delete c;
        return PSSFSS_EHIP;
    }
    c->scratch_dev = scratch_address(device);
    if (!c->scratch_dev) {
        (void)hipStreamDestroy(c->st_aux);
        (void)hipStreamDestroy(c->st);
        delete c;
        return PSSFSS_EHIP;
    }
    init_overlap(c);
    *out = c;
    return PSSFSS_OK;
}

int pssfss_ctx_destroy(pssfss_ctx *ctx) {
    if (!ctx) return PSSFSS_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->st);
    (void)hipStreamSynchronize(ctx->st_aux);
    // sheets the caller did not free go with their context
    for (pssfss_sheet *sh : ctx->sheets) delete sh;
    ctx->sheets.clear();
    ctx->dft.clear();
    destroy_overlap(ctx->lu_ov);
    (void)hipStreamDestroy(ctx->st_aux);
    (void)hipStreamDestroy(ctx->st);
    delete ctx;
    return PSSFSS_OK;
}

const char *pssfss_last_error(const pssfss_ctx *ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

int pssfss_device_info(pssfss_ctx *ctx, char *name, int name_len, int *num_cu, double *hbm_gib) {
    return guarded(ctx, [&] {
        hipDeviceProp_t p;
        HIPCHECK(hipGetDeviceProperties(&p, ctx->device));
        if (name && name_len > 0) {
            snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
        }
        if (num_cu) *num_cu = p.multiProcessorCount;
        if (hbm_gib) *hbm_gib = (double)p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0);
    });
}

int pssfss_sheet_upload(pssfss_ctx *ctx, const pssfss_sheet_desc *desc, pssfss_sheet **out) {
    return guarded(ctx, [&] {
        if (!out) throw ArgError("null output");
        *out = nullptr;
        HIPCHECK(hipSetDevice(ctx->device));
        pssfss_sheet *sh = new pssfss_sheet();
        try {
            upload_sheet(ctx, desc, sh);
            ctx->sheets.insert(sh);
        } catch (...) {
            delete sh;
            throw;
        }
        *out = sh;
    });
}

int pssfss_sheet_free(pssfss_ctx *ctx, pssfss_sheet *sheet) {
    return guarded(ctx, [&] {
        if (!sheet) return;
        if (!ctx->sheets.erase(sheet)) throw ArgError("sheet is not live in this context");
        HIPCHECK(hipSetDevice(ctx->device));
        sync(ctx);
        if (ctx->pf.sheet == sheet) ctx->pf.valid = false;
        delete sheet;
    });
}

int pssfss_sheet_gsm_batch(pssfss_ctx *ctx, pssfss_sheet *sheet, int npts,
                           const pssfss_point *pts, pssfss_gsm *gsms, int precision,
                           pssfss_point_info *infos) {
    return guarded(ctx, [&] {
        if (!sheet || !pts || !gsms || npts < 0) throw ArgError("null argument");
        if (precision != PSSFSS_FP64 && precision != PSSFSS_MIXED)
            throw ArgError("precision must be PSSFSS_FP64 or PSSFSS_MIXED");
        HIPCHECK(hipSetDevice(ctx->device));
        for (int p = 0; p < npts; ++p) check_point(sheet, &pts[p], true);
        if (infos) memset(infos, 0, sizeof(pssfss_point_info) * npts);
        int c = npts > 0 ? chunk_size(ctx, sheet, npts, precision) : 1;
        for (int p0 = 0; p0 < npts; p0 += c) {
            int np = std::min(c, npts - p0);
            gsm_chunk(ctx, sheet, pts + p0, gsms + p0, np, precision, infos ? infos + p0 : nullptr);
        }
    });
}

int pssfss_sheet_prefetch_jk(pssfss_ctx *ctx, pssfss_sheet *sheet, int npts, const double *psi1,
                             const double *psi2, const double *u, int precision) {
    return guarded(ctx, [&] {
        if (!sheet || npts < 0 || (npts > 0 && (!psi1 || !psi2 || !u)))
            throw ArgError("null argument");
        if (precision != PSSFSS_FP64 && precision != PSSFSS_MIXED)
            throw ArgError("precision must be PSSFSS_FP64 or PSSFSS_MIXED");
        if (!ctx->sheets.count(sheet)) throw ArgError("sheet is not live in this context");
        HIPCHECK(hipSetDevice(ctx->device));
        if (npts == 0) return;
        std::vector<double> ul(npts);
        for (int p = 0; p < npts; ++p) {
            if (!(u[p] > 0)) throw ArgError("u must be positive");
            ul[p] = u[p] / sheet->upm;
        }
        const int mixed = precision == PSSFSS_MIXED ? 1 : 0;
        bool one = true;   // every point reuses the first point's key
        for (int p = 1; p < npts && one; ++p)
            one = key_matches(sheet, true, psi1[0], psi2[0], ul[0], psi1[p], psi2[p], ul[p]);
        if (one) {         // into the sheet's cache, where the GSM call finds it
            const bool valid = sheet->jk_valid && (mixed || !sheet->jk_mixed);
            if (!key_matches(sheet, valid, sheet->jk_psi1, sheet->jk_psi2, sheet->jk_u, psi1[0],
                             psi2[0], ul[0])) {
                sheet->jk.ensure((size_t)6 * sheet->nufp);
                compute_jk(ctx, sheet, ul[0], {psi1[0], psi2[0]}, sheet->jk.p, mixed);
                sheet->jk_valid = true;
                sheet->jk_mixed = mixed != 0;
                sheet->jk_psi1 = psi1[0];
                sheet->jk_psi2 = psi2[0];
                sheet->jk_u = ul[0];
            }
            ensure_singular(ctx, sheet, ul[0]);
            return;
        }
        // several keys: the plan of the GSM call's single chunk into the context's
        // slots, used there when that call plans the same keys from the same cache
        if (chunk_size(ctx, sheet, npts, precision) < npts) return;
        const JkPlan pl = plan_jk(sheet, npts, psi1, psi2, ul.data(), mixed);
        ctx->pf.valid = false;
        compute_plan(ctx, sheet, pl, mixed, ctx->wjk);
        auto &f = ctx->pf;
        f.sheet = sheet;
        f.mixed = mixed;
        f.keys = pl.keys;
        f.uses_old = pl.uses_old;
        f.c_valid = sheet->jk_valid;
        f.c_mixed = sheet->jk_mixed;
        f.c_psi1 = sheet->jk_psi1;
        f.c_psi2 = sheet->jk_psi2;
        f.c_u = sheet->jk_u;
        f.valid = true;
        ensure_singular(ctx, sheet, pl.slot[0] < 0 ? sheet->jk_u : pl.keys[3 * pl.slot[0] + 2]);
    });
}

int pssfss_sheets_gsm_batch(pssfss_ctx *ctx, int nsheet, pssfss_sheet *const *sheets,
                            const int *npts, const pssfss_point *pts, pssfss_gsm *gsms,
                            int precision, pssfss_point_info *infos) {
    return guarded(ctx, [&] {
        if (nsheet < 0 || (nsheet > 0 && (!sheets || !npts)) || !gsms || !pts)
            throw ArgError("null argument");
        if (precision != PSSFSS_FP64 && precision != PSSFSS_MIXED)
            throw ArgError("precision must be PSSFSS_FP64 or PSSFSS_MIXED");
        HIPCHECK(hipSetDevice(ctx->device));
        int total = 0;
        pssfss_sheet *big = nullptr;
        for (int g = 0; g < nsheet; ++g) {
            if (!sheets[g] || npts[g] < 0) throw ArgError("null sheet or negative point count");
            for (int p = 0; p < npts[g]; ++p) check_point(sheets[g], &pts[total + p], true);
            total += npts[g];
            if (npts[g] > 0 && (!big || sheets[g]->nbf > big->nbf)) big = sheets[g];
        }
        if (infos) memset(infos, 0, sizeof(pssfss_point_info) * total);
        if (total == 0) return;
        // batches sized for the largest sheet (every system is held at its size)
        const int c = chunk_size(ctx, big, total, precision);
        int g = 0, p = 0;        // next point: sheet g, its point p
        int done = 0;
        while (done < total) {
            std::vector<GsmGroup> groups;
            int take = std::min(c, total - done);
            while (take > 0) {
                while (p >= npts[g]) { ++g; p = 0; }
                const int k = std::min(take, npts[g] - p);
                groups.push_back(GsmGroup{sheets[g], pts + done, gsms + done,
                                          infos ? infos + done : nullptr, k});
                p += k;
                done += k;
                take -= k;
            }
            gsm_chunk_groups(ctx, groups, precision);
        }
    });
}

int pssfss_sheet_adjoint(pssfss_ctx *ctx, pssfss_sheet *sheet, int npts,
                         const pssfss_point *pts, int nparam, const double *vtan, int nobj,
                         const double *wbar, double *grad, int precision) {
    return guarded(ctx, [&] {
        if (!sheet || !pts || !vtan || !wbar || !grad || npts < 0 || nparam < 0)
            throw ArgError("null argument");
        if (precision != PSSFSS_FP64 && precision != PSSFSS_MIXED)
            throw ArgError("precision must be PSSFSS_FP64 or PSSFSS_MIXED");
        if (nobj < 1 || nobj > 4) throw ArgError("adjoint: 1..4 objectives");
        if (!sheet->identity)
            throw ArgError("adjoint: every face pair must be its own class (no ufp classes)");
        HIPCHECK(hipSetDevice(ctx->device));
        for (int p = 0; p < npts; ++p) check_point(sheet, &pts[p], true);
        if (npts == 0 || nparam == 0) return;
        int c = chunk_size(ctx, sheet, npts, precision);
        const int ncol = pts[0].nmodes[0] + pts[0].nmodes[1];
        for (int p0 = 0; p0 < npts; p0 += c) {
            int np = std::min(c, npts - p0);
            adjoint_chunk(ctx, sheet, pts + p0, np, nparam, vtan, nobj,
                          wbar + 2 * (size_t)p0 * nobj * ncol * ncol,
                          grad + 2 * (size_t)p0 * nobj * nparam, precision);
        }
    });
}

int pssfss_face_pair_classes(pssfss_ctx *ctx, int64_t nnode, const double *rho,
                             int64_t nface, const int64_t *fv, int64_t *nufp, int64_t *offsets,
                             int64_t *members, int64_t *ufpm) {
    return guarded(ctx, [&] {
        if (!rho || !fv || !nufp || !offsets || !members || nnode <= 0 || nface < 2)
            throw ArgError("null argument or fewer than two faces");
        HIPCHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->st;
        for (int64_t i = 0; i < 3 * nface; ++i)
            if (fv[i] < 1 || fv[i] > nnode) throw ArgError("fv index out of range");
        // grouping radius of the reference: 1e-6 |c1 - c2| (first two centroids)
        double c[2][2];
        for (int f = 0; f < 2; ++f)
            for (int d = 0; d < 2; ++d)
                c[f][d] = ((rho[2 * (fv[3 * f] - 1) + d] + rho[2 * (fv[3 * f + 1] - 1) + d]) +
                           rho[2 * (fv[3 * f + 2] - 1) + d]) / 3.0;
        const double radius = 1e-6 * hypot(c[0][0] - c[1][0], c[0][1] - c[1][1]);
        const int64_t npair = nface * nface;
        DBuf<double> drho(2 * nnode);
        DBuf<int64_t> dfv(3 * nface), dufpm(npair), dmem(npair);
        DBuf<uint64_t> doff(npair + 1);
        drho.upload(rho, 2 * nnode, st);
        dfv.upload(fv, 3 * nface, st);
        const int64_t n = face_pair_classes_dev(nnode, drho.p, nface, dfv.p, radius, dufpm.p,
                                                dmem.p, doff.p, st);
        *nufp = n;
        HIPCHECK(hipMemcpyAsync(offsets, doff.p, 8 * (n + 1), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(members, dmem.p, 8 * npair, hipMemcpyDeviceToHost, st));
        if (ufpm) HIPCHECK(hipMemcpyAsync(ufpm, dufpm.p, 8 * npair, hipMemcpyDeviceToHost, st));
        sync(ctx);
    });
}

int pssfss_sheet_gsm(pssfss_ctx *ctx, pssfss_sheet *sheet, const pssfss_point *pt,
                     pssfss_gsm *gsm, int precision, pssfss_point_info *info) {
    return pssfss_sheet_gsm_batch(ctx, sheet, 1, pt, gsm, precision, info);
}

int pssfss_fill(pssfss_ctx *ctx, pssfss_sheet *sheet, const pssfss_point *pt, double *zmat,
                pssfss_point_info *info) {
    return guarded(ctx, [&] {
        if (!sheet || !zmat) throw ArgError("null argument");
        HIPCHECK(hipSetDevice(ctx->device));
        check_point(sheet, pt, false);
        ChunkResult cr;
        fill_chunk(ctx, sheet, pt, 1, 0, cr);
        const int64_t n = sheet->nbf;
        std::vector<cd> h((size_t)n * n);
        d2h(ctx, h.data(), ctx->wZ.p, h.size());
        sync(ctx);
        cd *o = (cd *)zmat;
        for (int64_t m = 0; m < n; ++m)
            for (int64_t s = 0; s < n; ++s) o[m + n * s] = h[m * n + s];
        if (info) {
            memset(info, 0, sizeof(*info));
            info->mmax = cr.M[0];
            info->table_converged = cr.conv[0];
        }
    });
}

int pssfss_fill_dk0(pssfss_ctx *ctx, pssfss_sheet *sheet, const pssfss_point *pt,
                    double *zmat, double *dzmat, pssfss_point_info *info) {
    return guarded(ctx, [&] {
        if (!sheet || !zmat || !dzmat) throw ArgError("null argument");
        HIPCHECK(hipSetDevice(ctx->device));
        check_point(sheet, pt, false);
        hipStream_t st = ctx->st;
        ChunkResult cr;
        TableSet tabs;
        std::vector<FillParams> fp;
        fill_chunk(ctx, sheet, pt, 1, 0, cr, &tabs, &fp);
        const int64_t n = sheet->nbf;
        std::vector<cd> h((size_t)n * n);
        d2h(ctx, h.data(), ctx->wZ.p, h.size());
        sync(ctx);
        cd *o = (cd *)zmat;
        for (int64_t m = 0; m < n; ++m)
            for (int64_t s = 0; s < n; ++s) o[m + n * s] = h[m * n + s];
        std::vector<TabInput> tin = {sheet_tab_input(sheet, pt)};
        TableSet dt;
        build_table_dk0(ctx, tin, tabs, dt);
        FillParams f = fp[0];
        const int sl = pt->s;
        const cd e1 = lay(pt->epsr, sl - 1), m1 = lay(pt->mur, sl - 1);
        const cd e2 = lay(pt->epsr, sl), m2 = lay(pt->mur, sl);
        f.jk = sheet->jk.p;   // the sheet's cache holds this point's spatial sums now
        f.dtab1 = dt.tab1(0);
        f.dtab2 = dt.tab2(0);
        const double k0 = pt->k0, u = pt->u;
        const int mag = sheet->cls == 'M' ? 1 : 0;
        if (!mag) {
            f.dc3 = c3_dk0(k0, m1, e1, m2, e2);
            f.dd3 = d3_dk0(k0, m1, e1, m2, e2);
        } else {   // derivatives of KFfact and KPfact of make_fill_params
            f.dc3 = (c3_dk0(k0, m1, e1, m1, e1) * e1 + c3_dk0(k0, m2, e2, m2, e2) * e2) /
                    ((2.0 * f.epsb) * u);
            f.dd3 = (f.mut / (2.0 * u)) *
                    (d3_dk0(k0, m1, e1, m1, e1) / m1 + d3_dk0(k0, m2, e2, m2, e2) / m2);
        }
        DBuf<cd> dZ((size_t)n * n);
        f.Z = dZ.p;
        DBuf<FillParams> dfp(1);
        dfp.upload(&f, 1, st);
        const int64_t gstride = 9 * sheet->nufp;
        ctx->wG.ensure((size_t)gstride);
        launch_fill_classes_dk0(sheet->dev(), dfp.p, sheet->rho_r.p, sheet->rinv.p, mag,
                                ctx->wG.p, st);
        // the surface-resistance term does not depend on k0
        launch_fill_gather(sheet->dev(), dfp.p, 1, sheet->bf_face.p, sheet->bf_loc.p,
                           sheet->fpcls.p, ctx->wG.p, gstride, 0, mag, st);
        d2h(ctx, h.data(), dZ.p, h.size());
        sync(ctx);
        o = (cd *)dzmat;
        for (int64_t m = 0; m < n; ++m)
            for (int64_t s = 0; s < n; ++s) o[m + n * s] = h[m * n + s];
        if (info) {
            memset(info, 0, sizeof(*info));
            info->mmax = cr.M[0];
            info->table_converged = cr.conv[0];
        }
    });
}

int pssfss_filljk(pssfss_ctx *ctx, pssfss_sheet *sheet, double u_local, double psi1,
                  double psi2, double *jk, double *rho_r, double *rinv) {
    return guarded(ctx, [&] {
        if (!sheet || !jk || !rho_r || !rinv) throw ArgError("null argument");
        if (!(u_local > 0)) throw ArgError("u_local must be positive");
        HIPCHECK(hipSetDevice(ctx->device));
        const int64_t nu = sheet->nufp;
        DBuf<cd> buf((size_t)6 * nu);
        compute_jk(ctx, sheet, u_local, {psi1, psi2}, buf.p, 0);
        ensure_singular(ctx, sheet, u_local);
        std::vector<cd> h((size_t)6 * nu);
        d2h(ctx, h.data(), buf.p, h.size());
        sync(ctx);
        flush_jk_count(ctx);
        // device order -> reference class order (ufp2fp: ifm + nface*ifs)
        const int64_t nf = sheet->nface;
        std::vector<double> rr(2 * nu), ri(nu);
        d2h(ctx, rr.data(), sheet->rho_r.p, 2 * nu);
        d2h(ctx, ri.data(), sheet->rinv.p, nu);
        sync(ctx);
        cd *o = (cd *)jk;
        for (int64_t i = 0; i < nu; ++i) {
            int64_t j = i;
            if (sheet->identity) j = (i / nf) + nf * (i % nf);
            for (int q = 0; q < 6; ++q) o[6 * j + q] = h[q * nu + i];
            rho_r[2 * j] = rr[2 * i];
            rho_r[2 * j + 1] = rr[2 * i + 1];
            rinv[j] = ri[i];
        }
    });
}

int pssfss_modal_tables(pssfss_ctx *ctx, int magnetic, const pssfss_point *pt,
                        const double *beta1, const double *beta2, const double *beta00,
                        double convtest, int *mmax, int *converged, double *tab1,
                        double *tab2) {
    return guarded(ctx, [&] {
        if (!pt || !beta1 || !beta2 || !beta00 || !tab1 || !tab2) throw ArgError("null argument");
        HIPCHECK(hipSetDevice(ctx->device));
        check_point(nullptr, pt, false);
        std::vector<TabInput> in = {make_tab_input(pt, magnetic != 0, beta1, beta2, beta00)};
        TableSet ts;
        build_tables(ctx, in, convtest, ts, ctx->st);
        const int64_t E = (int64_t)ts.M[0] + 2;
        d2h(ctx, (cd *)tab1, ts.tab1(0), (size_t)(E * E));
        d2h(ctx, (cd *)tab2, ts.tab2(0), (size_t)(E * E));
        sync(ctx);
        if (mmax) *mmax = ts.M[0];
        if (converged) *converged = ts.converged[0];
    });
}

int pssfss_rwgbfft(pssfss_ctx *ctx, pssfss_sheet *sheet, const double *k, double psi1,
                   double psi2, double *ft) {
    return guarded(ctx, [&] {
        if (!sheet || !k || !ft) throw ArgError("null argument");
        HIPCHECK(hipSetDevice(ctx->device));
        BfftJob J;
        J.k[0] = k[0];
        J.k[1] = k[1];
        J.ff[0] = cd{1, 0}; J.ff[1] = cd{1, 0}; J.ff[2] = hcis(-psi1);
        J.ff[3] = cd{1, 0}; J.ff[4] = hcis(-psi2);
        DBuf<BfftJob> dj(1);
        dj.upload(&J, 1, ctx->st);
        DBuf<cd> out((size_t)2 * sheet->nbf);
        launch_bfft(sheet->dev(), sheet->bf_face.p, sheet->bf_loc.p, dj.p, 1, out.p, ctx->st);
        d2h(ctx, (cd *)ft, out.p, (size_t)2 * sheet->nbf);
        sync(ctx);
    });
}

int pssfss_lu_solve_batch(pssfss_ctx *ctx, int64_t n, int nrhs, int nbatch, const double *A,
                          double *B, int precision, int *iters, double *resid) {
    return guarded(ctx, [&] {
        if (n <= 0 || nrhs <= 0 || nbatch <= 0 || !A || !B) throw ArgError("bad arguments");
        if (precision != PSSFSS_FP64 && precision != PSSFSS_MIXED)
            throw ArgError("precision must be PSSFSS_FP64 or PSSFSS_MIXED");
        HIPCHECK(hipSetDevice(ctx->device));
        const int64_t nb2 = n * n;
        std::vector<cd> rm((size_t)nb2 * nbatch);
        const cd *a = (const cd *)A;
        for (int b = 0; b < nbatch; ++b)
            for (int64_t i = 0; i < n; ++i)
                for (int64_t j = 0; j < n; ++j) rm[b * nb2 + i * n + j] = a[b * nb2 + i + n * j];
        ctx->wZ.upload(rm.data(), rm.size(), ctx->st);
        ctx->wB.upload((const cd *)B, (size_t)n * nrhs * nbatch, ctx->st);
        std::vector<pssfss_point_info> inf(nbatch);
        solve_chunk(ctx, n, nrhs, nbatch, precision, inf.data());
        d2h(ctx, (cd *)B, ctx->wB.p, (size_t)n * nrhs * nbatch);
        sync(ctx);
        for (int b = 0; b < nbatch; ++b) {
            if (iters) iters[b] = inf[b].refine_iters;
            if (resid) resid[b] = inf[b].refine_resid;
        }
    });
}

int pssfss_lu_solve_batch_noalloc(pssfss_ctx *ctx, int64_t n, int nrhs, int nbatch, double *A,
                                  double *B) {
    return guarded(ctx, [&] {
        g_stage_rec = nullptr;      // no stage events: nothing is created or allocated
        if (n <= 0 || nrhs <= 0 || nbatch <= 0 || !A || !B) throw ArgError("bad arguments");
        if (n > 40960) throw ArgError("n above 40960 (row permutation held in LDS)");
        auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
        const size_t nb = (size_t)nbatch;
        const size_t b_piv = up(sizeof(int32_t) * (size_t)n * nb);
        const size_t b_ws = up(sizeof(cd) * (size_t)lu_factor_ws_elems(n) * nb);
        const size_t b_t = up(sizeof(cd) * (size_t)n * (size_t)nrhs * nb);
        const size_t need = 2 * b_piv + b_ws + b_t;
        if (need > NOALLOC_SCRATCH_BYTES) {
            char msg[160];
            snprintf(msg, sizeof msg,
                     "n = %lld, nrhs = %d, nbatch = %d need %zu bytes of scratch; the limit is %zu",
                     (long long)n, nrhs, nbatch, need, NOALLOC_SCRATCH_BYTES);
            throw ArgError(msg);
        }
        HIPCHECK(hipSetDevice(ctx->device));
        unsigned char *base = (unsigned char *)ctx->scratch_dev;
        int32_t *ipiv = (int32_t *)base;
        int32_t *perm = (int32_t *)(base + b_piv);
        cd *ws = (cd *)(base + 2 * b_piv);
        cd *T = (cd *)(base + 2 * b_piv + b_ws);
        // A column-major is A^T row-major: factor that in place, then solve
        // (A^T)^T X = B with the transposed triangular solves.
        cd *a = (cd *)A;
        const int64_t nb2 = n * n;
        lu_factor_batched_ws(n, nbatch, a, nb2, ipiv, n, ws, ctx->st);
        lu_make_perm(n, nbatch, ipiv, n, perm, ctx->st);
        lu_solve_batched(n, nrhs, nbatch, a, nb2, perm, n, (cd *)B, n * nrhs, T, ctx->st, nullptr,
                         true);
        sync(ctx);
    });
}

int pssfss_cascade_batch(pssfss_ctx *ctx, int nbatch, int n1, int n2, int n3, const double *a,
                         const double *b, double *c) {
    return guarded(ctx, [&] {
        if (nbatch <= 0 || n1 <= 0 || n2 <= 0 || n3 <= 0 || !a || !b || !c)
            throw ArgError("bad arguments");
        HIPCHECK(hipSetDevice(ctx->device));
        const int64_t la = (int64_t)n1 * n1 + 2 * (int64_t)n1 * n2 + (int64_t)n2 * n2;
        const int64_t lb = (int64_t)n2 * n2 + 2 * (int64_t)n2 * n3 + (int64_t)n3 * n3;
        const int64_t lc = (int64_t)n1 * n1 + 2 * (int64_t)n1 * n3 + (int64_t)n3 * n3;
        DBuf<cd> da((size_t)la * nbatch), db((size_t)lb * nbatch), dc((size_t)lc * nbatch);
        da.upload((const cd *)a, (size_t)la * nbatch, ctx->st);
        db.upload((const cd *)b, (size_t)lb * nbatch, ctx->st);
        cascade_batched_dev(nbatch, n1, n2, n3, da.p, db.p, dc.p, ctx->st);
        d2h(ctx, (cd *)c, dc.p, (size_t)lc * nbatch);
        sync(ctx);
    });
}

int pssfss_ctx_set_chunk(pssfss_ctx *ctx, int max_points, double mem_fraction) {
    return guarded(ctx, [&] {
        if (max_points < 0 || max_points > 4096 || !(mem_fraction >= 0) || mem_fraction > 0.9)
            throw ArgError("max_points must be in 0..4096 and mem_fraction in 0..0.9");
        ctx->chunk_max = max_points > 0 ? max_points : 128;
        ctx->mem_frac = mem_fraction > 0 ? mem_fraction : 0.45;
    });
}

int pssfss_stats_enable(pssfss_ctx *ctx, int on) {
    return guarded(ctx, [&] { ctx->stats_on = on != 0; });
}

int pssfss_stats_read(pssfss_ctx *ctx, pssfss_stats *out, int reset) {
    return guarded(ctx, [&] {
        if (!out) throw ArgError("null output");
        sync(ctx);
        ctx->rec.flush();
        for (int i = 0; i < NSTAGE; ++i) {
            out->ms[i] = ctx->rec.ms[i];
            out->flops[i] = ctx->rec.flops[i];
            out->launches[i] = ctx->rec.launches[i];
        }
        if (reset) ctx->rec.reset();
    });
}

}  // extern "C"
