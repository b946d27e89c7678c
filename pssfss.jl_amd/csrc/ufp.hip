// Unique-face-pair classes of setup_rwg (src/RWG.jl:270-322) on the device.
// Face pair mn = m + nface n (m match face, n source face) is characterised by the
// three vectors from the match centroid to the source vertices; pairs whose
// vectors agree belong to one class (source triangles superimposed by a
// translation, same vertex order, same relative match position).  The vectors are
// quantised (step 64e-6 |c1 - c2|), mixed into a 64-bit key and radix-sorted
// (stable: members stay ascending); every member is then checked against its
// class's first member, exactly on the quantised vector and within the
// reference's radius 1e-6 |c1 - c2| on the unquantised one.  Classes are numbered
// by their smallest member, the order of the reference's first-appearance loop.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "internal.hpp"

namespace {

struct PairGeom {
    const double *c;      // nface x 2 centroids
    const double *v;      // nface x 6 vertices (x, y) x 3
    int64_t nface;
    double q, radius;
};

__device__ __forceinline__ void pair_vec(const PairGeom &g, int64_t mn, double d[6]) {
    const int64_t m = mn % g.nface, n = mn / g.nface;
    const double cx = g.c[2 * m], cy = g.c[2 * m + 1];
    const double *v = g.v + 6 * n;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        d[2 * k] = cx - v[2 * k];
        d[2 * k + 1] = cy - v[2 * k + 1];
    }
}

__device__ __forceinline__ void pair_key(const PairGeom &g, int64_t mn, int64_t key[6]) {
    double d[6];
    pair_vec(g, mn, d);
#pragma unroll
    for (int j = 0; j < 6; ++j) key[j] = (int64_t)rint(d[j] / g.q);
}

__global__ void k_face_geom(int64_t nface, const double *rho, const int64_t *fv, double *c,
                            double *v) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nface) return;
    double sx = 0, sy = 0;
    for (int k = 0; k < 3; ++k) {
        const int64_t node = fv[3 * f + k] - 1;
        const double x = rho[2 * node], y = rho[2 * node + 1];
        v[6 * f + 2 * k] = x;
        v[6 * f + 2 * k + 1] = y;
        if (k == 0) { sx = x; sy = y; } else { sx = sx + x; sy = sy + y; }
    }
    c[2 * f] = sx / 3.0;
    c[2 * f + 1] = sy / 3.0;
}

__global__ void k_pair_hash(PairGeom g, int64_t npair, uint64_t *h, uint32_t *idx) {
    const int64_t mn = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (mn >= npair) return;
    int64_t key[6];
    pair_key(g, mn, key);
    const uint64_t mult[6] = {0x9E3779B97F4A7C15ull, 0xC2B2AE3D27D4EB4Full, 0x165667B19E3779F9ull,
                              0xD6E8FEB86659FD93ull, 0xFF51AFD7ED558CCDull, 0xC4CEB9FE1A85EC53ull};
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) s += (uint64_t)key[j] * mult[j];
    s ^= s >> 31;
    h[mn] = s;
    idx[mn] = (uint32_t)mn;
}

// head position of each sorted element's run (0 for non-heads before the scan)
__global__ void k_heads(int64_t npair, const uint64_t *hs, const uint32_t *is, uint32_t *headpos,
                        uint32_t *first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npair) return;
    const bool head = i == 0 || hs[i] != hs[i - 1];
    headpos[i] = head ? (uint32_t)i : 0u;
    if (head) first[is[i]] = 1u;
}

// every member equal to its run's first pair (exact key, reference radius)
__global__ void k_verify(PairGeom g, int64_t npair, const uint32_t *is, const uint32_t *runhead,
                         unsigned long long *bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npair) return;
    const int64_t a = is[i], b = is[runhead[i]];
    if (a == b) return;
    int64_t ka[6], kb[6];
    pair_key(g, a, ka);
    pair_key(g, b, kb);
    double da[6], db[6];
    pair_vec(g, a, da);
    pair_vec(g, b, db);
    double d2 = 0;
    bool same = true;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        same = same && ka[j] == kb[j];
        d2 += (da[j] - db[j]) * (da[j] - db[j]);
    }
    if (!same || !(sqrt(d2) < g.radius)) atomicAdd(bad, 1ull);
}

// class of each run from the rank of its first member; run lengths per class
__global__ void k_classes(int64_t npair, const uint32_t *is, const uint32_t *runhead,
                          const uint32_t *rank, int64_t *ufpm, uint32_t *count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npair) return;
    const uint32_t cls = rank[is[runhead[i]]];
    ufpm[is[i]] = (int64_t)cls + 1;
    if (runhead[i] == (uint32_t)i) {
        int64_t e = i + 1;
        // run length: scan forward to the next head (runs are short in practice)
        while (e < npair && runhead[e] == (uint32_t)i) ++e;
        count[cls] = (uint32_t)(e - i);
    }
}

__global__ void k_members(int64_t npair, const uint32_t *is, const uint32_t *runhead,
                          const uint32_t *rank, const uint64_t *offs, int64_t *members) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npair) return;
    const uint32_t h = runhead[i];
    const uint32_t cls = rank[is[h]];
    members[offs[cls] + (i - h)] = (int64_t)is[i];
}

struct MaxOp {
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const {
        return a > b ? a : b;
    }
};

unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// Returns nufp; offsets (nufp + 1), members (npair) and ufpm (npair, 1-based) are
// device arrays of the caller's buffers.
int64_t face_pair_classes_dev(int64_t nnode, const double *d_rho, int64_t nface,
                              const int64_t *d_fv, double radius, int64_t *d_ufpm,
                              int64_t *d_members, uint64_t *d_offsets_u64, hipStream_t st) {
    const int64_t npair = nface * nface;
    if (nface > 65535) throw ArgError("face_pair_classes: at most 65535 faces");
    (void)nnode;
    DBuf<double> c(2 * nface), v(6 * nface);
    hipLaunchKernelGGL(k_face_geom, dim3(grid_of(nface)), dim3(256), 0, st, nface, d_rho, d_fv,
                       c.p, v.p);
    HIPCHECK(hipGetLastError());
    PairGeom g{c.p, v.p, nface, 64.0 * radius, radius};
    DBuf<uint64_t> h(npair), hs(npair);
    DBuf<uint32_t> idx(npair), is(npair);
    hipLaunchKernelGGL(k_pair_hash, dim3(grid_of(npair)), dim3(256), 0, st, g, npair, h.p, idx.p);
    HIPCHECK(hipGetLastError());
    // runs: head positions propagated by a max-scan; first members flagged
    // (first is padded to whole 16-byte words for the zero fill)
    DBuf<uint32_t> headpos(npair), runhead(npair), first(npair + 4), rank(npair);
    // one temporary buffer sized for the sort and both scans, allocated before any
    // of them runs
    size_t b_sort = 0, b_max = 0, b_sum = 0;
    HIPCHECK(rocprim::radix_sort_pairs(nullptr, b_sort, h.p, hs.p, idx.p, is.p, (size_t)npair, 0,
                                       64, st));
    HIPCHECK(rocprim::inclusive_scan(nullptr, b_max, headpos.p, runhead.p, (size_t)npair,
                                     MaxOp(), st));
    HIPCHECK(rocprim::exclusive_scan(nullptr, b_sum, first.p, rank.p, 0u, (size_t)npair,
                                     rocprim::plus<uint32_t>(), st));
    size_t tb = std::max(b_sort, std::max(b_max, b_sum));
    DBuf<char> tmp(tb + 16);
    HIPCHECK(rocprim::radix_sort_pairs(tmp.p, b_sort, h.p, hs.p, idx.p, is.p, (size_t)npair, 0,
                                       64, st));
    launch_zero((cd *)first.p, (npair * 4 + 15) / 16, st);
    hipLaunchKernelGGL(k_heads, dim3(grid_of(npair)), dim3(256), 0, st, npair, hs.p, is.p,
                       headpos.p, first.p);
    HIPCHECK(hipGetLastError());
    HIPCHECK(rocprim::inclusive_scan(tmp.p, b_max, headpos.p, runhead.p, (size_t)npair, MaxOp(),
                                     st));
    DBuf<unsigned long long> bad(2);
    launch_zero((cd *)bad.p, 1, st);
    hipLaunchKernelGGL(k_verify, dim3(grid_of(npair)), dim3(256), 0, st, g, npair, is.p,
                       runhead.p, bad.p);
    HIPCHECK(hipGetLastError());
    // class numbers: exclusive scan of the first-member flags in pair order
    HIPCHECK(rocprim::exclusive_scan(tmp.p, b_sum, first.p, rank.p, 0u, (size_t)npair,
                                     rocprim::plus<uint32_t>(), st));
    uint32_t last_rank = 0, last_first = 0;
    HIPCHECK(hipMemcpyAsync(&last_rank, rank.p + npair - 1, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&last_first, first.p + npair - 1, 4, hipMemcpyDeviceToHost, st));
    unsigned long long nbad = 0;
    HIPCHECK(hipMemcpyAsync(&nbad, bad.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (nbad)
        throw ArgError("unique-face-pair classes are not within the grouping radius; mesh has "
                       "near-coincident face pairs at the quantization scale");
    const int64_t nufp = (int64_t)last_rank + last_first;
    DBuf<uint32_t> count(nufp);
    hipLaunchKernelGGL(k_classes, dim3(grid_of(npair)), dim3(256), 0, st, npair, is.p,
                       runhead.p, rank.p, d_ufpm, count.p);
    HIPCHECK(hipGetLastError());
    // offsets = exclusive scan of the class sizes (64-bit), offsets[nufp] = npair
    {
        std::vector<uint32_t> hc(nufp);
        HIPCHECK(hipMemcpyAsync(hc.data(), count.p, 4 * nufp, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        std::vector<uint64_t> ho(nufp + 1);
        ho[0] = 0;
        for (int64_t i = 0; i < nufp; ++i) ho[i + 1] = ho[i] + hc[i];
        if ((int64_t)ho[nufp] != npair) throw HipError("face_pair_classes: class sizes mismatch");
        HIPCHECK(hipMemcpyAsync(d_offsets_u64, ho.data(), 8 * (nufp + 1), hipMemcpyHostToDevice,
                                st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    hipLaunchKernelGGL(k_members, dim3(grid_of(npair)), dim3(256), 0, st, npair, is.p,
                       runhead.p, rank.p, d_offsets_u64, d_members);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    return nufp;
}
