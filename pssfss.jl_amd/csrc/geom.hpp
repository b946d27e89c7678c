// Quadrature rule and lattice ring enumeration shared by the spatial-sum and
// adjoint kernels.
//   7-point rule of src/Zint.jl:55-62 (weights sum to 1/2, the reference
//   triangle's area)
//   ring order of src/Rings.jl:9-63
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__constant__ double QXI[7] = {0.33333333333333330, 0.10128650732345633, 0.79742698535308730,
                              0.10128650732345633, 0.47014206410511505, 0.05971587178976989,
                              0.47014206410511505};
__constant__ double QETA[7] = {0.33333333333333333, 0.79742698535308730, 0.10128650732345633,
                               0.10128650732345633, 0.05971587178976989, 0.47014206410511505,
                               0.47014206410511505};
__constant__ double QW[7] = {0.1125, 0.06296959027241358, 0.06296959027241358,
                             0.06296959027241358, 0.06619707639425308, 0.06619707639425308,
                             0.06619707639425308};

// Ring(r) element k: m=r column, m=-r column (n ascending), then n=-r and n=r rows
// (m ascending, corners excluded) -- the order of src/Rings.jl:9-63.
__device__ __forceinline__ void ring_mn(int r, int k, int &m, int &n) {
    if (r == 0) { m = 0; n = 0; return; }
    int side = 2 * r + 1;
    if (k < side) { m = r; n = -r + k; return; }
    k -= side;
    if (k < side) { m = -r; n = -r + k; return; }
    k -= side;
    if (k < 2 * r - 1) { m = -r + 1 + k; n = -r; return; }
    k -= 2 * r - 1;
    m = -r + 1 + k;
    n = r;
}

__device__ __forceinline__ int64_t ring_base(int r) {
    return r == 0 ? 0 : (int64_t)(2 * r - 1) * (2 * r - 1);
}

__device__ __forceinline__ int ring_of(int64_t e) {
    if (e == 0) return 0;
    int r = (int)((sqrt((double)e) + 1.0) * 0.5);
    while (ring_base(r + 1) <= e) ++r;
    while (r > 0 && ring_base(r) > e) --r;
    return r;
}

__device__ __forceinline__ int64_t ring_index(int m, int n) {
    int r = max(abs(m), abs(n));
    if (r == 0) return 0;
    int64_t b = ring_base(r);
    int side = 2 * r + 1;
    if (m == r) return b + (n + r);
    if (m == -r) return b + side + (n + r);
    if (n == -r) return b + 2 * side + (m + r - 1);
    return b + 2 * side + (2 * r - 1) + (m + r - 1);
}

}  // namespace
