// common.h -- shared device helpers for the CMS-Flow implicit step on gfx950.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Value of a reference default-real (single precision) literal in the FP64
// build: the reference writes e.g. relax(2)=0.8 and max(dk,1.e-6) without the
// _ikind suffix, so an ikind=8 build evaluates them as float then widens.
#define F4(x) ((double)(float)(x))

namespace cms {

constexpr int KN = 0, KE = 1, KS = 2, KW = 3;
constexpr int BLOCK = 256;
// Threads per block of the kernels on the fixed Krylov grid (8 waves per
// SIMD with 4 blocks per CU).
constexpr int KBLOCK = 512;
constexpr double GRAV = 9.81;                       // 9.81_ikind, flow.F90:35
constexpr double GRAVINV = 1.0 / 9.81;              // flow.F90:37

// Geometry of a uniform Cartesian grid with a one-cell ghost ring.  Cell (i,j),
// i in [-1,nx], j in [-1,ny], lives at c = (j+1)*nxp + (i+1); active cells have
// 0<=i<nx, 0<=j<ny.  Per-direction constants follow geo_init_cart
// (geospatial.F90:1469-1611) evaluated on a uniform grid.
struct Geo {
    int nx, ny, nxp, nyp;
    int nact;                 // nx*ny
    double dx, dy, area;
    double ds[4], dn[4], dsxy[4], fnx[4], fny[4], fintp[4], rx[4], ry[4], dsx[4], dsy[4];
    int side[4];              // boundary type per side (CMS2D_BND_*)
};

__device__ __forceinline__ int cidx(const Geo &g, int i, int j) { return (j + 1) * g.nxp + (i + 1); }
__device__ __forceinline__ int nbr(const Geo &g, int c, int k) {
    return k == KN ? c + g.nxp : k == KE ? c + 1 : k == KS ? c - g.nxp : c - 1;
}
__device__ __forceinline__ bool active_ij(const Geo &g, int i, int j) {
    return i >= 0 && i < g.nx && j >= 0 && j < g.ny;
}

// Block index remapped so that consecutive logical blocks (adjacent rows of the
// grid) run on the same XCD: the hardware deals workgroups to the 8 XCDs
// round-robin, and each XCD has its own L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    constexpr int NXCD = 8;
    int q = nb / NXCD, r = nb % NXCD;
    int x = b % NXCD, k = b / NXCD;
    // blocks of XCD x: logical range [x*q + min(x,r), ...) of length q + (x<r)
    int base = x * q + (x < r ? x : r);
    int len = q + (x < r ? 1 : 0);
    return k < len ? base + k : b;
}

// Active-cell thread mapping: returns false for threads past the end.
__device__ __forceinline__ bool act_thread(const Geo &g, int &i, int &j, int &c) {
    int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= g.nact) return false;
    j = t / g.nx;
    i = t - j * g.nx;
    c = cidx(g, i, j);
    return true;
}

// --- deterministic reductions -------------------------------------------------
template <int NV, int BS = BLOCK>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[NV][BS / 64];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double x = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[k][w] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = sh[k][0];
#pragma unroll
        for (int q = 1; q < BS / 64; q++) s += sh[k][q];
        v[k] = s;
    }
}

}  // namespace cms
