// d3q15.hpp -- D3Q15 lattice tables and small device helpers shared by the
// FP64 and mixed-precision kernels.
//
// Tables follow CPU/common.hpp:3-45: c_i for i = 0 (rest), 1..6 (+x,+y,+z,
// -x,-y,-z), 7..14 (the 8 diagonals); weights E = 2/9, 1/9, 1/72;
// H = (1, 0, ...); F = (-7/3, 1/3 x6, 1/24 x8).  All loops over q are fully
// unrolled, so every table entry becomes an immediate.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace lbm {

constexpr int Q = 15;
constexpr int CX[Q] = {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1};
constexpr int CY[Q] = {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1};
constexpr int CZ[Q] = {0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1};
constexpr double UC[Q] = {0.0, 1.0, 0.0, 0.0, -1.0, 0.0, 0.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0};
constexpr double VC[Q] = {0.0, 0.0, 1.0, 0.0, 0.0, -1.0, 0.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0};
constexpr double WC[Q] = {0.0, 0.0, 0.0, 1.0, 0.0, 0.0, -1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0, 1.0};
constexpr double EW[Q] = {2.0 / 9.0,
                          1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                          1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0,
                          1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
constexpr double HW[Q] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
constexpr double FW[Q] = {-7.0 / 3.0,
                          1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0,
                          1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0,
                          1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0};

// Specular x-wall pairs (+x direction, -x direction), CPU/Inamuro.cpp:592-593.
constexpr int SLIP_POS[5] = {1, 7, 9, 10, 12};
constexpr int SLIP_NEG[5] = {4, 8, 14, 13, 11};
// Opposite direction, c_opp = -c (bounce-back at the optional z walls).
constexpr int OPP[Q] = {0, 4, 5, 6, 1, 2, 3, 11, 12, 13, 14, 7, 8, 9, 10};

constexpr int BLOCK = 256;       // threads per block: 4 wave64s
constexpr int RED_BLOCKS = 1024; // fixed grid of the deterministic reductions

struct Grid {
    int lx, ly, lz;
    int64_t n;      // lx * ly * lz
    int zwall = 0;  // 1: no-slip walls at z = 0 and lz-1 (extension); 0: periodic z
};

// Cell coordinates of linear index c = (z*ly + y)*lx + x.
__device__ __forceinline__ void cell_xyz(int64_t c, const Grid& g, int& x, int& y, int& z)
{
    const uint32_t cc = static_cast<uint32_t>(c);
    const uint32_t t = cc / static_cast<uint32_t>(g.lx);
    x = static_cast<int>(cc - t * static_cast<uint32_t>(g.lx));
    const uint32_t zz = t / static_cast<uint32_t>(g.ly);
    y = static_cast<int>(t - zz * static_cast<uint32_t>(g.ly));
    z = static_cast<int>(zz);
}

// Periodic neighbour of coordinate a in direction d (-1, 0, +1).
__device__ __forceinline__ int wrap(int a, int d, int l)
{
    int r = a + d;
    r = r < 0 ? r + l : r;
    r = r >= l ? r - l : r;
    return r;
}

// Stencil neighbour on x: specular mirror lx -> lx-2, -1 -> 1
// (CPU/Inamuro.hpp:234-238).
__device__ __forceinline__ int mirror_x(int x, int d, int lx)
{
    int r = x + d;
    r = r >= lx ? lx - 2 : r;
    r = r < 0 ? 1 : r;
    return r;
}

// z neighbour of the stencils: periodic, or (z walls) the cell itself beyond
// a wall, i.e. zero normal gradient about a wall half a cell outside.
__device__ __forceinline__ int z_stencil(int z, int d, const Grid& g)
{
    if (g.zwall) {
        const int r = z + d;
        return r < 0 ? 0 : (r >= g.lz ? g.lz - 1 : r);
    }
    return wrap(z, d, g.lz);
}

// Halfway bounce-back at the z walls: the population arriving at z in
// direction q would come from beyond a wall (pull side) ...
__device__ __forceinline__ bool z_bounced_in(int q, int z, const Grid& g)
{
    return g.zwall && ((CZ[q] == 1 && z == 0) || (CZ[q] == -1 && z == g.lz - 1));
}
// ... or the population leaving z in direction q would cross a wall (push side).
__device__ __forceinline__ bool z_bounced_out(int q, int z, const Grid& g)
{
    return g.zwall && ((CZ[q] == 1 && z == g.lz - 1) || (CZ[q] == -1 && z == 0));
}

__device__ __forceinline__ int64_t lin(int x, int y, int z, const Grid& g)
{
    return (static_cast<int64_t>(z) * g.ly + y) * g.lx + x;
}

// The collision kernels' tile: STENCIL_TX x STENCIL_TY cells (one wave64 per
// x-row segment) marching through STENCIL_ZCHUNK z layers.  The five stencil
// fields of each layer are staged in LDS once, with a one-cell halo in x and
// y, in a ring of three planes (layers z-1, z, z+1 of the cell being
// collided).
constexpr int STENCIL_TX = 64, STENCIL_TY = 4, STENCIL_ZCHUNK = 16;
constexpr int STENCIL_SX = STENCIL_TX + 2, STENCIL_SY = STENCIL_TY + 2;
constexpr int STENCIL_PLANE = STENCIL_SX * STENCIL_SY;

// Global cell held at slot e of a staged plane of layer gz for the tile at
// (x0, y0): x mirrored at the walls (as mirror_x), y periodic.
__device__ __forceinline__ int64_t stencil_source(int e, int x0, int y0, int gz, const Grid& g)
{
    const int sx = e % STENCIL_SX, sy = e / STENCIL_SX;
    int gx = x0 - 1 + sx;
    gx = gx < 0 ? 1 : (gx >= g.lx ? g.lx - 2 : gx);
    const int gy = (y0 - 1 + sy + g.ly) % g.ly;
    return lin(gx, gy, gz, g);
}

// Slot of neighbour k (k = 0: the cell) of tile cell (tx, ty) within its plane.
__device__ __forceinline__ int stencil_slot(int tx, int ty, int k)
{
    return (tx + 1 + CX[k]) + STENCIL_SX * (ty + 1 + CY[k]);
}

// Deterministic wave64 butterfly sum.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Deterministic sum over a 256-thread block; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds4)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    if (lane == 0) lds4[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = ((lds4[0] + lds4[1]) + lds4[2]) + lds4[3];
    __syncthreads();
    return r;
}

} // namespace lbm
