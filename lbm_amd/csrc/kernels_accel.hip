// kernels_accel.hip -- the reference GPU solver's optional Poisson
// accelerators (GPU/InamuroCUDA.hpp:60-76), FP64.
//
// They leave the strict contract (WHILE_BASELINE.md:3-13), so they run on
// their own state layout: h is held explicitly after the x-slip rule
// (post-boundary, [q][z][y][x]) and p is written by every iteration, as in
// the reference.  Per-cell arithmetic keeps the reference's association
// (this unit is built with -ffp-contract=off):
//   image  (GPU/InamuroCUDA.cu:539-583, 778-846): h_old - s (h_old - E p)/tau_h
//          - (E/3) S at the source cell x - c_q, pulled, then the x-slip rule
//   onepass: h_old + w (image - h_old), p = sum_q
//   scalar (:868-908): (sum_{q=1..14} p(x - c_qb) - 5 scale rho S) / 14
//   Anderson(1) (:930-1013, 2238-2280): the dot products are block partials
//          of the onepass kernel itself and the mixing coefficient is formed
//          on the device, so an iteration needs no host round trip
//   two-grid (:1198-1323): 4^3 block sums of p - p_prev, one wave64 per block
//   source-aware start (:905-928)
// Every global sum is a fixed-shape tree, so results are deterministic.
#include "kernels.hpp"

namespace lbm {
namespace {

__device__ __forceinline__ double reduce_parts(const double* __restrict__ parts, int count, int stride,
                                               int offset, double* lds)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < count; i += BLOCK) acc += parts[i * stride + offset];
    return block_sum(acc, lds);
}

// p, tau_h and S of one z layer and its one-cell x/y halo, staged in LDS in
// a ring of three layers (the collision kernels' tile: STENCIL_TX x
// STENCIL_TY cells marching zc layers; see pull_zchunk).  The x halo is the
// stencils' mirror (x = -1 -> 1, x = lx -> lx-2); y and z wrap.  Loads of the
// next layer go to registers before the current layer is computed and land
// in LDS after it.
template <int NF>
struct AccStage {
    static constexpr int PER = (STENCIL_PLANE + BLOCK - 1) / BLOCK;
    double next[NF][PER];
    __device__ __forceinline__ void fetch(const double* const (&src)[NF], int x0, int y0, int gz, const Grid& g)
    {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = threadIdx.x + j * BLOCK;
            if (e < STENCIL_PLANE) {
                const int64_t i = stencil_source(e, x0, y0, gz, g);
#pragma unroll
                for (int f = 0; f < NF; ++f) next[f][j] = src[f][i];
            }
        }
    }
    __device__ __forceinline__ void land(double (&staged)[NF][3][STENCIL_PLANE], int slot)
    {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = threadIdx.x + j * BLOCK;
            if (e < STENCIL_PLANE) {
#pragma unroll
                for (int f = 0; f < NF; ++f) staged[f][slot][e] = next[f][j];
            }
        }
    }
};

// LDS slot (ring slot times plane, plus place in the plane) of the source
// x - c_q of tile cell (tx, ty) in the layer whose ring slots are sm, s0, sp
// (z - 1, z, z + 1); cx is the x offset actually taken (c_x of q, or -c_x
// for the mirror image).
__device__ __forceinline__ int pull_slot(int q, int cx, int tx, int ty, int sm, int s0, int sp)
{
    const int ring = CZ[q] == 1 ? sm : (CZ[q] == -1 ? sp : s0);
    return ring * STENCIL_PLANE + (tx + 1 - cx) + STENCIL_SX * (ty + 1 - CY[q]);
}

// Slip partner of direction q (the same c_y, c_z and weight, c_x reversed).
__device__ __forceinline__ constexpr int slip_partner(int q)
{
    return q == 1 ? 4 : q == 4 ? 1 : q == 7 ? 8 : q == 8 ? 7 : q == 9 ? 14 : q == 14 ? 9
         : q == 10 ? 13 : q == 13 ? 10 : q == 12 ? 11 : q == 11 ? 12 : q;
}

// One iteration of the pull form.  ONEPASS: fixed-point relaxation against
// the cell's own h.  AND (onepass only): 0 none; 1 the first Anderson
// iteration, which also stores the history (residual image - h, image);
// 2 later iterations, which add the block partials of <R_prev, R - R_prev>
// and |R - R_prev|^2 and leave p to the mixing kernel.  The 15 source cells'
// p, tau_h and S come from LDS; the source populations from HBM.  The x-slip
// rule is folded into each direction: at x = 0 (x = lx-1) a direction whose
// source lies beyond the wall takes the image of its slip partner, i.e. the
// partner's population pulled from the mirror cell x + c_x; the image the
// reference's periodic pull forms for it is overwritten, so it is not formed.
// Each direction is then independent, and every access is a per-direction
// base (uniform across the block) plus a 32-bit cell index (grids hold at
// most 2^31 - 1 cells), which keeps the register count low.
template <int ONEPASS, int AND>
__global__ __launch_bounds__(BLOCK) void k_acc_pull(
    Grid g, int zc, double scale, double omega, const double* __restrict__ h, const double* __restrict__ p,
    const double* __restrict__ tau, const double* __restrict__ src, double* __restrict__ hn,
    double* __restrict__ pn, double* __restrict__ pr, double* __restrict__ pi,
    double* __restrict__ parts)
{
    __shared__ double lds[4];
    __shared__ double staged[3][3][STENCIL_PLANE];  // p, tau_h, S; ring of 3 layers
    const int ntx = (g.lx + STENCIL_TX - 1) / STENCIL_TX, nty = (g.ly + STENCIL_TY - 1) / STENCIL_TY;
    const int b = static_cast<int>(blockIdx.x);
    const int x0 = (b % ntx) * STENCIL_TX, y0 = ((b / ntx) % nty) * STENCIL_TY;
    const int zb = (b / (ntx * nty)) * zc, ze = min(g.lz, zb + zc);
    const double* const fields[3] = {p, tau, src};
    AccStage<3> st;
    st.fetch(fields, x0, y0, wrap(zb, -1, g.lz), g);
    st.land(staged, 0);
    st.fetch(fields, x0, y0, zb, g);
    st.land(staged, 1);
    st.fetch(fields, x0, y0, wrap(zb, 1, g.lz), g);
    const int tx = threadIdx.x % STENCIL_TX, ty = threadIdx.x / STENCIL_TX;
    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < g.lx && y < g.ly;
    const int xm = wrap(x, -1, g.lx), xp = wrap(x, 1, g.lx);
    const int ym = wrap(y, -1, g.ly), yp = wrap(y, 1, g.ly);
    double num = 0.0, den = 0.0;
    for (int z = zb; z < ze; ++z) {
        const int sm = (z - zb) % 3, s0 = (z - zb + 1) % 3, sp = (z - zb + 2) % 3;
        st.land(staged, sp);
        __syncthreads();
        if (z + 1 < ze) st.fetch(fields, x0, y0, wrap(z + 1, 1, g.lz), g);
        if (valid) {
            const int zm = wrap(z, -1, g.lz), zp = wrap(z, 1, g.lz);
            const int64_t c = lin(x, y, z, g);
            double p_sum = 0.0;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const bool mirrored = (x == 0 && CX[q] == 1) || (x == g.lx - 1 && CX[q] == -1);
                const int cx = mirrored ? -CX[q] : CX[q];
                const int qa = mirrored ? slip_partner(q) : q;
                const int sx = cx == 1 ? xm : (cx == -1 ? xp : x);
                const int sy = CY[q] == 1 ? ym : (CY[q] == -1 ? yp : y);
                const int sz = CZ[q] == 1 ? zm : (CZ[q] == -1 ? zp : z);
                const int slot = pull_slot(q, cx, tx, ty, sm, s0, sp);
                const double tauh = staged[1][0][slot];
                const double hequ = EW[q] * staged[0][0][slot];
                const uint32_t src_cell = static_cast<uint32_t>((sz * g.ly + sy) * g.lx + sx);
                const double h_old = (h + qa * g.n)[src_cell];
                const uint32_t cc = static_cast<uint32_t>(c);
                double v = h_old - scale * (h_old - hequ) / tauh - (EW[q] / 3.0) * staged[2][0][slot];
                double h_own = 0.0;
                if (ONEPASS) {
                    h_own = (h + q * g.n)[cc];
                    v = h_own + omega * (v - h_own);
                }
                (hn + q * g.n)[cc] = v;
                p_sum += v;
                if (AND == 1) {
                    (pr + q * g.n)[cc] = v - h_own;
                    (pi + q * g.n)[cc] = v;
                } else if (AND == 2) {
                    const double residual = v - h_own;
                    const double prev = (pr + q * g.n)[cc];
                    const double delta = residual - prev;
                    num += prev * delta;
                    den += delta * delta;
                }
            }
            if (AND != 2) pn[c] = p_sum;
        }
        __syncthreads();  // the ring slot of layer z - 1 is overwritten next
    }
    if (AND == 2) {
        const double sn = block_sum(num, lds);
        const double sd = block_sum(den, lds);
        if (threadIdx.x == 0) {
            parts[2 * blockIdx.x] = sn;
            parts[2 * blockIdx.x + 1] = sd;
        }
    }
}

// The mixing coefficient (GPU/InamuroCUDA.cu:2254-2266):
// alpha = -num / den when den > 1e-300, else 1; 1 if not finite; clamped to
// [-beta_max, 1 + beta_max].
__global__ __launch_bounds__(BLOCK) void k_acc_alpha(const double* __restrict__ parts, int nparts,
                                                     double beta_max, DevState* st)
{
    __shared__ double lds[4];
    const double num = reduce_parts(parts, nparts, 2, 0, lds);
    const double den = reduce_parts(parts, nparts, 2, 1, lds);
    if (threadIdx.x != 0) return;
    double alpha = 1.0;
    if (den > 1.0e-300) alpha = -num / den;
    if (!isfinite(alpha)) alpha = 1.0;
    const double lo = -beta_max;
    const double hi = 1.0 + beta_max;
    if (alpha > hi) alpha = hi;
    if (alpha < lo) alpha = lo;
    st->acc_alpha = alpha;
}

// andersonM1ApplyKernel (GPU/InamuroCUDA.cu:983-1013): the mixed state
// (1 - alpha) I_prev + alpha I goes into the previous image's buffer, which
// becomes the current h; I becomes the previous image without a copy.
__global__ __launch_bounds__(BLOCK) void k_acc_apply(Grid g, const DevState* __restrict__ st,
                                                     const double* __restrict__ h,
                                                     const double* __restrict__ img,
                                                     double* __restrict__ pi_to_a,
                                                     double* __restrict__ pr, double* __restrict__ pn)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    const double alpha = st->acc_alpha;
    double p_sum = 0.0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int64_t k = q * g.n + c;
        const double current_image = img[k];
        const double accelerated = (1.0 - alpha) * pi_to_a[k] + alpha * current_image;
        pr[k] = current_image - h[k];
        pi_to_a[k] = accelerated;
        p_sum += accelerated;
    }
    pn[c] = p_sum;
}

// scalarPressureJacobiKernel (GPU/InamuroCUDA.cu:848-908): the 14 moving
// neighbours x - c_q in q order, p staged in LDS like k_acc_pull's fields.
// pressure_bounce_direction replaces a direction whose source lies beyond an
// x wall by its slip partner, which differs only in c_x, so the neighbour is
// the mirror image: exactly what the staged halo holds there.
__global__ __launch_bounds__(BLOCK) void k_acc_scalar(Grid g, int zc, double scale, const double* __restrict__ p,
                                                      const double* __restrict__ rho,
                                                      const double* __restrict__ src,
                                                      double* __restrict__ pn)
{
    __shared__ double staged[1][3][STENCIL_PLANE];
    const int ntx = (g.lx + STENCIL_TX - 1) / STENCIL_TX, nty = (g.ly + STENCIL_TY - 1) / STENCIL_TY;
    const int b = static_cast<int>(blockIdx.x);
    const int x0 = (b % ntx) * STENCIL_TX, y0 = ((b / ntx) % nty) * STENCIL_TY;
    const int zb = (b / (ntx * nty)) * zc, ze = min(g.lz, zb + zc);
    const double* const fields[1] = {p};
    AccStage<1> st;
    st.fetch(fields, x0, y0, wrap(zb, -1, g.lz), g);
    st.land(staged, 0);
    st.fetch(fields, x0, y0, zb, g);
    st.land(staged, 1);
    st.fetch(fields, x0, y0, wrap(zb, 1, g.lz), g);
    const int tx = threadIdx.x % STENCIL_TX, ty = threadIdx.x / STENCIL_TX;
    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < g.lx && y < g.ly;
    for (int z = zb; z < ze; ++z) {
        const int sm = (z - zb) % 3, s0 = (z - zb + 1) % 3, sp = (z - zb + 2) % 3;
        st.land(staged, sp);
        __syncthreads();
        if (z + 1 < ze) st.fetch(fields, x0, y0, wrap(z + 1, 1, g.lz), g);
        if (valid) {
            const int64_t c = lin(x, y, z, g);
            double neighbor_sum = 0.0;
#pragma unroll
            for (int q = 1; q < Q; ++q) neighbor_sum += staged[0][0][pull_slot(q, CX[q], tx, ty, sm, s0, sp)];
            const double rhs = scale * rho[c] * src[c];
            pn[c] = (neighbor_sum - 5.0 * rhs) / 14.0;
        }
        __syncthreads();
    }
}

// sourceAwareHHInitKernel (GPU/InamuroCUDA.cu:905-928).
__global__ __launch_bounds__(BLOCK) void k_acc_source_init(Grid g, double scale,
                                                           const double* __restrict__ p,
                                                           const double* __restrict__ tau,
                                                           const double* __restrict__ src,
                                                           double* __restrict__ h)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    const double tauh = tau[c];
    const double div_u = src[c];
    const double loc_p = p[c];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double hequ = EW[q] * loc_p;
        const double source = (EW[q] / 3.0) * div_u;
        h[q * g.n + c] = hequ - scale * tauh * source;
    }
}

__global__ __launch_bounds__(BLOCK) void k_acc_sum(Grid g, const double* __restrict__ a,
                                                   double* __restrict__ parts)
{
    __shared__ double lds[4];
    double acc = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x; i < g.n;
         i += static_cast<int64_t>(gridDim.x) * BLOCK)
        acc += a[i];
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// The gauge's h half (applyPressureGaugeKernel, GPU/InamuroCUDA.cu:716-730).
__global__ __launch_bounds__(BLOCK) void k_acc_gauge_h(Grid g, const DevState* __restrict__ st,
                                                       double* __restrict__ h)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    const double shift = st->shift;
#pragma unroll
    for (int q = 0; q < Q; ++q) h[q * g.n + c] -= EW[q] * shift;
}

// Gauge of p, relative change and p_prev <- p (as k_gauge_error) together
// with the 4^3 block sums of p - p_prev (pressureErrorSpatialKernel,
// GPU/InamuroCUDA.cu:1198-1248).  A block of 4^3 cells is one wave64, lane
// (i, j, k) = (lane & 3, lane >> 2 & 3, lane >> 4); block b = (bz * by + by)
// * bx + bx, the reference's numbering.
__global__ __launch_bounds__(BLOCK) void k_acc_gauge_error_blocks(Grid g, int nbx, int nby, int nblocks,
                                                                  double* __restrict__ p,
                                                                  double* __restrict__ p_prev,
                                                                  const DevState* __restrict__ st,
                                                                  double* __restrict__ blocks,
                                                                  double* __restrict__ parts)
{
    __shared__ double lds[4];
    const double shift = st->shift;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    double e1 = 0.0, e2 = 0.0;
    for (int b = blockIdx.x * 4 + wave; b < nblocks; b += gridDim.x * 4) {
        const int bxi = b % nbx;
        const int t = b / nbx;
        const int byi = t % nby;
        const int bzi = t / nby;
        const int x = bxi * 4 + (lane & 3);
        const int y = byi * 4 + ((lane >> 2) & 3);
        const int z = bzi * 4 + (lane >> 4);
        double delta = 0.0;
        if (x < g.lx && y < g.ly && z < g.lz) {
            const int64_t i = lin(x, y, z, g);
            const double pc = p[i] - shift;
            p[i] = pc;
            delta = pc - p_prev[i];
            e1 += fabs(delta);
            e2 += fabs(pc);
            p_prev[i] = pc;
        }
        const double bs = wave_sum(delta);
        if (lane == 0) blocks[b] = bs;
    }
    const double s1 = block_sum(e1, lds);
    const double s2 = block_sum(e2, lds);
    if (threadIdx.x == 0) {
        parts[2 * blockIdx.x] = s1;
        parts[2 * blockIdx.x + 1] = s2;
    }
}

// sum_b |B_b| (blockAbsSumKernel, GPU/InamuroCUDA.cu:1250-1275).
__global__ __launch_bounds__(BLOCK) void k_acc_block_abs(const double* __restrict__ blocks, int nblocks,
                                                         DevState* st)
{
    __shared__ double lds[4];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += BLOCK) acc += fabs(blocks[i]);
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) st->acc_block_abs = s;
}

// coarsePressureCorrectionKernel (GPU/InamuroCUDA.cu:1277-1323), applied when
// the check just made has not converged and sum |p - p_prev| > 0
// (:2149-2158); h is nullptr in scalar mode, where h is not iterated.
__global__ __launch_bounds__(BLOCK) void k_acc_two_grid(Grid g, int nbx, int nby, double strength,
                                                        const double* __restrict__ blocks,
                                                        const DevState* __restrict__ st,
                                                        double* __restrict__ p,
                                                        double* __restrict__ p_prev,
                                                        double* __restrict__ h)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    if (st->converged || !(st->e1 > 0.0)) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    const int bxi = x / 4, byi = y / 4, bzi = z / 4;
    const int b = (bzi * nby + byi) * nbx + bxi;
    const int x0 = bxi * 4, y0 = byi * 4, z0 = bzi * 4;
    const int x1 = min(x0 + 4, g.lx), y1 = min(y0 + 4, g.ly), z1 = min(z0 + 4, g.lz);
    const int block_cells = max(1, (x1 - x0) * (y1 - y0) * (z1 - z0));
    const double correction = strength * blocks[b] / static_cast<double>(block_cells);
    const double corrected_p = p[c] + correction;
    p[c] = corrected_p;
    p_prev[c] = corrected_p;
    if (h) {
#pragma unroll
        for (int q = 0; q < Q; ++q) h[q * g.n + c] += EW[q] * correction;
    }
}

} // namespace

namespace acc {

// Layers a block of the staged kernels marches: up to STENCIL_ZCHUNK, fewer
// on small grids so that the launch still has about four blocks per CU
// (1024); the staging of the two extra layers is then spread over fewer
// computed ones.
int pull_zchunk(const Grid& g)
{
    const int64_t tiles = static_cast<int64_t>((g.lx + STENCIL_TX - 1) / STENCIL_TX) *
                          ((g.ly + STENCIL_TY - 1) / STENCIL_TY);
    const int64_t zc = tiles * g.lz / 1024;
    return static_cast<int>(zc < 1 ? 1 : (zc > STENCIL_ZCHUNK ? STENCIL_ZCHUNK : zc));
}

int pull_blocks(const Grid& g)
{
    const int zc = pull_zchunk(g);
    return ((g.lx + STENCIL_TX - 1) / STENCIL_TX) * ((g.ly + STENCIL_TY - 1) / STENCIL_TY) *
           ((g.lz + zc - 1) / zc);
}

void pull(hipStream_t s, const Grid& g, int onepass, int anderson, double scale, double omega,
          const double* h, const double* p, const double* tau, const double* src, double* hn,
          double* pn, double* pr, double* pi, double* parts)
{
    const int nb = pull_blocks(g), zc = pull_zchunk(g);
    if (!onepass)
        k_acc_pull<0, 0><<<nb, BLOCK, 0, s>>>(g, zc, scale, omega, h, p, tau, src, hn, pn, pr, pi, parts);
    else if (anderson == 0)
        k_acc_pull<1, 0><<<nb, BLOCK, 0, s>>>(g, zc, scale, omega, h, p, tau, src, hn, pn, pr, pi, parts);
    else if (anderson == 1)
        k_acc_pull<1, 1><<<nb, BLOCK, 0, s>>>(g, zc, scale, omega, h, p, tau, src, hn, pn, pr, pi, parts);
    else
        k_acc_pull<1, 2><<<nb, BLOCK, 0, s>>>(g, zc, scale, omega, h, p, tau, src, hn, pn, pr, pi, parts);
}

void anderson_mix(hipStream_t s, const Grid& g, double beta_max, const double* parts, DevState* st,
                  const double* h, const double* img, double* pi_to_a, double* pr, double* pn)
{
    k_acc_alpha<<<1, BLOCK, 0, s>>>(parts, pull_blocks(g), beta_max, st);
    k_acc_apply<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, st, h, img, pi_to_a, pr, pn);
}

void scalar(hipStream_t s, const Grid& g, double scale, const double* p, const double* rho,
            const double* src, double* pn)
{
    k_acc_scalar<<<pull_blocks(g), BLOCK, 0, s>>>(g, pull_zchunk(g), scale, p, rho, src, pn);
}

void source_init(hipStream_t s, const Grid& g, double scale, const double* p, const double* tau,
                 const double* src, double* h)
{
    k_acc_source_init<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, scale, p, tau, src, h);
}

int block_count(const Grid& g, int* nbx, int* nby)
{
    *nbx = (g.lx + 3) / 4;
    *nby = (g.ly + 3) / 4;
    return *nbx * *nby * ((g.lz + 3) / 4);
}

void check(hipStream_t s, const Grid& g, const Phys& ph, double tol, int blocks_needed,
           int spatial, int two_grid, double strength, double* p, double* p_prev, double* h,
           double* blocks, double* partials, DevState* st)
{
    const int nb = red_blocks(g.n);
    double* parts_sum = partials;
    double* parts_err = partials + RED_BLOCKS;
    k_acc_sum<<<nb, BLOCK, 0, s>>>(g, p, parts_sum);
    f64::gauge_finalize(s, ph, parts_sum, nb, st);
    if (h) k_acc_gauge_h<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, st, h);
    int nbx = 0, nby = 0;
    const int nblk = block_count(g, &nbx, &nby);
    int nparts = nb;
    if (blocks_needed) {
        const int wg = (nblk + 3) / 4;
        nparts = wg < RED_BLOCKS ? wg : RED_BLOCKS;
        k_acc_gauge_error_blocks<<<nparts, BLOCK, 0, s>>>(g, nbx, nby, nblk, p, p_prev, st, blocks,
                                                           parts_err);
    } else {
        f64::gauge_error(s, g, p, p_prev, st, parts_err);
    }
    f64::check_finalize(s, parts_err, nullptr, nparts, 0, tol, st);
    if (spatial) k_acc_block_abs<<<1, BLOCK, 0, s>>>(blocks, nblk, st);
    if (two_grid)
        k_acc_two_grid<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, nbx, nby, strength, blocks, st, p, p_prev, h);
}

void check_streamed(hipStream_t s, const Grid& g, const Phys& ph, double tol, int spatial, int two_grid,
                    double strength, const double* hin_const, double* h_stored, double* p, double* p_prev,
                    double* blocks, double* partials, DevState* st)
{
    const int nb = red_blocks(g.n);
    double* parts_sum = partials;
    double* parts_err = partials + RED_BLOCKS;
    f64::getp_partials(s, g, hin_const, p, parts_sum);
    f64::gauge_finalize(s, ph, parts_sum, nb, st);
    int nbx = 0, nby = 0;
    const int nblk = block_count(g, &nbx, &nby);
    const int wg = (nblk + 3) / 4;
    const int nparts = wg < RED_BLOCKS ? wg : RED_BLOCKS;
    k_acc_gauge_error_blocks<<<nparts, BLOCK, 0, s>>>(g, nbx, nby, nblk, p, p_prev, st, blocks, parts_err);
    f64::check_finalize(s, parts_err, nullptr, nparts, 0, tol, st);
    if (spatial) k_acc_block_abs<<<1, BLOCK, 0, s>>>(blocks, nblk, st);
    if (two_grid)
        k_acc_two_grid<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, nbx, nby, strength, blocks, st, p, p_prev,
                                                           h_stored);
}

} // namespace acc
} // namespace lbm
