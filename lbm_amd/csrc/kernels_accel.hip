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

// One iteration of the pull form.  ONEPASS: fixed-point relaxation against
// the cell's own h.  AND (onepass only): 0 none; 1 the first Anderson
// iteration, which also stores the history (residual image - h, image);
// 2 later iterations, which add the block partials of <R_prev, R - R_prev>
// and |R - R_prev|^2 and leave p to the mixing kernel.
template <int ONEPASS, int AND>
__global__ __launch_bounds__(BLOCK) void k_acc_pull(
    Grid g, double scale, double omega, const double* __restrict__ h, const double* __restrict__ p,
    const double* __restrict__ tau, const double* __restrict__ src, double* __restrict__ hn,
    double* __restrict__ pn, double* __restrict__ pr, double* __restrict__ pi,
    double* __restrict__ parts)
{
    __shared__ double lds[4];
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    double num = 0.0, den = 0.0;
    if (c < g.n) {
        int x, y, z;
        cell_xyz(c, g, x, y, z);
        const int xm = wrap(x, -1, g.lx), xp = wrap(x, 1, g.lx);
        const int ym = wrap(y, -1, g.ly), yp = wrap(y, 1, g.ly);
        const int zm = wrap(z, -1, g.lz), zp = wrap(z, 1, g.lz);
        double im[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int sx = CX[q] == 1 ? xm : (CX[q] == -1 ? xp : x);
            const int sy = CY[q] == 1 ? ym : (CY[q] == -1 ? yp : y);
            const int sz = CZ[q] == 1 ? zm : (CZ[q] == -1 ? zp : z);
            const int64_t s = lin(sx, sy, sz, g);
            const double tauh = tau[s];
            const double hequ = EW[q] * p[s];
            const double h_old = h[q * g.n + s];
            im[q] = h_old - scale * (h_old - hequ) / tauh - (EW[q] / 3.0) * src[s];
        }
        if (x == 0) {
#pragma unroll
            for (int r = 0; r < 5; ++r) im[SLIP_POS[r]] = im[SLIP_NEG[r]];
        } else if (x == g.lx - 1) {
#pragma unroll
            for (int r = 0; r < 5; ++r) im[SLIP_NEG[r]] = im[SLIP_POS[r]];
        }
        double p_sum = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int64_t k = q * g.n + c;
            double v = im[q];
            double h_own = 0.0;
            if (ONEPASS) {
                h_own = h[k];
                v = h_own + omega * (v - h_own);
            }
            hn[k] = v;
            p_sum += v;
            if (AND == 1) {
                pr[k] = v - h_own;
                pi[k] = v;
            } else if (AND == 2) {
                const double residual = v - h_own;
                const double prev = pr[k];
                const double delta = residual - prev;
                num += prev * delta;
                den += delta * delta;
            }
        }
        if (AND != 2) pn[c] = p_sum;
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
// neighbours x - c_q in q order.  pressure_bounce_direction replaces a
// direction whose source lies beyond an x wall (c_x = +1 at x = 0, c_x = -1
// at x = lx-1) by its slip partner, which differs only in c_x, so the
// neighbour is the mirror image x + c_x.
__global__ __launch_bounds__(BLOCK) void k_acc_scalar(Grid g, double scale, const double* __restrict__ p,
                                                      const double* __restrict__ rho,
                                                      const double* __restrict__ src,
                                                      double* __restrict__ pn)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    double neighbor_sum = 0.0;
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        const bool mirrored = (x == 0 && CX[q] == 1) || (x == g.lx - 1 && CX[q] == -1);
        const int sx = wrap(x, mirrored ? CX[q] : -CX[q], g.lx);
        const int sy = wrap(y, -CY[q], g.ly);
        const int sz = wrap(z, -CZ[q], g.lz);
        neighbor_sum += p[lin(sx, sy, sz, g)];
    }
    const double rhs = scale * rho[c] * src[c];
    pn[c] = (neighbor_sum - 5.0 * rhs) / 14.0;
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

void pull(hipStream_t s, const Grid& g, int onepass, int anderson, double scale, double omega,
          const double* h, const double* p, const double* tau, const double* src, double* hn,
          double* pn, double* pr, double* pi, double* parts)
{
    const int nb = grid_blocks(g.n);
    if (!onepass)
        k_acc_pull<0, 0><<<nb, BLOCK, 0, s>>>(g, scale, omega, h, p, tau, src, hn, pn, pr, pi, parts);
    else if (anderson == 0)
        k_acc_pull<1, 0><<<nb, BLOCK, 0, s>>>(g, scale, omega, h, p, tau, src, hn, pn, pr, pi, parts);
    else if (anderson == 1)
        k_acc_pull<1, 1><<<nb, BLOCK, 0, s>>>(g, scale, omega, h, p, tau, src, hn, pn, pr, pi, parts);
    else
        k_acc_pull<1, 2><<<nb, BLOCK, 0, s>>>(g, scale, omega, h, p, tau, src, hn, pn, pr, pi, parts);
}

void anderson_mix(hipStream_t s, const Grid& g, double beta_max, const double* parts, DevState* st,
                  const double* h, const double* img, double* pi_to_a, double* pr, double* pn)
{
    k_acc_alpha<<<1, BLOCK, 0, s>>>(parts, grid_blocks(g.n), beta_max, st);
    k_acc_apply<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, st, h, img, pi_to_a, pr, pn);
}

void scalar(hipStream_t s, const Grid& g, double scale, const double* p, const double* rho,
            const double* src, double* pn)
{
    k_acc_scalar<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, scale, p, rho, src, pn);
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
