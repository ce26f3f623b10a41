// kernels_adjoint.hip -- discrete adjoint of the FP64 time step for gfx950.
//
// The reference has no adjoint.  These kernels transpose, primitive by
// primitive, the forward kernels of kernels_f64.hip as they execute with a
// frozen Poisson iteration count per step:
//
//   forward                                 adjoint here
//   correct_uvw (CPU/Inamuro.cpp:794-823)   k_adj_correct    gather of the
//                                           mirrored / periodic differences
//   Poisson iteration, push form            k_adj_poisson    push form too:
//   (collide, stream, x-slip, getp,         own-cell loads, slip transposed
//   lazy gauge; CPU/Inamuro.cpp:673-793)    in place, shifted stores; lazy gauge
//   gauge p -= s, h -= E s                  k_adj_lam_finalize
//   getMacro (CPU/Inamuro.cpp:294-346),     k_adj_source + the gather inside
//   pull + slip of f, g                     k_adj_collide
//   collision (CPU/Inamuro.cpp:134-250)     k_adj_collide
//   firstord / secondord                    k_adj_stencil   gather over the
//   (CPU/Inamuro.hpp:213-314)               cells whose stencil reads x
//
// Every kernel is a gather or a permutation (each output written by exactly
// one thread), so no atomics are needed and every result is deterministic.
// Reductions use the same fixed-shape wave64 trees as the forward path.  Compiled with
// -ffp-contract=off like the forward FP64 kernels.
//
// Data layout: as the forward path, macro fields [z][y][x] (x fastest), FP64,
// distributions and their adjoints SoA [q][z][y][x].
#include "kernels.hpp"

namespace lbm {
namespace adj {

namespace {

// Mirror partner of each x-moving direction (+x <-> -x with y, z kept),
// from the slip pairs; -1 for directions with c_x = 0.
constexpr int PARTNER[Q] = {-1, 4, -1, -1, 1, -1, -1, 8, 7, 14, 13, 12, 11, 10, 9};

// Transpose of "pull from x - c_q (across a z wall: the cell's own opposite
// population), then x-slip overwrite" (the view every reader of f, g and h
// has).  view(q, d) is the adjoint of the viewed value of direction q at cell
// d; out[r] becomes the adjoint of the unstreamed value r at this cell, which
// lands at d = x + c_r as direction r, or, when that crosses a z wall, at this
// cell as direction opp(r).  Where it lands as direction rr at d:
//   d on the x = 0 wall:      +x values there are overwritten  -> 0
//                             -x value rr also feeds its partner -> + view(partner)
//   d on the x = lx-1 wall:   the same with the signs exchanged.
template <typename View>
__device__ __forceinline__ double slip_T(int rr, int dx, int64_t d, const Grid& g, View view)
{
    if (CX[rr] == 0) return view(rr, d);
    const bool low = dx == 0, high = dx == g.lx - 1;
    if ((low && CX[rr] == 1) || (high && CX[rr] == -1)) return 0.0;
    if (low || high) return view(rr, d) + view(PARTNER[rr], d);
    return view(rr, d);
}

template <typename View>
__device__ __forceinline__ void gather_stream_slip_T(int x, int y, int z, const Grid& g, View view,
                                                     double (&out)[Q])
{
    const int xm = wrap(x, -1, g.lx), xp = wrap(x, 1, g.lx);
    const int ym = wrap(y, -1, g.ly), yp = wrap(y, 1, g.ly);
    const int zm = wrap(z, -1, g.lz), zp = wrap(z, 1, g.lz);
#pragma unroll
    for (int r = 0; r < Q; ++r) {
        if (z_bounced_out(r, z, g)) {
            out[r] = slip_T(OPP[r], x, lin(x, y, z, g), g, view);
        } else {
            const int dx = CX[r] == 1 ? xp : (CX[r] == -1 ? xm : x);
            const int dy = CY[r] == 1 ? yp : (CY[r] == -1 ? ym : y);
            const int dz = CZ[r] == 1 ? zp : (CZ[r] == -1 ? zm : z);
            out[r] = slip_T(r, dx, lin(dx, dy, dz, g), g, view);
        }
    }
}

// How many viewed values read the unstreamed value r at cell x (0, 1 or 2):
// the x-slip count of where it lands (see slip_T).
__device__ __forceinline__ double view_count(int r, int x, int z, const Grid& g)
{
    const bool b = z_bounced_out(r, z, g);
    const int rr = b ? OPP[r] : r;
    const int dx = b ? x : wrap(x, CX[r], g.lx);
    if (CX[rr] == 1) return dx == 0 ? 0.0 : (dx == g.lx - 1 ? 2.0 : 1.0);
    if (CX[rr] == -1) return dx == g.lx - 1 ? 0.0 : (dx == 0 ? 2.0 : 1.0);
    return 1.0;
}

// The 15 streamed values of cell c seen through the x-slip rule (the forward
// Poisson kernels' load_streamed).  A recorded iterate is read once, so the
// load carries a non-temporal hint and leaves the cache to the adjoint state.
__device__ __forceinline__ void load_view(const double* __restrict__ hin, int64_t c, int x,
                                          const Grid& g, double (&hv)[Q])
{
#pragma unroll
    for (int q = 0; q < Q; ++q) hv[q] = __builtin_nontemporal_load(&hin[q * g.n + c]);
    if (x == 0) {
#pragma unroll
        for (int r = 0; r < 5; ++r) hv[SLIP_POS[r]] = hv[SLIP_NEG[r]];
    }
    if (x == g.lx - 1) {
#pragma unroll
        for (int r = 0; r < 5; ++r) hv[SLIP_NEG[r]] = hv[SLIP_POS[r]];
    }
}

__device__ __forceinline__ double reduce_partials(const double* __restrict__ parts, int count,
                                                  double* lds)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < count; i += BLOCK) acc += parts[i];
    return block_sum(acc, lds);
}

// ---- objective --------------------------------------------------------------

// J = sum (fei - target)^2 (block partials) and the seed dJ/dfei = 2 (fei - target).
__global__ __launch_bounds__(BLOCK) void k_adj_seed(Grid g, const double* __restrict__ fei,
                                                    const double* __restrict__ target,
                                                    double* __restrict__ lfei,
                                                    double* __restrict__ parts)
{
    __shared__ double lds[4];
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    double j = 0.0;
    if (c < g.n) {
        const double d = fei[c] - target[c];
        j = d * d;
        lfei[c] = 2.0 * d;
    }
    const double s = block_sum(j, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

__global__ __launch_bounds__(BLOCK) void k_adj_set_objective(const double* __restrict__ parts,
                                                             int nparts, AdjState* as)
{
    __shared__ double lds[4];
    const double s = reduce_partials(parts, nparts, lds);
    if (threadIdx.x == 0) {
        as->J = s;
        as->kg = 0.0;
    }
}

// Centroid height J = sum z w / sum w, w = fei - fei_G: block partials of
// sum w (parts[0..nb)) and sum z w (parts[nb..2nb)).
__global__ __launch_bounds__(BLOCK) void k_adj_centroid_parts(Grid g, double fei_G,
                                                              const double* __restrict__ fei,
                                                              double* __restrict__ parts)
{
    __shared__ double lds[4];
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    double w = 0.0, zw = 0.0;
    if (c < g.n) {
        int x, y, z;
        cell_xyz(c, g, x, y, z);
        w = fei[c] - fei_G;
        zw = static_cast<double>(z) * w;
    }
    const double sw = block_sum(w, lds);
    const double szw = block_sum(zw, lds);
    if (threadIdx.x == 0) {
        parts[blockIdx.x] = sw;
        parts[gridDim.x + blockIdx.x] = szw;
    }
}

__global__ __launch_bounds__(BLOCK) void k_adj_centroid_final(const double* __restrict__ parts,
                                                              int nparts, AdjState* as)
{
    __shared__ double lds[4];
    const double W = reduce_partials(parts, nparts, lds);
    const double Z = reduce_partials(parts + nparts, nparts, lds);
    if (threadIdx.x == 0) {
        as->J = Z / W;
        as->weight = W;
        as->kg = 0.0;
    }
}

// dJ/dfei = (z - J) / W.
__global__ __launch_bounds__(BLOCK) void k_adj_centroid_seed(Grid g, const AdjState* __restrict__ as,
                                                             double* __restrict__ lfei)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    lfei[c] = (static_cast<double>(z) - as->J) / as->weight;
}

// ---- correct_uvw --------------------------------------------------------------

// u -= (p_e - p_w) / (2 rho') and likewise v, w (CPU/Inamuro.cpp:794-823).
// Writes lpf = adjoint of the solve's final p (incoming lp plus the gathered
// difference terms) and adds the rho' terms to lrho.  The x rows 0 and lx-1
// use the mirrored pair (e = w), whose two terms cancel; so do periodic axes
// of length <= 2.  Block partials of lpf feed the adjoint of the last gauge.
__global__ __launch_bounds__(BLOCK) void k_adj_correct(
    Grid g, const double* __restrict__ p, const double* __restrict__ rho,
    const double* __restrict__ lu, const double* __restrict__ lv, const double* __restrict__ lw,
    const double* __restrict__ lp, double* __restrict__ lrho, double* __restrict__ lpf,
    double* __restrict__ parts)
{
    __shared__ double lds[4];
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    double out = 0.0;
    if (c < g.n) {
        int x, y, z;
        cell_xyz(c, g, x, y, z);
        double acc = lp[c];
        // x: cell x-1 subtracts via its east neighbour x, cell x+1 adds via
        // its west neighbour x; only interior rows 1..lx-2 contribute.
        if (x - 1 >= 1 && x - 1 <= g.lx - 2) {
            const int64_t i = lin(x - 1, y, z, g);
            acc -= lu[i] * (1.0 / (2.0 * rho[i]));
        }
        if (x + 1 >= 1 && x + 1 <= g.lx - 2) {
            const int64_t i = lin(x + 1, y, z, g);
            acc += lu[i] * (1.0 / (2.0 * rho[i]));
        }
        if (g.ly > 2) {
            const int64_t is = lin(x, wrap(y, -1, g.ly), z, g), in = lin(x, wrap(y, 1, g.ly), z, g);
            acc -= lv[is] * (1.0 / (2.0 * rho[is]));
            acc += lv[in] * (1.0 / (2.0 * rho[in]));
        }
        if (g.zwall) {
            // a wall layer's outer neighbour is the layer itself
            if (z >= 1) {
                const int64_t ib = lin(x, y, z - 1, g);
                acc -= lw[ib] * (1.0 / (2.0 * rho[ib]));
            }
            if (z == g.lz - 1) acc -= lw[c] * (1.0 / (2.0 * rho[c]));
            if (z <= g.lz - 2) {
                const int64_t it = lin(x, y, z + 1, g);
                acc += lw[it] * (1.0 / (2.0 * rho[it]));
            }
            if (z == 0) acc += lw[c] * (1.0 / (2.0 * rho[c]));
        } else if (g.lz > 2) {
            const int64_t ib = lin(x, y, wrap(z, -1, g.lz), g), it = lin(x, y, wrap(z, 1, g.lz), g);
            acc -= lw[ib] * (1.0 / (2.0 * rho[ib]));
            acc += lw[it] * (1.0 / (2.0 * rho[it]));
        }
        lpf[c] = acc;
        out = acc;
        // d u / d rho' = (p_e - p_w) / (2 rho'^2)
        const int x_e = (x == g.lx - 1) ? g.lx - 2 : x + 1;
        const int x_w = (x == 0) ? 1 : x - 1;
        const int y_n = (y == g.ly - 1) ? 0 : y + 1;
        const int y_s = (y == 0) ? g.ly - 1 : y - 1;
        const int z_n = z_stencil(z, 1, g);
        const int z_s = z_stencil(z, -1, g);
        const double r = rho[c];
        lrho[c] += (lu[c] * (p[lin(x_e, y, z, g)] - p[lin(x_w, y, z, g)]) +
                    lv[c] * (p[lin(x, y_n, z, g)] - p[lin(x, y_s, z, g)]) +
                    lw[c] * (p[lin(x, y, z_n, g)] - p[lin(x, y, z_s, g)])) /
                   (2.0 * r * r);
    }
    const double s = block_sum(out, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// Adjoint of a gauge (p -= s, h -= E s, s = mean(p) - target): with
// Lambda = sum over cells of the adjoints of p and of sum_q E_q h_q after the
// gauge, every pre-gauge adjoint loses Lambda/N (applied lazily by the next
// k_adj_poisson) and the target, mean(p_start), gains Lambda.
__global__ __launch_bounds__(BLOCK) void k_adj_lam_finalize(const double* __restrict__ parts,
                                                            int nparts, double cell_count,
                                                            int reset, AdjState* as)
{
    __shared__ double lds[4];
    const double L = reduce_partials(parts, nparts, lds);
    if (threadIdx.x == 0) {
        as->lam_shift = L / cell_count;
        as->lam_target = reset ? L : as->lam_target + L;
    }
}

// ---- Poisson iteration ------------------------------------------------------

enum { SRC_PLAIN = 0, SRC_LAZY = 1, SRC_INIT = 2 };

// Adjoint of forward iteration k, in push form (the transpose of the forward
// push: loads at the cell itself, stores shifted).  Bin holds, at each cell x,
// lh*_q(x): the adjoint of the value iteration k's collision produced there,
// already streamed back from where it landed (x + c_q) and slip-transposed.
// With SRC_LAZY the pending gauge adjoint of check k is applied on read: each
// such value carries -Lambda/N once per viewed value it collects (0 for a
// value the slip overwrote, 2 where it also fed its mirror partner).  For the
// last iteration (SRC_INIT) every viewed value at y carries lpf(y), gathered.
// MODE is how forward iteration k obtained its input (h_{k-1}, p_{k-1}):
//   H_EQUILIBRIUM: h = E p_start (first iteration); writes lp_start = dJ/dp_start
//   H_GAUGED:      h = view(raw) - E shift, p = gauged p of the check; block
//                  partials of sum_q E_q L_q for that check's adjoint
//   H_STREAMED:    h = view(raw), p = sum_q h_q
// Forward collision: h* = h - (h - E p)/tau - (E/3) S, so
//   lv = lh* (1 - 1/tau),  lp = sum lh* E / tau,  lS -= sum lh* E / 3,
//   ltau += sum lh* (h - E p) / tau^2,
// and L = lv + lp is the adjoint of iteration k-1's viewed output; its slip
// transpose at x is pushed to x - c_q for the next (earlier) iteration.
// TB is the storage type of the carried adjoint Bin/Bout: double, or float
// for the FP32 adjoint state (arithmetic, lS, ltau and the gauge partials stay
// FP64 either way).
template <int SRC, int MODE, typename TB, int PB>
__global__ __launch_bounds__(PB) void k_adj_poisson(
    Grid g, const TB* __restrict__ Bin, const double* __restrict__ lpf,
    const AdjState* __restrict__ as, const double* __restrict__ raw,
    const double* __restrict__ pin, const double* __restrict__ shift_ptr,
    const double* __restrict__ tau, double* __restrict__ lS, double* __restrict__ ltau,
    TB* __restrict__ Bout, double* __restrict__ lp_start, double* __restrict__ parts)
{
    __shared__ double lds[4];
    // the gauged iteration's block sum assumes BLOCK threads
    static_assert(MODE != H_GAUGED || PB == BLOCK, "gauged adjoint iterations run BLOCK-thread blocks");
    const int64_t c = static_cast<int64_t>(blockIdx.x) * PB + threadIdx.x;
    double part = 0.0;
    if (c < g.n) {
        int x, y, z;
        cell_xyz(c, g, x, y, z);
        double lh[Q];
        if (SRC == SRC_INIT) {
            const double ls = as->lam_shift;
            gather_stream_slip_T(x, y, z, g, [&](int, int64_t d) { return lpf[d] - ls; }, lh);
        } else {
#pragma unroll
            for (int q = 0; q < Q; ++q) lh[q] = static_cast<double>(__builtin_nontemporal_load(&Bin[q * g.n + c]));
            if (SRC == SRC_LAZY) {
                const double ls = as->lam_shift;
#pragma unroll
                for (int q = 0; q < Q; ++q) lh[q] -= view_count(q, x, z, g) * ls;
            }
        }
        double hv[Q];
        double p;
        if (MODE == H_EQUILIBRIUM) {
            p = pin[c];
#pragma unroll
            for (int q = 0; q < Q; ++q) hv[q] = p * EW[q];
        } else if (MODE == H_GAUGED) {
            load_view(raw, c, x, g, hv);
            const double shift = *shift_ptr;
#pragma unroll
            for (int q = 0; q < Q; ++q) hv[q] -= EW[q] * shift;
            p = pin[c];
        } else {
            load_view(raw, c, x, g, hv);
            p = 0.0;
#pragma unroll
            for (int q = 0; q < Q; ++q) p += hv[q];
        }
        const double t = tau[c];
        double np = 0.0, ns = 0.0, nt = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            np += lh[q] * EW[q] / t;
            ns -= lh[q] * EW[q] / 3.0;
            nt += lh[q] * (hv[q] - EW[q] * p) / (t * t);
            lh[q] = lh[q] * (1.0 - 1.0 / t);
        }
        lS[c] += ns;
        ltau[c] += nt;
        if (MODE == H_EQUILIBRIUM) {
            double b = np;
#pragma unroll
            for (int q = 0; q < Q; ++q) b += EW[q] * lh[q];
            lp_start[c] = b;
        } else {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                lh[q] += np;
                part += EW[q] * lh[q];
            }
            // slip transpose at x: the overwritten value gets nothing, its
            // source collects both
            if (x == 0) {
#pragma unroll
                for (int r = 0; r < 5; ++r) {
                    lh[SLIP_NEG[r]] += lh[SLIP_POS[r]];
                    lh[SLIP_POS[r]] = 0.0;
                }
            }
            if (x == g.lx - 1) {
#pragma unroll
                for (int r = 0; r < 5; ++r) {
                    lh[SLIP_POS[r]] += lh[SLIP_NEG[r]];
                    lh[SLIP_NEG[r]] = 0.0;
                }
            }
            const int xm = wrap(x, -1, g.lx), xp = wrap(x, 1, g.lx);
            const int ym = wrap(y, -1, g.ly), yp = wrap(y, 1, g.ly);
            const int zm = wrap(z, -1, g.lz), zp = wrap(z, 1, g.lz);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                // the streamed value q at x came from x - c_q, or, across a
                // z wall, from this cell's opposite population
                if (z_bounced_in(q, z, g)) {
                    Bout[OPP[q] * g.n + c] = static_cast<TB>(lh[q]);
                } else {
                    const int sx = CX[q] == 1 ? xm : (CX[q] == -1 ? xp : x);
                    const int sy = CY[q] == 1 ? ym : (CY[q] == -1 ? yp : y);
                    const int sz = CZ[q] == 1 ? zm : (CZ[q] == -1 ? zp : z);
                    Bout[q * g.n + lin(sx, sy, sz, g)] = static_cast<TB>(lh[q]);
                }
            }
        }
    }
    if (MODE == H_GAUGED) {
        const double s = block_sum(part, lds);
        if (threadIdx.x == 0) parts[blockIdx.x] = s;
    }
}

// ---- convergence of the reverse solve -------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_adj_change(Grid g, const double* __restrict__ lS,
                                                      double* __restrict__ lS_prev,
                                                      double* __restrict__ parts)
{
    __shared__ double lds[4];
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    double d = 0.0, a = 0.0;
    if (c < g.n) {
        const double v = lS[c];
        d = fabs(v - lS_prev[c]);
        a = fabs(v);
        lS_prev[c] = v;
    }
    const double sd = block_sum(d, lds);
    const double sa = block_sum(a, lds);
    if (threadIdx.x == 0) {
        parts[blockIdx.x] = sd;
        parts[gridDim.x + blockIdx.x] = sa;
    }
}

__global__ __launch_bounds__(BLOCK) void k_adj_change_finalize(const double* __restrict__ parts, int nparts,
                                                               double done, double tol, int first,
                                                               AdjState* as)
{
    __shared__ double lds[4];
    const double sd = reduce_partials(parts, nparts, lds);
    const double sa = reduce_partials(parts + nparts, nparts, lds);
    if (threadIdx.x == 0) {
        const double rel = sa > 0.0 ? sd / sa : 0.0;
        if (first) as->conv_reached = -1.0;
        as->conv_change = rel;
        if (rel < tol && as->conv_reached < 0.0) as->conv_reached = done;
    }
}

__global__ void k_adj_change_record(const AdjState* __restrict__ as, double* __restrict__ out, int step)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        out[2 * step] = as->conv_change;
        out[2 * step + 1] = as->conv_reached;
    }
}

// ---- entry of the solve and getMacro ------------------------------------------

// After all iterations: lps += lambda_target / N (gauge target = mean p_start),
// lrho' -= ltau / rho'^2 (tau_h = 1/rho' + 0.5), and the per-cell adjoint of
// every streamed f population from the macro sum and the density map,
//   lmac = lfei' + lrho' * d rho / d fei (fei')   (CPU/Inamuro.cpp:296-327),
// plus block partials of lS for the source projection S = div - mean(div).
__global__ __launch_bounds__(BLOCK) void k_adj_source(
    Grid g, Phys ph, const double* __restrict__ fei, const double* __restrict__ rho,
    const double* __restrict__ ltau, const double* __restrict__ lS,
    const double* __restrict__ lfei, const AdjState* __restrict__ as, double* __restrict__ lrho,
    double* __restrict__ lps, double* __restrict__ lmac, double* __restrict__ parts)
{
    __shared__ double lds[4];
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    double s = 0.0;
    if (c < g.n) {
        lps[c] += as->lam_target / ph.cell_count;
        const double r = rho[c];
        const double lr = lrho[c] - ltau[c] / (r * r);
        lrho[c] = lr;
        const double f = fei[c];
        double dr = 0.0;
        if (f > ph.fei_G && f < ph.fei_L) {
            const double k = M_PI / (ph.fei_L - ph.fei_G);
            dr = (ph.rho_L - ph.rho_G) / 2.0 * cos((f - ph.fei_avg) * k) * k;
        }
        lmac[c] = lfei[c] + lr * dr;
        s = lS[c];
    }
    const double t = block_sum(s, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = t;
}

__global__ __launch_bounds__(BLOCK) void k_adj_mean(const double* __restrict__ parts, int nparts,
                                                    double cell_count, AdjState* as)
{
    __shared__ double lds[4];
    const double s = reduce_partials(parts, nparts, lds);
    if (threadIdx.x == 0) as->lS_mean = s / cell_count;
}

// ---- collision -----------------------------------------------------------

// Slots of the per-cell adjoint array written by k_adj_collide and gathered by
// k_adj_stencil: centre values, then the gradient and Laplacian adjoints.
enum {
    PC_FEI, PC_RHO, PC_U, PC_V, PC_W,
    PC_UX, PC_UY, PC_UZ, PC_VX, PC_VY, PC_VZ, PC_WX, PC_WY, PC_WZ,
    PC_RX, PC_RY, PC_RZ, PC_FX, PC_FY, PC_FZ,
    PC_ULAP, PC_VLAP, PC_WLAP, PC_FLAP
};

// Adjoint of the f/g collision at one cell (CPU/Inamuro.cpp:134-250), with
// the adjoints of its outputs gathered from the cells they stream to:
//   f: cf (the next step's collision input adjoint) + lmac
//   g: cg + c_q . (lu', lv', lw')   (u' = sum_q c_q g_q)
// Writes the collision-input adjoints cf_o, cg_o (zero on the first step,
// where f = feq), the 24 per-cell adjoints, and block partials of dJ/dk_g.
template <int FIRST>
__global__ __launch_bounds__(BLOCK) void k_adj_collide(
    Grid g, Phys ph,
    const double* __restrict__ fei, const double* __restrict__ rho,
    const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
    const double* __restrict__ cf, const double* __restrict__ cg,
    const double* __restrict__ lmac, const double* __restrict__ lu,
    const double* __restrict__ lv, const double* __restrict__ lw,
    const double* __restrict__ lS, const AdjState* __restrict__ as,
    double* __restrict__ cf_o, double* __restrict__ cg_o, double* __restrict__ pc,
    double* __restrict__ parts)
{
    __shared__ double lds[4];
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    double kg = 0.0;
    if (c < g.n) {
        int x, y, z;
        cell_xyz(c, g, x, y, z);
        double af[Q], ag[Q];
        gather_stream_slip_T(x, y, z, g, [&](int q, int64_t d) { return cf[q * g.n + d] + lmac[d]; }, af);
        gather_stream_slip_T(x, y, z, g, [&](int q, int64_t d) {
            return cg[q * g.n + d] + (UC[q] * lu[d] + VC[q] * lv[d] + WC[q] * lw[d]);
        }, ag);

        // Step-start fields and their stencils, as the forward collision.
        int64_t idx[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k)
            idx[k] = lin(mirror_x(x, CX[k], g.lx), wrap(y, CY[k], g.ly), z_stencil(z, CZ[k], g), g);
        double Lu[3], gp[3], gr[3], Lphi;
        {
            double au = 0.0, av = 0.0, aw = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
            double rx = 0.0, ry = 0.0, rz = 0.0, lf = 0.0;
#pragma unroll
            for (int k = 1; k < Q; ++k) {
                au += u[idx[k]];
                av += v[idx[k]];
                aw += w[idx[k]];
                const double fk = fei[idx[k]];
                ax += fk * UC[k];
                ay += fk * VC[k];
                az += fk * WC[k];
                lf += fk;
                const double rk = rho[idx[k]];
                rx += rk * UC[k];
                ry += rk * VC[k];
                rz += rk * WC[k];
            }
            Lu[0] = (au - 14.0 * u[c]) / 5.0;
            Lu[1] = (av - 14.0 * v[c]) / 5.0;
            Lu[2] = (aw - 14.0 * w[c]) / 5.0;
            Lphi = (lf - 14.0 * fei[c]) / 5.0;
            gp[0] = ax / 10.0; gp[1] = ay / 10.0; gp[2] = az / 10.0;
            gr[0] = rx / 10.0; gr[1] = ry / 10.0; gr[2] = rz / 10.0;
        }
        const double phi = fei[c], rc = rho[c];
        const double uc[3] = {u[c], v[c], w[c]};
        const double sum_rho = gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2];
        const double mu = (rc - ph.rho_G) * ph.tmp1 + ph.mu_G;
        const double dp0 = ph.T / ((1.0 - ph.b * phi) * (1.0 - ph.b * phi)) - 2.0 * ph.a * phi;
        const double V = 1.5 * (ph.taug - 0.5) * 2.0;
        double a_phi = 0.0, a_rho = 0.0, a_u[3] = {0.0, 0.0, 0.0};
        double a_gp[3] = {0.0, 0.0, 0.0}, a_gr[3] = {0.0, 0.0, 0.0}, a_Lu[3] = {0.0, 0.0, 0.0};
        double a_D[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        double a_p0 = 0.0, a_sf = 0.0, a_sr = 0.0, a_vel = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double cq[3] = {UC[q], VC[q], WC[q]};
            double fq = af[q], gq = ag[q];
            // gravity: g_q += 3 E_q c_qz gz (1 - rho_G/rho)
            if (CZ[q] != 0 && ph.gz != 0.0) a_rho += gq * 3.0 * EW[q] * WC[q] * ph.gz * ph.rho_G / (rc * rc);
            if (FIRST) {
                cf_o[q * g.n + c] = 0.0;
                cg_o[q * g.n + c] = 0.0;
            } else {
                // f_out = f_in - (f_in - feq)/tau_f;
                // g_out = g_in - (g_in - geq)/tau_g + 3 E/rho mu (c . lap u)
                cf_o[q * g.n + c] = fq * (1.0 - 1.0 / ph.tauf);
                cg_o[q * g.n + c] = gq * (1.0 - 1.0 / ph.taug);
                const double cl = cq[0] * Lu[0] + cq[1] * Lu[1] + cq[2] * Lu[2];
#pragma unroll
                for (int a = 0; a < 3; ++a) a_Lu[a] += gq * 3.0 * EW[q] * mu / rc * cq[a];
                a_rho += gq * 3.0 * EW[q] * cl * (ph.tmp1 / rc - mu / (rc * rc));
                fq /= ph.tauf;
                gq /= ph.taug;
            }
            const double un = cq[0] * uc[0] + cq[1] * uc[1] + cq[2] * uc[2];
            const double fe = cq[0] * gp[0] + cq[1] * gp[1] + cq[2] * gp[2];
            const double re = cq[0] * gr[0] + cq[1] * gr[1] + cq[2] * gr[2];
            const double cs = cq[0] * cq[0] + cq[1] * cq[1] + cq[2] * cq[2];
            const double Gr = 4.5 * re * re - 1.5 * sum_rho * cs;
            // feq = H fei + F (p0 - k_f fei lap - k_f |grad fei|^2 / 6)
            //       + 3 E fei (c.u) + E k_f (4.5 (c.grad fei)^2 - 1.5 |grad fei|^2 |c|^2)
            a_phi += fq * (HW[q] + 3.0 * EW[q] * un);
            a_p0 += fq * FW[q];
            a_sf += fq * EW[q] * ph.k_f * (-1.5 * cs);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                a_gp[a] += fq * EW[q] * ph.k_f * 9.0 * fe * cq[a];
                a_u[a] += fq * 3.0 * EW[q] * phi * cq[a];
            }
            // geq = E velPart + E k_g/rho Grho - 2/3 F k_g/rho |grad rho|^2
            const double vb = gq * EW[q];
            a_vel += vb;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                a_u[a] += vb * (3.0 + 9.0 * un) * cq[a];
#pragma unroll
                for (int b = 0; b < 3; ++b) a_D[a * 3 + b] += vb * V * cq[a] * cq[b];
            }
            const double gb = vb * ph.k_g / rc;
#pragma unroll
            for (int a = 0; a < 3; ++a) a_gr[a] += gb * 9.0 * re * cq[a];
            a_sr += gb * (-1.5 * cs) - gq * 2.0 / 3.0 * FW[q] * ph.k_g / rc;
            a_rho += -vb * ph.k_g * Gr / (rc * rc) + gq * 2.0 / 3.0 * FW[q] * ph.k_g * sum_rho / (rc * rc);
            kg += vb * Gr / rc - gq * 2.0 / 3.0 * FW[q] * sum_rho / rc;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) a_u[a] -= 3.0 * uc[a] * a_vel;
        a_phi += a_p0 * (dp0 - ph.k_f * Lphi);
        a_sf += a_p0 * (-ph.k_f / 6.0);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            a_gp[a] += a_sf * 2.0 * gp[a];
            a_gr[a] += a_sr * 2.0 * gr[a];
        }
        // The projected Poisson source reads u_x, v_y, w_z.
        const double lsp = lS[c] - as->lS_mean;
        const int64_t n = g.n;
        pc[PC_FEI * n + c] = a_phi;
        pc[PC_RHO * n + c] = a_rho;
        pc[PC_U * n + c] = a_u[0];
        pc[PC_V * n + c] = a_u[1];
        pc[PC_W * n + c] = a_u[2];
        pc[PC_UX * n + c] = lsp + a_D[0];
        pc[PC_UY * n + c] = a_D[1];
        pc[PC_UZ * n + c] = a_D[2];
        pc[PC_VX * n + c] = a_D[3];
        pc[PC_VY * n + c] = lsp + a_D[4];
        pc[PC_VZ * n + c] = a_D[5];
        pc[PC_WX * n + c] = a_D[6];
        pc[PC_WY * n + c] = a_D[7];
        pc[PC_WZ * n + c] = lsp + a_D[8];
        pc[PC_RX * n + c] = a_gr[0];
        pc[PC_RY * n + c] = a_gr[1];
        pc[PC_RZ * n + c] = a_gr[2];
        pc[PC_FX * n + c] = a_gp[0];
        pc[PC_FY * n + c] = a_gp[1];
        pc[PC_FZ * n + c] = a_gp[2];
        pc[PC_ULAP * n + c] = a_Lu[0];
        pc[PC_VLAP * n + c] = a_Lu[1];
        pc[PC_WLAP * n + c] = a_Lu[2];
        pc[PC_FLAP * n + c] = a_p0 * (-ph.k_f * phi);
    }
    const double s = block_sum(kg, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

__global__ __launch_bounds__(BLOCK) void k_adj_kg(const double* __restrict__ parts, int nparts,
                                                  AdjState* as)
{
    __shared__ double lds[4];
    const double s = reduce_partials(parts, nparts, lds);
    if (threadIdx.x == 0) as->kg += s;
}

// ---- stencils -------------------------------------------------------------

// Transpose of firstord / secondord (CPU/Inamuro.hpp:213-314): the cells x'
// whose stencil direction k reads cell x are x - c_k (y, z periodic) when that
// lies inside [0, lx-1], plus the mirrored reader (x' = lx-1 reads lx-2 for
// c_x = +1, x' = 0 reads 1 for c_x = -1); with the z walls, the same in z
// about the wall layers (x' with z' = lz-1 reads itself for c_z = +1, z' = 0
// for c_z = -1).  Each output is the centre adjoint
// plus sum over readers of (c_k . grad adjoint)/10 + lap adjoint/5, minus
// 14/5 of its own lap adjoint.
__global__ __launch_bounds__(BLOCK) void k_adj_stencil(Grid g, const double* __restrict__ pc,
                                                       double* __restrict__ lfei,
                                                       double* __restrict__ lrho,
                                                       double* __restrict__ lu,
                                                       double* __restrict__ lv,
                                                       double* __restrict__ lw)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    const int64_t n = g.n;
    double o[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    auto add = [&](int k, int64_t s) {
        const double ck[3] = {UC[k], VC[k], WC[k]};
        o[0] += (ck[0] * pc[PC_FX * n + s] + ck[1] * pc[PC_FY * n + s] + ck[2] * pc[PC_FZ * n + s]) / 10.0 +
                pc[PC_FLAP * n + s] / 5.0;
        o[1] += (ck[0] * pc[PC_RX * n + s] + ck[1] * pc[PC_RY * n + s] + ck[2] * pc[PC_RZ * n + s]) / 10.0;
        o[2] += (ck[0] * pc[PC_UX * n + s] + ck[1] * pc[PC_UY * n + s] + ck[2] * pc[PC_UZ * n + s]) / 10.0 +
                pc[PC_ULAP * n + s] / 5.0;
        o[3] += (ck[0] * pc[PC_VX * n + s] + ck[1] * pc[PC_VY * n + s] + ck[2] * pc[PC_VZ * n + s]) / 10.0 +
                pc[PC_VLAP * n + s] / 5.0;
        o[4] += (ck[0] * pc[PC_WX * n + s] + ck[1] * pc[PC_WY * n + s] + ck[2] * pc[PC_WZ * n + s]) / 10.0 +
                pc[PC_WLAP * n + s] / 5.0;
    };
    auto add_x = [&](int k, int sy, int sz) {
        const int sx = x - CX[k];
        if (sx >= 0 && sx <= g.lx - 1) add(k, lin(sx, sy, sz, g));
        if (CX[k] == 1 && x == g.lx - 2) add(k, lin(g.lx - 1, sy, sz, g));
        if (CX[k] == -1 && x == 1) add(k, lin(0, sy, sz, g));
    };
#pragma unroll
    for (int k = 1; k < Q; ++k) {
        const int sy = wrap(y, -CY[k], g.ly);
        if (!g.zwall) {
            add_x(k, sy, wrap(z, -CZ[k], g.lz));
        } else {
            // z readers with the walls: z - c_z inside the box, and a wall
            // layer, whose stencil point beyond the wall is itself
            const int sz = z - CZ[k];
            if (sz >= 0 && sz <= g.lz - 1) add_x(k, sy, sz);
            if ((CZ[k] == 1 && z == g.lz - 1) || (CZ[k] == -1 && z == 0)) add_x(k, sy, z);
        }
    }
    lfei[c] = pc[PC_FEI * n + c] + o[0] - 14.0 * pc[PC_FLAP * n + c] / 5.0;
    lrho[c] = pc[PC_RHO * n + c] + o[1];
    lu[c] = pc[PC_U * n + c] + o[2] - 14.0 * pc[PC_ULAP * n + c] / 5.0;
    lv[c] = pc[PC_V * n + c] + o[3] - 14.0 * pc[PC_VLAP * n + c] / 5.0;
    lw[c] = pc[PC_W * n + c] + o[4] - 14.0 * pc[PC_WLAP * n + c] / 5.0;
}

} // namespace

void seed(hipStream_t s, const Grid& g, const Phys& ph, int objective, const double* fei,
          const double* target, double* lfei, double* parts, AdjState* as)
{
    const int nb = grid_blocks(g.n);
    if (objective == ADJ_OBJ_CENTROID_Z) {
        k_adj_centroid_parts<<<nb, BLOCK, 0, s>>>(g, ph.fei_G, fei, parts);
        k_adj_centroid_final<<<1, BLOCK, 0, s>>>(parts, nb, as);
        k_adj_centroid_seed<<<nb, BLOCK, 0, s>>>(g, as, lfei);
        return;
    }
    k_adj_seed<<<nb, BLOCK, 0, s>>>(g, fei, target, lfei, parts);
    k_adj_set_objective<<<1, BLOCK, 0, s>>>(parts, nb, as);
}

void correct(hipStream_t s, const Grid& g, const Phys& ph, const double* p, const double* rho,
             const double* lu, const double* lv, const double* lw, const double* lp, double* lrho,
             double* lpf, double* parts, AdjState* as)
{
    const int nb = grid_blocks(g.n);
    k_adj_correct<<<nb, BLOCK, 0, s>>>(g, p, rho, lu, lv, lw, lp, lrho, lpf, parts);
    k_adj_lam_finalize<<<1, BLOCK, 0, s>>>(parts, nb, ph.cell_count, 1, as);
}

template <typename TB>
static void poisson_t(hipStream_t s, const Grid& g, int src, int mode, const TB* Bin,
                      const double* lpf, AdjState* as, const double* raw, const double* pin,
                      const double* shift, const double* tau, double* lS, double* ltau, TB* Bout,
                      double* lp_start, double* parts, int block)
{
    // The gauged iterations reduce block sums over BLOCK threads (their count
    // and order fixed by the grid); the others may use wider blocks.
#define LBM_ADJ_AT(S, M, PB)                                                                          \
    k_adj_poisson<S, M, TB, PB><<<static_cast<int>((g.n + PB - 1) / PB), PB, 0, s>>>(                 \
        g, Bin, lpf, as, raw, pin, shift, tau, lS, ltau, Bout, lp_start, parts)
#define LBM_ADJ_LAUNCH(S, M)                                         \
    do {                                                             \
        if constexpr (M != H_GAUGED) {                               \
            if (block == 768) {                                      \
                LBM_ADJ_AT(S, M, 768);                               \
                break;                                               \
            }                                                        \
        }                                                            \
        LBM_ADJ_AT(S, M, BLOCK);                                     \
    } while (0)
#define LBM_ADJ_MODES(S)                                         \
    do {                                                         \
        if (mode == H_EQUILIBRIUM) LBM_ADJ_LAUNCH(S, H_EQUILIBRIUM); \
        else if (mode == H_GAUGED) LBM_ADJ_LAUNCH(S, H_GAUGED);      \
        else LBM_ADJ_LAUNCH(S, H_STREAMED);                          \
    } while (0)
    if (src == ADJ_SRC_INIT) LBM_ADJ_MODES(SRC_INIT);
    else if (src == ADJ_SRC_LAZY) LBM_ADJ_MODES(SRC_LAZY);
    else LBM_ADJ_MODES(SRC_PLAIN);
#undef LBM_ADJ_MODES
#undef LBM_ADJ_LAUNCH
#undef LBM_ADJ_AT
}

void poisson(hipStream_t s, const Grid& g, const Phys& ph, int src, int mode, bool state_f32,
             const void* Bin, const double* lpf, AdjState* as, const double* raw, const double* pin,
             const double* shift, const double* tau, double* lS, double* ltau, void* Bout,
             double* lp_start, double* parts, int block)
{
    if (state_f32)
        poisson_t(s, g, src, mode, static_cast<const float*>(Bin), lpf, as, raw, pin, shift, tau, lS,
                  ltau, static_cast<float*>(Bout), lp_start, parts, block);
    else
        poisson_t(s, g, src, mode, static_cast<const double*>(Bin), lpf, as, raw, pin, shift, tau, lS,
                  ltau, static_cast<double*>(Bout), lp_start, parts, block);
    if (mode == H_GAUGED) k_adj_lam_finalize<<<1, BLOCK, 0, s>>>(parts, grid_blocks(g.n), ph.cell_count, 0, as);
}

void source(hipStream_t s, const Grid& g, const Phys& ph, const double* fei, const double* rho,
            const double* ltau, const double* lS, const double* lfei, AdjState* as, double* lrho,
            double* lps, double* lmac, double* parts)
{
    const int nb = grid_blocks(g.n);
    k_adj_source<<<nb, BLOCK, 0, s>>>(g, ph, fei, rho, ltau, lS, lfei, as, lrho, lps, lmac, parts);
    k_adj_mean<<<1, BLOCK, 0, s>>>(parts, nb, ph.cell_count, as);
}

void collide(hipStream_t s, int first, const Grid& g, const Phys& ph, const double* fei,
             const double* rho, const double* u, const double* v, const double* w, const double* cf,
             const double* cg, const double* lmac, const double* lu, const double* lv,
             const double* lw, const double* lS, AdjState* as, double* cf_o, double* cg_o,
             double* pc, double* parts)
{
    const int nb = grid_blocks(g.n);
    if (first)
        k_adj_collide<1><<<nb, BLOCK, 0, s>>>(g, ph, fei, rho, u, v, w, cf, cg, lmac, lu, lv, lw, lS,
                                              as, cf_o, cg_o, pc, parts);
    else
        k_adj_collide<0><<<nb, BLOCK, 0, s>>>(g, ph, fei, rho, u, v, w, cf, cg, lmac, lu, lv, lw, lS,
                                              as, cf_o, cg_o, pc, parts);
    k_adj_kg<<<1, BLOCK, 0, s>>>(parts, nb, as);
}

void converge(hipStream_t s, const Grid& g, const double* lS, double* lS_prev, uint64_t done, double tol,
              int first, double* parts, AdjState* as)
{
    const int nb = grid_blocks(g.n);
    k_adj_change<<<nb, BLOCK, 0, s>>>(g, lS, lS_prev, parts);
    k_adj_change_finalize<<<1, BLOCK, 0, s>>>(parts, nb, static_cast<double>(done), tol, first, as);
}

void converge_record(hipStream_t s, const AdjState* as, double* out, int step)
{
    k_adj_change_record<<<1, 64, 0, s>>>(as, out, step);
}

void stencil(hipStream_t s, const Grid& g, const double* pc, double* lfei, double* lrho, double* lu,
             double* lv, double* lw)
{
    k_adj_stencil<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, pc, lfei, lrho, lu, lv, lw);
}

} // namespace adj
} // namespace lbm
