// kernels_f64.hip -- plain FP64 time step for gfx950.
//
// Compiled with -ffp-contract=off: every expression keeps the association of
// the reference CPU solver (CPU/Inamuro.cpp, CPU/Inamuro.hpp), divisions stay
// divisions and reciprocal-multiplies stay reciprocal-multiplies, so each
// per-cell value is the one the reference computes.  Only global sums differ
// in association (fixed-shape wave64 trees instead of one sequential loop).
//
// Data layout in HBM: macro fields [z][y][x] (x fastest) in FP64,
// distributions SoA [q][z][y][x].  f and g are kept post-collision and
// unstreamed; the macro and collision kernels pull f_q(x - c_q) and apply the
// x-slip rule on the fly, so streaming never costs a pass of its own.
#include "kernels.hpp"

namespace lbm {
namespace f64 {

namespace {

// Pulls the 15 populations that stream into cell (x,y,z) from the
// post-collision array `in` (periodic on all axes, CPU/Inamuro.cpp:252-287)
// and applies the x-slip overwrite (CPU/Inamuro.cpp:589-623).
__device__ __forceinline__ void pull_slip(const double* __restrict__ in, int x, int y, int z,
                                          const Grid& g, double (&o)[Q])
{
    const int xm = wrap(x, -1, g.lx), xp = wrap(x, 1, g.lx);
    const int ym = wrap(y, -1, g.ly), yp = wrap(y, 1, g.ly);
    const int zm = wrap(z, -1, g.lz), zp = wrap(z, 1, g.lz);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int sx = CX[q] == 1 ? xm : (CX[q] == -1 ? xp : x);
        const int sy = CY[q] == 1 ? ym : (CY[q] == -1 ? yp : y);
        const int sz = CZ[q] == 1 ? zm : (CZ[q] == -1 ? zp : z);
        if (z_bounced_in(q, z, g))
            o[q] = in[OPP[q] * g.n + lin(x, y, z, g)];  // bounce-back at a z wall
        else
            o[q] = in[q * g.n + lin(sx, sy, sz, g)];
    }
    if (x == 0) {
#pragma unroll
        for (int r = 0; r < 5; ++r) o[SLIP_POS[r]] = o[SLIP_NEG[r]];
    }
    if (x == g.lx - 1) {
#pragma unroll
        for (int r = 0; r < 5; ++r) o[SLIP_NEG[r]] = o[SLIP_POS[r]];
    }
}

// firstord (CPU/Inamuro.hpp:213-266) and secondord (CPU/Inamuro.hpp:268-314)
// of one field at one cell.  nb[k] holds the field at the D3Q15 neighbour
// k (k = 1..14, x mirrored, y/z periodic); nb[0] is the cell itself.
__device__ __forceinline__ void grad(const double (&nb)[Q], double& gx, double& gy, double& gz)
{
    double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
    for (int k = 1; k < Q; ++k) {
        ax += nb[k] * UC[k];
        ay += nb[k] * VC[k];
        az += nb[k] * WC[k];
    }
    gx = ax / 10.0;
    gy = ay / 10.0;
    gz = az / 10.0;
}

__device__ __forceinline__ double lap(const double (&nb)[Q])
{
    double acc = 0.0;
#pragma unroll
    for (int k = 1; k < Q; ++k) acc += nb[k];
    return (acc - 14.0 * nb[0]) / 5.0;
}

__device__ __forceinline__ void load_nb(const double* a, const int (&idx)[Q], double (&nb)[Q])
{
#pragma unroll
    for (int k = 0; k < Q; ++k) nb[k] = a[idx[k]];
}

// Inamuro::collision, CPU/Inamuro.cpp:134-250, fused with the pull of the
// streamed f/g (mode FG_PULL) so that one pass reads the stencils, the
// incoming populations and writes the post-collision populations.  A block
// owns a STENCIL_TX x STENCIL_TY tile and marches it through STENCIL_ZCHUNK
// z layers, staging u, v, w, rho and phi of each layer and its x/y halo in
// LDS once (a ring of three layers), so the 5 x 15 stencil reads of a cell
// come from LDS, in the reference's k order, bit for bit.
template <int MODE>
__global__ __launch_bounds__(BLOCK) void k_collide(
    Grid g, Phys ph,
    const double* __restrict__ fei, const double* __restrict__ rho,
    const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
    const double* __restrict__ p,
    const double* __restrict__ fin, const double* __restrict__ gin,
    double* __restrict__ fout, double* __restrict__ gout,
    double* __restrict__ ux_o, double* __restrict__ vy_o, double* __restrict__ wz_o,
    double* __restrict__ partials)
{
    __shared__ double lds[4];
    __shared__ double staged[5][3][STENCIL_PLANE];  // u, v, w, rho, phi; ring of 3 layers
    const int ntx = (g.lx + STENCIL_TX - 1) / STENCIL_TX, nty = (g.ly + STENCIL_TY - 1) / STENCIL_TY;
    const int b = static_cast<int>(blockIdx.x);
    const int x0 = (b % ntx) * STENCIL_TX, y0 = ((b / ntx) % nty) * STENCIL_TY;
    const int zb = (b / (ntx * nty)) * STENCIL_ZCHUNK, ze = min(g.lz, zb + STENCIL_ZCHUNK);
    // A plane is staged in two halves: its loads go to registers (next[])
    // before the current layer is collided, and land in LDS after it, so the
    // loads' latency overlaps the collision arithmetic.
    constexpr int PER = (STENCIL_PLANE + BLOCK - 1) / BLOCK;
    double next[5][PER];
    auto fetch = [&](int gz) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = threadIdx.x + j * BLOCK;
            if (e < STENCIL_PLANE) {
                const int64_t i = stencil_source(e, x0, y0, gz, g);
                next[0][j] = u[i];
                next[1][j] = v[i];
                next[2][j] = w[i];
                next[3][j] = rho[i];
                next[4][j] = fei[i];
            }
        }
    };
    auto land = [&](int slot) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = threadIdx.x + j * BLOCK;
            if (e < STENCIL_PLANE) {
#pragma unroll
                for (int f = 0; f < 5; ++f) staged[f][slot][e] = next[f][j];
            }
        }
    };
    // layer L sits in ring slot (L - zb + 1) mod 3
    fetch(z_stencil(zb, -1, g));
    land(0);
    fetch(zb);
    land(1);
    fetch(z_stencil(zb, 1, g));
    const int tx = threadIdx.x % STENCIL_TX, ty = threadIdx.x / STENCIL_TX;
    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < g.lx && y < g.ly;
    double div_c = 0.0, p_c = 0.0, adiv_c = 0.0;
    for (int z = zb; z < ze; ++z) {
        const int sm = (z - zb) % 3, s0 = (z - zb + 1) % 3, sp = (z - zb + 2) % 3;
        land(sp);
        __syncthreads();
        if (z + 1 < ze) fetch(z_stencil(z + 1, 1, g));
        if (valid) {
            const int64_t c = lin(x, y, z, g);
            // flat index into staged[f]: ring slot of the neighbour's layer, then
            // its place in the plane
            int idx[Q];
#pragma unroll
            for (int k = 0; k < Q; ++k)
                idx[k] = (CZ[k] == 1 ? sp : (CZ[k] == -1 ? sm : s0)) * STENCIL_PLANE + stencil_slot(tx, ty, k);

            double nb[Q];
            double u_x, u_y, u_z, v_x, v_y, v_z, w_x, w_y, w_z;
            double rho_x, rho_y, rho_z, fei_x, fei_y, fei_z;
            load_nb(&staged[0][0][0], idx, nb);
            grad(nb, u_x, u_y, u_z);
            const double u_lap = lap(nb);
            const double uc = nb[0];
            load_nb(&staged[1][0][0], idx, nb);
            grad(nb, v_x, v_y, v_z);
            const double v_lap = lap(nb);
            const double vc = nb[0];
            load_nb(&staged[2][0][0], idx, nb);
            grad(nb, w_x, w_y, w_z);
            const double w_lap = lap(nb);
            const double wc = nb[0];
            load_nb(&staged[3][0][0], idx, nb);
            grad(nb, rho_x, rho_y, rho_z);
            const double rhoc = nb[0];
            load_nb(&staged[4][0][0], idx, nb);
            grad(nb, fei_x, fei_y, fei_z);
            const double fei_lap = lap(nb);
            const double feic = nb[0];

            const double mu = (rhoc - ph.rho_G) * ph.tmp1 + ph.mu_G;
            const double sum_fei = fei_x * fei_x + fei_y * fei_y + fei_z * fei_z;
            const double sum_rho = rho_x * rho_x + rho_y * rho_y + rho_z * rho_z;
            const double fei2sum = fei_lap;
            const double usq = uc * uc + vc * vc + wc * wc;
            const double p0 = feic * ph.T / (1.0 - ph.b * feic) - ph.a * feic * feic;

            double fv[Q], gv[Q];
            if (MODE == FG_PULL) {
                pull_slip(fin, x, y, z, g, fv);
                pull_slip(gin, x, y, z, g, gv);
            } else if (MODE == FG_LOCAL) {
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    fv[q] = fin[q * g.n + c];
                    gv[q] = gin[q * g.n + c];
                }
            }

#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const double un = UC[i] * uc + VC[i] * vc + WC[i] * wc;
                const double fei_ei = UC[i] * fei_x + VC[i] * fei_y + WC[i] * fei_z;
                const double rho_ei = UC[i] * rho_x + VC[i] * rho_y + WC[i] * rho_z;
                const double ci_sq = UC[i] * UC[i] + VC[i] * VC[i] + WC[i] * WC[i];
                const double Gfei = 4.5 * (fei_ei * fei_ei) - 1.5 * sum_fei * ci_sq;
                const double Grho = 4.5 * (rho_ei * rho_ei) - 1.5 * sum_rho * ci_sq;
                const double fequ = HW[i] * feic +
                                    FW[i] * (p0 - ph.k_f * feic * fei2sum - ph.k_f * sum_fei / 6.0) +
                                    3.0 * EW[i] * feic * un +
                                    EW[i] * ph.k_f * Gfei;
                const double velPart = 1.0 + 3.0 * un - 1.5 * usq + 4.5 * un * un +
                    1.5 * (ph.taug - 0.5) * 2.0 *
                        (UC[i] * UC[i] * u_x + UC[i] * VC[i] * u_y + UC[i] * WC[i] * u_z +
                         VC[i] * UC[i] * v_x + VC[i] * VC[i] * v_y + VC[i] * WC[i] * v_z +
                         WC[i] * UC[i] * w_x + WC[i] * VC[i] * w_y + WC[i] * WC[i] * w_z);
                const double gequ = EW[i] * velPart +
                                    EW[i] * ph.k_g / rhoc * Grho -
                                    2.0 / 3.0 * FW[i] * ph.k_g / rhoc * sum_rho;
                double fo, go;
                if (MODE == FG_FIRST) {
                    fo = fequ;
                    go = gequ;
                } else {
                    fo = fv[i] - (fv[i] - fequ) / ph.tauf;
                    go = gv[i] - (gv[i] - gequ) / ph.taug +
                         3.0 * EW[i] / rhoc * mu * (UC[i] * u_lap + VC[i] * v_lap + WC[i] * w_lap);
                }
                if (CZ[i] != 0 && ph.gz != 0.0) {
                    // extension: buoyant acceleration gz (1 - rho_G/rho) on the
                    // predicted velocity
                    const double az = ph.gz * (1.0 - ph.rho_G / rhoc);
                    go += 3.0 * EW[i] * WC[i] * az;
                }
                fout[i * g.n + c] = fo;
                gout[i * g.n + c] = go;
            }
            ux_o[c] = u_x;
            vy_o[c] = v_y;
            wz_o[c] = w_z;
            const double div = u_x + v_y + w_z;
            div_c += div;
            adiv_c += fabs(div);
            p_c += p[c];
        }
        __syncthreads();  // the next layer's staging reuses the oldest slot
    }
    const double sd = block_sum(div_c, lds);
    const double sp = block_sum(p_c, lds);
    const double sa = block_sum(adiv_c, lds);
    if (threadIdx.x == 0) {
        partials[SOURCE_TERMS * blockIdx.x] = sd;
        partials[SOURCE_TERMS * blockIdx.x + 1] = sp;
        partials[SOURCE_TERMS * blockIdx.x + 2] = sa;
    }
}

// Deterministic fixed-order sum of `count` strided partials in one block.
__device__ __forceinline__ double reduce_partials(const double* __restrict__ parts, int count,
                                                  int stride, int offset, double* lds)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < count; i += BLOCK) acc += parts[i * stride + offset];
    return block_sum(acc, lds);
}

// Source statistics of solvePressurePoisson, CPU/Inamuro.cpp:652-664, with
// the GPU solver's sum |div| (GPU/InamuroCUDA.cu:1868-1893).
__global__ __launch_bounds__(BLOCK) void k_source_finalize(Phys ph, const double* __restrict__ parts,
                                                           int nparts, DevState* st)
{
    __shared__ double lds[4];
    const double sd = reduce_partials(parts, nparts, SOURCE_TERMS, 0, lds);
    const double sp = reduce_partials(parts, nparts, SOURCE_TERMS, 1, lds);
    const double sa = reduce_partials(parts, nparts, SOURCE_TERMS, 2, lds);
    if (threadIdx.x == 0) {
        const double dm = sd / ph.cell_count;
        const double gt = sp / ph.cell_count;
        st->div_mean = dm;
        st->div_abs_sum = sa;
        st->gauge_target = gt;
        st->shift = 0.0;
        st->max_shift = 0.0;
        st->converged = 0;
        st->pressure_converged = 0;
        st->fp_converged = 0;
        st->fp_evaluated = 0;
        if (!isfinite(dm) || !isfinite(gt)) {
            if (st->status == ST_OK) st->status = ST_SOURCE_NONFINITE;
        }
    }
}

// getMacro (CPU/Inamuro.cpp:294-346) on the pulled populations, plus the two
// per-cell Poisson inputs that stay fixed for the whole solve:
//   tau_h = 1.0/rho + 0.5                           (CPU/Inamuro.cpp:764)
//   S     = ((u_x - mean) + v_y) + w_z              (CPU/Inamuro.cpp:665-668, 773)
template <int PULL>
__device__ __forceinline__ double macro_cell(
    int64_t c, const Grid& g, const Phys& ph, const double* __restrict__ fin,
    const double* __restrict__ gin, const double* __restrict__ ux, const double* __restrict__ vy,
    const double* __restrict__ wz, const DevState* __restrict__ st,
    double* __restrict__ fei, double* __restrict__ rho, double* __restrict__ u,
    double* __restrict__ v, double* __restrict__ w, double* __restrict__ tau,
    double* __restrict__ src)
{
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    double fv[Q], gv[Q];
    if (PULL) {
        pull_slip(fin, x, y, z, g, fv);
        pull_slip(gin, x, y, z, g, gv);
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            fv[q] = fin[q * g.n + c];
            gv[q] = gin[q * g.n + c];
        }
    }
    double fei_local = 0.0;
#pragma unroll
    for (int i = 0; i < Q; ++i) fei_local += fv[i];
    double density;
    if (fei_local <= ph.fei_G) {
        density = ph.rho_G;
    } else if (fei_local >= ph.fei_L) {
        density = ph.rho_L;
    } else {
        const double arg = (fei_local - ph.fei_avg) / (ph.fei_L - ph.fei_G) * M_PI;
        density = (ph.rho_L - ph.rho_G) / 2.0 * (sin(arg) + 1.0) + ph.rho_G;
    }
    double uu = 0.0, vv = 0.0, ww = 0.0;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
        uu += UC[i] * gv[i];
        vv += VC[i] * gv[i];
        ww += WC[i] * gv[i];
    }
    fei[c] = fei_local;
    rho[c] = density;
    u[c] = uu;
    v[c] = vv;
    w[c] = ww;
    tau[c] = 1.0 / density + 0.5;
    const double sc = ((ux[c] - st->div_mean) + vy[c]) + wz[c];
    src[c] = sc;
    return sc;
}

template <int PULL>
__global__ __launch_bounds__(BLOCK) void k_macro(
    Grid g, Phys ph, const double* __restrict__ fin, const double* __restrict__ gin,
    const double* __restrict__ ux, const double* __restrict__ vy, const double* __restrict__ wz,
    const DevState* __restrict__ st,
    double* __restrict__ fei, double* __restrict__ rho, double* __restrict__ u,
    double* __restrict__ v, double* __restrict__ w, double* __restrict__ tau,
    double* __restrict__ src, double* __restrict__ parts)
{
    __shared__ double lds[4];
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    double s_c = 0.0;
    if (c < g.n) s_c = macro_cell<PULL>(c, g, ph, fin, gin, ux, vy, wz, st, fei, rho, u, v, w, tau, src);
    // Block sums of the projected source (GPU/InamuroCUDA.cu:1897-1920).
    const double bs = block_sum(s_c, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = bs;
}

// Loads the 15 streamed populations of cell c and applies the x-slip rule
// (CPU/Inamuro.cpp:589-623): the pressure populations are stored streamed
// but not yet bounced, so every reader sees them through this view.
template <int NT = 0>
__device__ __forceinline__ void load_streamed(const double* __restrict__ hin, int64_t c, int x,
                                              const Grid& g, double (&hv)[Q])
{
#pragma unroll
    for (int q = 0; q < Q; ++q) hv[q] = NT ? __builtin_nontemporal_load(&hin[q * g.n + c]) : hin[q * g.n + c];
    if (x == 0) {
#pragma unroll
        for (int r = 0; r < 5; ++r) hv[SLIP_POS[r]] = hv[SLIP_NEG[r]];
    }
    if (x == g.lx - 1) {
#pragma unroll
        for (int r = 0; r < 5; ++r) hv[SLIP_NEG[r]] = hv[SLIP_POS[r]];
    }
}

// One Poisson LBM iteration in push form.  Per cell: getp of the previous
// iteration (p = sum_q h_q in q order, CPU/Inamuro.cpp:779-793), collision_p
// with the cell's own tau_h and source (CPU/Inamuro.cpp:755-778), and the
// periodic stream (CPU/Inamuro.cpp:252-287) as 15 stores to the neighbours.
// Every value is the one the reference computes; only the place where the
// slip rule and the pressure sum are evaluated moves, so all loads are the
// cell's own (aligned, coalesced) and p is written only at checks.
// NT: the streamed h is loaded with a non-temporal hint (it is dead after
// this read).  PB: threads per block.  Which (PB, NT) pair is fastest depends
// on the grid's shape; the caller picks one of PUSH_CONFIGS.
// RELAX: the reference GPU solver's pressure relaxation scale s
// (collisionPressureKernel, GPU/InamuroCUDA.cu:503-537), h - s (h - E p) /
// tau_h - (E/3) S; the strict map is RELAX = 0.
template <int MODE, int NT, int PB, int RELAX>
__global__ __launch_bounds__(PB) void k_poisson_push(
    Grid g, const double* __restrict__ hin, double* __restrict__ hout,
    const double* __restrict__ pbuf, const double* __restrict__ tau,
    const double* __restrict__ src, const DevState* __restrict__ st, double relax)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * PB + threadIdx.x;
    if (c >= g.n) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    double hv[Q];
    double p;
    if (MODE == H_EQUILIBRIUM) {
        p = pbuf[c];
#pragma unroll
        for (int q = 0; q < Q; ++q) hv[q] = p * EW[q];
    } else if (MODE == H_EXPLICIT) {
#pragma unroll
        for (int q = 0; q < Q; ++q) hv[q] = hin[q * g.n + c];
        p = pbuf[c];
    } else {
        load_streamed<NT>(hin, c, x, g, hv);
        if (MODE == H_GAUGED) {
            const double shift = st->shift;
#pragma unroll
            for (int q = 0; q < Q; ++q) hv[q] -= EW[q] * shift;
            p = pbuf[c];
        } else {
            p = 0.0;
#pragma unroll
            for (int q = 0; q < Q; ++q) p += hv[q];
        }
    }
    const double tauh = tau[c];
    const double sc = src[c];
    const int xm = wrap(x, -1, g.lx), xp = wrap(x, 1, g.lx);
    const int ym = wrap(y, -1, g.ly), yp = wrap(y, 1, g.ly);
    const int zm = wrap(z, -1, g.lz), zp = wrap(z, 1, g.lz);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double hequ = EW[q] * p;
        const double o = RELAX ? hv[q] - relax * (hv[q] - hequ) / tauh - (EW[q] / 3.0) * sc
                               : hv[q] - (hv[q] - hequ) / tauh - (EW[q] / 3.0) * sc;
        const int dx = CX[q] == 1 ? xp : (CX[q] == -1 ? xm : x);
        const int dy = CY[q] == 1 ? yp : (CY[q] == -1 ? ym : y);
        const int dz = CZ[q] == 1 ? zp : (CZ[q] == -1 ? zm : z);
        if (z_bounced_out(q, z, g))
            hout[OPP[q] * g.n + c] = o;  // bounce-back at a z wall
        else
            hout[q * g.n + lin(dx, dy, dz, g)] = o;
    }
}

// getp after the last iteration of a segment (CPU/Inamuro.cpp:779-793), with
// block partial sums of p for the gauge (CPU/Inamuro.cpp:691-695).
__global__ __launch_bounds__(BLOCK) void k_getp_sum(Grid g, const double* __restrict__ hin,
                                                    double* __restrict__ p,
                                                    double* __restrict__ parts)
{
    __shared__ double lds[4];
    double acc = 0.0;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x; c < g.n;
         c += static_cast<int64_t>(gridDim.x) * BLOCK) {
        int x, y, z;
        cell_xyz(c, g, x, y, z);
        double hv[Q];
        load_streamed(hin, c, x, g, hv);
        double pc = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) pc += hv[q];
        p[c] = pc;
        acc += pc;
    }
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// ---- convergence check (CPU/Inamuro.cpp:689-715) -------------------------

__global__ __launch_bounds__(BLOCK) void k_sum_strided(Grid g, const double* __restrict__ a,
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

// gauge_shift = mean(p) - target (CPU/Inamuro.cpp:691-698).
__global__ __launch_bounds__(BLOCK) void k_gauge_finalize(Phys ph, const double* __restrict__ parts,
                                                          int nparts, DevState* st)
{
    __shared__ double lds[4];
    const double cur = reduce_partials(parts, nparts, 1, 0, lds);
    if (threadIdx.x == 0) {
        const double shift = cur / ph.cell_count - st->gauge_target;
        st->shift = shift;
        if (!isfinite(shift)) {
            if (st->status == ST_OK) st->status = ST_GAUGE_NONFINITE;
        } else if (fabs(shift) > st->max_shift) {
            st->max_shift = fabs(shift);
        }
    }
}

// p -= shift, then getError (CPU/Inamuro.cpp:699-706, 726-753): partial sums of
// |p - p_prev| and |p|, and p_prev <- p.
__global__ __launch_bounds__(BLOCK) void k_gauge_error(Grid g, double* __restrict__ p,
                                                       double* __restrict__ p_prev,
                                                       const DevState* __restrict__ st,
                                                       double* __restrict__ parts)
{
    __shared__ double lds[4];
    const double shift = st->shift;
    double e1 = 0.0, e2 = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x; i < g.n;
         i += static_cast<int64_t>(gridDim.x) * BLOCK) {
        const double pc = p[i] - shift;
        p[i] = pc;
        e1 += fabs(pc - p_prev[i]);
        e2 += fabs(pc);
        p_prev[i] = pc;
    }
    const double s1 = block_sum(e1, lds);
    const double s2 = block_sum(e2, lds);
    if (threadIdx.x == 0) {
        parts[2 * blockIdx.x] = s1;
        parts[2 * blockIdx.x + 1] = s2;
    }
}

// Kinetic fixed-point residual of the frozen split map after the gauge
// (GPU/InamuroCUDA.cu:1987-2057 and 1114-1171, DUAL_RESIDUAL_CONTRACT.md).
// F(h,p) = BS[C(h,p)] is evaluated on the fly in pull form from the gauged
// state h = slip(streamed) - E*shift, p = gauged p; nothing but the six block
// sums is written.
__global__ __launch_bounds__(BLOCK) void k_fixed_point(
    Grid g, const double* __restrict__ hin, const double* __restrict__ p,
    const double* __restrict__ tau, const double* __restrict__ src,
    const DevState* __restrict__ st, double* __restrict__ parts)
{
    __shared__ double lds[4];
    const double shift = st->shift;
    const double gt = st->gauge_target;
    double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x; c < g.n;
         c += static_cast<int64_t>(gridDim.x) * BLOCK) {
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
            // slip view at the source cell: a +x population at x=0 (or a -x
            // one at x=lx-1) holds its mirror's value.
            int qq = q;
            if ((sx == 0 && CX[q] == 1) || (sx == g.lx - 1 && CX[q] == -1)) {
#pragma unroll
                for (int r = 0; r < 5; ++r) {
                    if (q == SLIP_POS[r]) qq = SLIP_NEG[r];
                    if (q == SLIP_NEG[r]) qq = SLIP_POS[r];
                }
            }
            double hq = hin[qq * g.n + s];
            hq -= EW[q] * shift;
            const double hequ = EW[q] * p[s];
            im[q] = hq - (hq - hequ) / tau[s] - (EW[q] / 3.0) * src[s];
        }
        if (x == 0) {
#pragma unroll
            for (int r = 0; r < 5; ++r) im[SLIP_POS[r]] = im[SLIP_NEG[r]];
        }
        if (x == g.lx - 1) {
#pragma unroll
            for (int r = 0; r < 5; ++r) im[SLIP_NEG[r]] = im[SLIP_POS[r]];
        }
        double hv[Q];
        load_streamed(hin, c, x, g, hv);
        double l0 = 0.0, l1 = 0.0, l2 = 0.0, hsum = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            hv[q] -= EW[q] * shift;
            const double ge = EW[q] * gt;
            l0 += fabs(hv[q] - im[q]);
            l1 += fabs(hv[q] - ge);
            l2 += fabs(im[q] - ge);
            hsum += hv[q];
        }
        const double pv = p[c];
        t[0] += l0;
        t[1] += l1;
        t[2] += l2;
        t[3] += fabs(pv - hsum);
        t[4] += fabs(pv - gt);
        t[5] += fabs(hsum - gt);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double s = block_sum(t[k], lds);
        if (threadIdx.x == 0) parts[6 * blockIdx.x + k] = s;
    }
}

__device__ __forceinline__ double rel_or_inf(double num, double den)
{
    if (den > 0.0) return num / den;
    return num == 0.0 ? 0.0 : INFINITY;
}

// Finishes a check: relative change, optional dual residual, convergence.
__global__ __launch_bounds__(BLOCK) void k_check_finalize(const double* __restrict__ parts_err,
                                                          const double* __restrict__ parts_fp,
                                                          int nparts, int dual, double tol,
                                                          DevState* st)
{
    __shared__ double lds[4];
    const double e1 = reduce_partials(parts_err, nparts, 2, 0, lds);
    const double e2 = reduce_partials(parts_err, nparts, 2, 1, lds);
    double fp[6] = {0, 0, 0, 0, 0, 0};
    if (dual) {
#pragma unroll
        for (int k = 0; k < 6; ++k) fp[k] = reduce_partials(parts_fp, nparts, 6, k, lds);
    }
    if (threadIdx.x != 0) return;
    int fp_ok = 1;
    if (dual) {
        for (int k = 0; k < 6; ++k) st->fp[k] = fp[k];
        bool finite = true;
        for (int k = 0; k < 6; ++k) finite = finite && isfinite(fp[k]);
        if (!finite) {
            if (st->status == ST_OK) st->status = ST_FP_NONFINITE;
        }
        const double hs = fp[1] > fp[2] ? fp[1] : fp[2];
        const double ps = fp[4] > fp[5] ? fp[4] : fp[5];
        st->fp_h_rel = rel_or_inf(fp[0], hs);
        st->fp_p_rel = rel_or_inf(fp[3], ps);
        st->fp_rel = st->fp_h_rel > st->fp_p_rel ? st->fp_h_rel : st->fp_p_rel;
        if (!isfinite(st->fp_rel) && st->status == ST_OK) st->status = ST_FP_REL_NONFINITE;
        st->fp_evaluated = 1;
        st->fp_converged = st->fp_rel < tol;
        // dual == 1: the exit also needs r_fp < tol; dual == 2: shadow only
        // (GPU/main.cpp:185-190).
        if (dual == 1) fp_ok = st->fp_converged;
    }
    st->e1 = e1;
    st->e2 = e2;
    if (!isfinite(e1) || !isfinite(e2)) {
        if (st->status == ST_OK) st->status = ST_PRESSURE_NONFINITE;
    }
    const double rel = rel_or_inf(e1, e2);
    st->rel = rel;
    if (!isfinite(rel) && st->status == ST_OK) st->status = ST_RESIDUAL_NONFINITE;
    st->pressure_converged = rel < tol;
    st->converged = (rel < tol) && fp_ok;
}

// correct_uvw, CPU/Inamuro.cpp:794-823.
__global__ __launch_bounds__(BLOCK) void k_correct(Grid g, const double* __restrict__ p,
                                                   const double* __restrict__ rho,
                                                   double* __restrict__ u, double* __restrict__ v,
                                                   double* __restrict__ w)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    const int x_e = (x == g.lx - 1) ? g.lx - 2 : x + 1;
    const int x_w = (x == 0) ? 1 : x - 1;
    const int y_n = (y == g.ly - 1) ? 0 : y + 1;
    const int y_s = (y == 0) ? g.ly - 1 : y - 1;
    const int z_n = (z == g.lz - 1) ? (g.zwall ? g.lz - 1 : 0) : z + 1;
    const int z_s = (z == 0) ? (g.zwall ? 0 : g.lz - 1) : z - 1;
    const double inv_2rho = 1.0 / (2.0 * rho[c]);
    u[c] -= (p[lin(x_e, y, z, g)] - p[lin(x_w, y, z, g)]) * inv_2rho;
    v[c] -= (p[lin(x, y_n, z, g)] - p[lin(x, y_s, z, g)]) * inv_2rho;
    w[c] -= (p[lin(x, y, z_n, g)] - p[lin(x, y, z_s, g)]) * inv_2rho;
}

__global__ __launch_bounds__(BLOCK) void k_materialize(Grid g, const double* __restrict__ in,
                                                       double* __restrict__ out)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    double o[Q];
    pull_slip(in, x, y, z, g, o);
#pragma unroll
    for (int q = 0; q < Q; ++q) out[q * g.n + c] = o[q];
}

__global__ __launch_bounds__(BLOCK) void k_streamed_view(Grid g, const double* __restrict__ in,
                                                         double* __restrict__ out)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    double hv[Q];
    load_streamed(in, c, x, g, hv);
#pragma unroll
    for (int q = 0; q < Q; ++q) out[q * g.n + c] = hv[q];
}

__global__ __launch_bounds__(BLOCK) void k_equilibrium_h(Grid g, const double* __restrict__ p,
                                                         double* __restrict__ h)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    const double pc = p[c];
#pragma unroll
    for (int q = 0; q < Q; ++q) h[q * g.n + c] = pc * EW[q];
}

// Sum of the projected source over the grid, from k_macro's block sums
// (GPU/InamuroCUDA.cu:1897-1920).
__global__ __launch_bounds__(BLOCK) void k_projected_finalize(const double* __restrict__ parts,
                                                              int nparts, DevState* st)
{
    __shared__ double lds[4];
    const double s = reduce_partials(parts, nparts, 1, 0, lds);
    if (threadIdx.x == 0) st->proj_div_sum = s;
}

// Fixed-order total of nparts block sums into *out.
__global__ __launch_bounds__(BLOCK) void k_sum_finalize(const double* __restrict__ parts, int nparts,
                                                        double* __restrict__ out)
{
    __shared__ double lds[4];
    const double s = reduce_partials(parts, nparts, 1, 0, lds);
    if (threadIdx.x == 0) *out = s;
}

} // namespace

void collide(hipStream_t s, int mode, const Grid& g, const Phys& ph, const double* fei,
             const double* rho, const double* u, const double* v, const double* w,
             const double* p, const double* fin, const double* gin, double* fout,
             double* gout, double* ux, double* vy, double* wz, double* partials)
{
    const int nb = stencil_blocks(g);
    switch (mode) {
    case FG_FIRST:
        k_collide<FG_FIRST><<<nb, BLOCK, 0, s>>>(g, ph, fei, rho, u, v, w, p, fin, gin, fout, gout,
                                                 ux, vy, wz, partials);
        break;
    case FG_PULL:
        k_collide<FG_PULL><<<nb, BLOCK, 0, s>>>(g, ph, fei, rho, u, v, w, p, fin, gin, fout, gout,
                                                ux, vy, wz, partials);
        break;
    default:
        k_collide<FG_LOCAL><<<nb, BLOCK, 0, s>>>(g, ph, fei, rho, u, v, w, p, fin, gin, fout, gout,
                                                 ux, vy, wz, partials);
        break;
    }
}

void source_finalize(hipStream_t s, const Grid&, const Phys& ph, const double* partials,
                     int nparts, DevState* st)
{
    k_source_finalize<<<1, BLOCK, 0, s>>>(ph, partials, nparts, st);
}

void macro(hipStream_t s, int pull, const Grid& g, const Phys& ph, const double* fin,
           const double* gin, const double* ux, const double* vy, const double* wz,
           DevState* st, double* fei, double* rho, double* u, double* v, double* w,
           double* tau, double* src, double* partials)
{
    const int nb = grid_blocks(g.n);
    if (pull)
        k_macro<1><<<nb, BLOCK, 0, s>>>(g, ph, fin, gin, ux, vy, wz, st, fei, rho, u, v, w, tau, src,
                                        partials);
    else
        k_macro<0><<<nb, BLOCK, 0, s>>>(g, ph, fin, gin, ux, vy, wz, st, fei, rho, u, v, w, tau, src,
                                        partials);
    k_projected_finalize<<<1, BLOCK, 0, s>>>(partials, nb, st);
}

void projected_finalize(hipStream_t s, const double* parts, int nparts, DevState* st)
{
    k_projected_finalize<<<1, BLOCK, 0, s>>>(parts, nparts, st);
}

void poisson_iter(hipStream_t s, int mode, const Grid& g, const double* hin, double* hout,
                  const double* p, const double* tau, const double* src, const DevState* st,
                  int config, double relax)
{
    // PUSH_CONFIGS: 0 = 256 threads with non-temporal loads, 1 = 256 plain,
    // 2 = 768 plain.
#define LBM_PUSH_AT(M, NT, PB, R)                                                                     \
    k_poisson_push<M, NT, PB, R><<<static_cast<int>((g.n + PB - 1) / PB), PB, 0, s>>>(g, hin, hout, p, \
                                                                                     tau, src, st, relax)
#define LBM_PUSH_R(M, R)                                                                     \
    do {                                                                                     \
        if (config == 1)                                                                     \
            LBM_PUSH_AT(M, 0, 256, R);                                                       \
        else if (config == 2)                                                                \
            LBM_PUSH_AT(M, 0, 768, R);                                                       \
        else                                                                                 \
            LBM_PUSH_AT(M, 1, 256, R);                                                       \
    } while (0)
#define LBM_PUSH(M)                  \
    do {                             \
        if (relax != 1.0)            \
            LBM_PUSH_R(M, 1);        \
        else                         \
            LBM_PUSH_R(M, 0);        \
    } while (0)
    switch (mode) {
    case H_EQUILIBRIUM: LBM_PUSH(H_EQUILIBRIUM); break;
    case H_GAUGED: LBM_PUSH(H_GAUGED); break;
    case H_EXPLICIT: LBM_PUSH(H_EXPLICIT); break;
    default: LBM_PUSH(H_STREAMED); break;
    }
#undef LBM_PUSH
#undef LBM_PUSH_R
#undef LBM_PUSH_AT
}

void getp_partials(hipStream_t s, const Grid& g, const double* hin, double* p, double* parts)
{
    k_getp_sum<<<red_blocks(g.n), BLOCK, 0, s>>>(g, hin, p, parts);
}

void check(hipStream_t s, const Grid& g, const Phys& ph, double* p, double* p_prev,
           const double* hin, const double* tau, const double* src, int dual, double tol,
           double* partials, DevState* st)
{
    const int nb = red_blocks(g.n);
    double* parts_sum = partials;
    double* parts_err = partials + RED_BLOCKS;
    double* parts_fp = partials + 3 * RED_BLOCKS;
    k_getp_sum<<<nb, BLOCK, 0, s>>>(g, hin, p, parts_sum);
    k_gauge_finalize<<<1, BLOCK, 0, s>>>(ph, parts_sum, nb, st);
    k_gauge_error<<<nb, BLOCK, 0, s>>>(g, p, p_prev, st, parts_err);
    if (dual) k_fixed_point<<<nb, BLOCK, 0, s>>>(g, hin, p, tau, src, st, parts_fp);
    k_check_finalize<<<1, BLOCK, 0, s>>>(parts_err, parts_fp, nb, dual, tol, st);
}

void gauge(hipStream_t s, const Grid& g, const Phys& ph, const double* hin, double* p,
           double* p_prev, double* partials, DevState* st)
{
    const int nb = red_blocks(g.n);
    double* parts_sum = partials;
    double* parts_err = partials + RED_BLOCKS;
    k_getp_sum<<<nb, BLOCK, 0, s>>>(g, hin, p, parts_sum);
    k_gauge_finalize<<<1, BLOCK, 0, s>>>(ph, parts_sum, nb, st);
    k_gauge_error<<<nb, BLOCK, 0, s>>>(g, p, p_prev, st, parts_err);
}

void gauge_finalize(hipStream_t s, const Phys& ph, const double* parts, int nparts, DevState* st)
{
    k_gauge_finalize<<<1, BLOCK, 0, s>>>(ph, parts, nparts, st);
}

void gauge_error(hipStream_t s, const Grid& g, double* p, double* p_prev, const DevState* st,
                 double* parts)
{
    k_gauge_error<<<red_blocks(g.n), BLOCK, 0, s>>>(g, p, p_prev, st, parts);
}

void check_finalize(hipStream_t s, const double* parts_err, const double* parts_fp, int nparts,
                    int dual, double tol, DevState* st)
{
    k_check_finalize<<<1, BLOCK, 0, s>>>(parts_err, parts_fp, nparts, dual, tol, st);
}

void mass_finalize(hipStream_t s, const double* parts, int nparts, DevState* st)
{
    k_sum_finalize<<<1, BLOCK, 0, s>>>(parts, nparts, &st->mass);
}

void streamed_to_post_boundary(hipStream_t s, const Grid& g, const double* hin, double* out)
{
    k_streamed_view<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, hin, out);
}

void correct(hipStream_t s, const Grid& g, const double* p, const double* rho, double* u,
             double* v, double* w)
{
    k_correct<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, p, rho, u, v, w);
}

void materialize(hipStream_t s, const Grid& g, const double* in, double* out)
{
    k_materialize<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, in, out);
}

void equilibrium_h(hipStream_t s, const Grid& g, const double* p, double* h)
{
    k_equilibrium_h<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, p, h);
}

void sum_field(hipStream_t s, const Grid& g, const double* a, double* partials, DevState* st)
{
    const int nb = red_blocks(g.n);
    k_sum_strided<<<nb, BLOCK, 0, s>>>(g, a, partials);
    k_sum_finalize<<<1, BLOCK, 0, s>>>(partials, nb, &st->mass);
}

void sum_field_into(hipStream_t s, const Grid& g, const double* a, double* partials, double* out)
{
    const int nb = red_blocks(g.n);
    k_sum_strided<<<nb, BLOCK, 0, s>>>(g, a, partials);
    k_sum_finalize<<<1, BLOCK, 0, s>>>(partials, nb, out);
}

} // namespace f64
} // namespace lbm
