// kernels_mixed.hip -- mixed-precision time step for gfx950.
//
// Storage and collision arithmetic in FP32; the pressure, p_prev and every
// global sum in FP64.  Distributions are stored as small deviations so that
// FP32 keeps their information:
//   g'_q = g_q - E_q                    (rest-weight equilibrium of g, whose
//                                        zeroth moment is ~1)
//   f'_0 = f_0 - phi(x),  f'_q = f_q    (f's rest population carries phi;
//                                        the moving ones are O(kappa |grad phi|^2))
//   h'_q(x) = h_q(x) - E_q b(x - c_q)   (streamed pressure populations relative
//                                        to the equilibrium of the cell they came
//                                        from; b is that cell's pressure rounded
//                                        to FP32 and stored beside h, so encoder
//                                        and decoder use the same number)
// In these variables the BGK relaxations are exact rewrites of the reference
// ones (CPU/Inamuro.cpp:224-245, 755-778):
//   f' <- f' - (f' - feq')/tau_f,   g' <- g' - (g' - geq')/tau_g + force,
//   h* - E p = (h - E p)(1 - 1/tau_h) - E S/3,
// The FP64 pressure itself is formed only where it is needed: at the
// convergence checks (getp), for the gauge, the residual and the velocity
// correction.
// The macro kernel re-bases f'_0 on the new phi in place (q = 0 never moves),
// so no kernel needs a neighbour's phi to decode f.
#include "kernels.hpp"

namespace lbm {
namespace mx {

namespace {

constexpr float EWF[Q] = {2.0f / 9.0f,
                          1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f,
                          1.0f / 72.0f, 1.0f / 72.0f, 1.0f / 72.0f, 1.0f / 72.0f,
                          1.0f / 72.0f, 1.0f / 72.0f, 1.0f / 72.0f, 1.0f / 72.0f};
constexpr float FWF[Q] = {-7.0f / 3.0f,
                          1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f,
                          1.0f / 24.0f, 1.0f / 24.0f, 1.0f / 24.0f, 1.0f / 24.0f,
                          1.0f / 24.0f, 1.0f / 24.0f, 1.0f / 24.0f, 1.0f / 24.0f};

// Products with a lattice weight.  Rounded to FP32 the weights are all off
// the same way (2/9, 1/9 and 1/72 up by 7.5e-9 relative, 1/3 and 1/24 up by
// 3e-8, -7/3 down by 3.4e-8), so w x with the rounded weight carries one
// relative bias in every direction of every cell, and the moments of the
// equilibria (the momentum of g, the phi flux of f) gain that factor on
// every step; the bias then accumulates over a run instead of averaging
// out.  Each weight is kept as the pair hi + lo (hi the FP32 weight, lo its
// rounding error) and multiplied as fma(hi, x, lo x): exact but for the
// last rounding, which falls either way.
constexpr float lo_part(double w) { return static_cast<float>(w - static_cast<double>(static_cast<float>(w))); }
constexpr float EWF_LO[Q] = {lo_part(2.0 / 9.0),
                             lo_part(1.0 / 9.0), lo_part(1.0 / 9.0), lo_part(1.0 / 9.0),
                             lo_part(1.0 / 9.0), lo_part(1.0 / 9.0), lo_part(1.0 / 9.0),
                             lo_part(1.0 / 72.0), lo_part(1.0 / 72.0), lo_part(1.0 / 72.0), lo_part(1.0 / 72.0),
                             lo_part(1.0 / 72.0), lo_part(1.0 / 72.0), lo_part(1.0 / 72.0), lo_part(1.0 / 72.0)};
constexpr float FWF_LO[Q] = {lo_part(-7.0 / 3.0),
                             lo_part(1.0 / 3.0), lo_part(1.0 / 3.0), lo_part(1.0 / 3.0),
                             lo_part(1.0 / 3.0), lo_part(1.0 / 3.0), lo_part(1.0 / 3.0),
                             lo_part(1.0 / 24.0), lo_part(1.0 / 24.0), lo_part(1.0 / 24.0), lo_part(1.0 / 24.0),
                             lo_part(1.0 / 24.0), lo_part(1.0 / 24.0), lo_part(1.0 / 24.0), lo_part(1.0 / 24.0)};

// E_q x and F_q x
__device__ __forceinline__ float ew(int q, float x) { return fmaf(EWF[q], x, EWF_LO[q] * x); }
__device__ __forceinline__ float fw(int q, float x) { return fmaf(FWF[q], x, FWF_LO[q] * x); }

// c_a * x for a compile-time lattice component c_a in {-1, 0, 1}.
template <int C>
__device__ __forceinline__ float cm(float x)
{
    return C == 0 ? 0.0f : (C > 0 ? x : -x);
}

__device__ __forceinline__ void slip(float (&o)[Q], int x, int lx)
{
    if (x == 0) {
#pragma unroll
        for (int r = 0; r < 5; ++r) o[SLIP_POS[r]] = o[SLIP_NEG[r]];
    }
    if (x == lx - 1) {
#pragma unroll
        for (int r = 0; r < 5; ++r) o[SLIP_NEG[r]] = o[SLIP_POS[r]];
    }
}

__device__ __forceinline__ void slip(double (&o)[Q], int x, int lx)
{
    if (x == 0) {
#pragma unroll
        for (int r = 0; r < 5; ++r) o[SLIP_POS[r]] = o[SLIP_NEG[r]];
    }
    if (x == lx - 1) {
#pragma unroll
        for (int r = 0; r < 5; ++r) o[SLIP_NEG[r]] = o[SLIP_POS[r]];
    }
}

// Pull of the post-collision populations streaming into (x,y,z), periodic on
// every axis, then the x-slip overwrite.  Deviations copy like full values
// because mirror directions share their weight.
__device__ __forceinline__ void pull_slip(const float* __restrict__ in, int x, int y, int z,
                                          const Grid& g, float (&o)[Q])
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
    slip(o, x, g.lx);
}

template <int K>
__device__ __forceinline__ void grad_acc(float v, float& ax, float& ay, float& az, float& sum)
{
    ax += cm<CX[K]>(v);
    ay += cm<CY[K]>(v);
    az += cm<CZ[K]>(v);
    sum += v;
}

// D3Q15 gradient (sum_k c_k psi_k / 10) and Laplacian ((sum_k psi_k - 14 psi)/5)
// of one field, CPU/Inamuro.hpp:213-314, summed over the differences
// d_k = psi_k - psi: sum_k c_k psi_k = sum_k c_k d_k because sum_k c_k = 0,
// and sum_k psi_k - 14 psi = sum_k d_k.  A difference of two FP32 values
// within a factor of two of each other is exact, so where the field is smooth
// the sums round relative to the gradient and the Laplacian themselves;
// summed over the raw values they round relative to psi, and the Laplacian's
// sum_k psi_k - 14 psi cancels (in the liquid, phi ~ 0.09, to an absolute
// error of ~1e-8).
struct Stencil {
    float gx, gy, gz, lap, c;
};

__device__ __forceinline__ Stencil stencil(const float* a, const int (&idx)[Q])
{
    float ax = 0.f, ay = 0.f, az = 0.f, s = 0.f;
    float v[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) v[k] = a[idx[k]];
    const float c = v[0];
    grad_acc<1>(v[1] - c, ax, ay, az, s);  grad_acc<2>(v[2] - c, ax, ay, az, s);
    grad_acc<3>(v[3] - c, ax, ay, az, s);  grad_acc<4>(v[4] - c, ax, ay, az, s);
    grad_acc<5>(v[5] - c, ax, ay, az, s);  grad_acc<6>(v[6] - c, ax, ay, az, s);
    grad_acc<7>(v[7] - c, ax, ay, az, s);  grad_acc<8>(v[8] - c, ax, ay, az, s);
    grad_acc<9>(v[9] - c, ax, ay, az, s);  grad_acc<10>(v[10] - c, ax, ay, az, s);
    grad_acc<11>(v[11] - c, ax, ay, az, s); grad_acc<12>(v[12] - c, ax, ay, az, s);
    grad_acc<13>(v[13] - c, ax, ay, az, s); grad_acc<14>(v[14] - c, ax, ay, az, s);
    return Stencil{ax * 0.1f, ay * 0.1f, az * 0.1f, s * 0.2f, c};
}

template <int I>
__device__ __forceinline__ float dot3(float x, float y, float z)
{
    return cm<CX[I]>(x) + cm<CY[I]>(y) + cm<CZ[I]>(z);
}

// Collision of f and g in deviation form (Inamuro::collision,
// CPU/Inamuro.cpp:134-250) for direction I.
template <int I, int MODE>
__device__ __forceinline__ void collide_dir(const PhysF& ph, float feic, float uc, float vc,
                                            float wc, const Stencil& U, const Stencil& V,
                                            const Stencil& W, float rx, float ry, float rz,
                                            float fx, float fy, float fz, float P0p, float usq,
                                            float sum_fei, float sum_rho, float kg_rho,
                                            float force_scale, float& f, float& g)
{
    const float un = dot3<I>(uc, vc, wc);
    const float fe = dot3<I>(fx, fy, fz);
    const float re = dot3<I>(rx, ry, rz);
    constexpr float ci_sq = float(CX[I] * CX[I] + CY[I] * CY[I] + CZ[I] * CZ[I]);
    const float Gf = 4.5f * fe * fe - 1.5f * sum_fei * ci_sq;
    const float Gr = 4.5f * re * re - 1.5f * sum_rho * ci_sq;
    // f_eq without the H_q * phi term (it is the stored base of f'_0).
    const float feq = fw(I, P0p) + ew(I, 3.0f * feic * un + ph.k_f * Gf);
    // sum_ab c_a c_b d_b u_a for this direction.
    float cc = 0.f;
    if (CX[I] != 0) cc += cm<CX[I]>(cm<CX[I]>(U.gx) + cm<CY[I]>(U.gy) + cm<CZ[I]>(U.gz));
    if (CY[I] != 0) cc += cm<CY[I]>(cm<CX[I]>(V.gx) + cm<CY[I]>(V.gy) + cm<CZ[I]>(V.gz));
    if (CZ[I] != 0) cc += cm<CZ[I]>(cm<CX[I]>(W.gx) + cm<CY[I]>(W.gy) + cm<CZ[I]>(W.gz));
    // g_eq - E_q.
    const float geq = ew(I, 3.0f * un - 1.5f * usq + 4.5f * un * un + ph.visc * cc + kg_rho * Gr) -
                      fw(I, (2.0f / 3.0f) * kg_rho * sum_rho);
    if (MODE == FG_FIRST) {
        f = feq;
        g = geq;
    } else {
        const float lapc = dot3<I>(U.lap, V.lap, W.lap);
        f = f - (f - feq) * ph.tauf_inv;
        g = g - (g - geq) * ph.taug_inv + ew(I, 3.0f * force_scale * lapc);
    }
}

// The collision with its stencils (tile and LDS staging as the FP64 kernel:
// a STENCIL_TX x STENCIL_TY tile marching through STENCIL_ZCHUNK layers, u,
// v, w, rho and phi of each layer staged once with their halo).
template <int MODE>
__global__ __launch_bounds__(BLOCK) void k_collide(
    Grid g, PhysF ph, const float* __restrict__ fei, const float* __restrict__ rho,
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
    const double* __restrict__ p, const float* __restrict__ fin, const float* __restrict__ gin,
    float* __restrict__ fout, float* __restrict__ gout, float* __restrict__ ux_o,
    float* __restrict__ vy_o, float* __restrict__ wz_o, double* __restrict__ partials)
{
    __shared__ double lds[4];
    __shared__ float staged[5][3][STENCIL_PLANE];  // u, v, w, rho, phi; ring of 3 layers
    const int ntx = (g.lx + STENCIL_TX - 1) / STENCIL_TX, nty = (g.ly + STENCIL_TY - 1) / STENCIL_TY;
    const int b = static_cast<int>(blockIdx.x);
    const int x0 = (b % ntx) * STENCIL_TX, y0 = ((b / ntx) % nty) * STENCIL_TY;
    const int zb = (b / (ntx * nty)) * STENCIL_ZCHUNK, ze = min(g.lz, zb + STENCIL_ZCHUNK);
    auto stage = [&](int gz, int slot) {
        for (int e = threadIdx.x; e < STENCIL_PLANE; e += BLOCK) {
            const int64_t i = stencil_source(e, x0, y0, gz, g);
            staged[0][slot][e] = u[i];
            staged[1][slot][e] = v[i];
            staged[2][slot][e] = w[i];
            staged[3][slot][e] = rho[i];
            staged[4][slot][e] = fei[i];
        }
    };
    // layer L sits in ring slot (L - zb + 1) mod 3; the loads are issued
    // directly into LDS (holding a plane in registers as the FP64 kernel does
    // would cost this kernel two thirds of its occupancy)
    stage(z_stencil(zb, -1, g), 0);
    stage(zb, 1);
    const int tx = threadIdx.x % STENCIL_TX, ty = threadIdx.x / STENCIL_TX;
    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < g.lx && y < g.ly;
    double div_c = 0.0, p_c = 0.0, adiv_c = 0.0;
    for (int z = zb; z < ze; ++z) {
        const int sm = (z - zb) % 3, s0 = (z - zb + 1) % 3, sp = (z - zb + 2) % 3;
        stage(z_stencil(z, 1, g), sp);
        __syncthreads();
        if (valid) {
            const int64_t c = lin(x, y, z, g);
            // flat index into staged[f]: ring slot of the neighbour's layer, then
            // its place in the plane
            int idx[Q];
#pragma unroll
            for (int k = 0; k < Q; ++k)
                idx[k] = (CZ[k] == 1 ? sp : (CZ[k] == -1 ? sm : s0)) * STENCIL_PLANE + stencil_slot(tx, ty, k);
            const Stencil U = stencil(&staged[0][0][0], idx);
            const Stencil V = stencil(&staged[1][0][0], idx);
            const Stencil W = stencil(&staged[2][0][0], idx);
            const Stencil R = stencil(&staged[3][0][0], idx);
            const Stencil F = stencil(&staged[4][0][0], idx);
            const float feic = F.c, rhoc = R.c;
            const float mu = (rhoc - ph.rho_G) * ph.tmp1 + ph.mu_G;
            const float sum_fei = F.gx * F.gx + F.gy * F.gy + F.gz * F.gz;
            const float sum_rho = R.gx * R.gx + R.gy * R.gy + R.gz * R.gz;
            const float usq = U.c * U.c + V.c * V.c + W.c * W.c;
            const float p0 = feic * ph.T / (1.0f - ph.b * feic) - ph.a * feic * feic;
            const float P0p = p0 - ph.k_f * feic * F.lap - ph.k_f * sum_fei * (1.0f / 6.0f);
            const float inv_rho = 1.0f / rhoc;
            const float kg_rho = ph.k_g * inv_rho;
            const float force_scale = inv_rho * mu;
            float fv[Q], gv[Q];
            if (MODE == FG_PULL) {
                pull_slip(fin, x, y, z, g, fv);
                pull_slip(gin, x, y, z, g, gv);
            } else if (MODE == FG_LOCAL) {
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    fv[q] = fin[q * g.n + c];
                    gv[q] = gin[q * g.n + c];
                }
            } else {
#pragma unroll
                for (int q = 0; q < Q; ++q) fv[q] = gv[q] = 0.f;
            }
#define LBM_MX_DIR(I)                                                                           \
        collide_dir<I, MODE>(ph, feic, U.c, V.c, W.c, U, V, W, R.gx, R.gy, R.gz, F.gx, F.gy, F.gz, \
                             P0p, usq, sum_fei, sum_rho, kg_rho, force_scale, fv[I], gv[I]);
            LBM_MX_DIR(0) LBM_MX_DIR(1) LBM_MX_DIR(2) LBM_MX_DIR(3) LBM_MX_DIR(4)
            LBM_MX_DIR(5) LBM_MX_DIR(6) LBM_MX_DIR(7) LBM_MX_DIR(8) LBM_MX_DIR(9)
            LBM_MX_DIR(10) LBM_MX_DIR(11) LBM_MX_DIR(12) LBM_MX_DIR(13) LBM_MX_DIR(14)
#undef LBM_MX_DIR
            if (ph.gz != 0.f) {
                // extension: buoyant acceleration gz (1 - rho_G/rho) on the
                // predicted velocity
                const float az = ph.gz * (1.0f - ph.rho_G * inv_rho);
#pragma unroll
                for (int q = 0; q < Q; ++q)
                    if (CZ[q] != 0) gv[q] += ew(q, 3.0f * static_cast<float>(CZ[q]) * az);
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                fout[q * g.n + c] = fv[q];
                gout[q * g.n + c] = gv[q];
            }
            ux_o[c] = U.gx;
            vy_o[c] = V.gy;
            wz_o[c] = W.gz;
            const double div = static_cast<double>(U.gx) + static_cast<double>(V.gy) +
                               static_cast<double>(W.gz);
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

// getMacro (CPU/Inamuro.cpp:294-346) on the pulled deviations; re-bases f'_0
// on the new phi; writes omega = 2 - 1/tau_h (the relaxation factor
// 1 - 1/tau_h plus one) and S/3 for the Poisson solve.
__device__ __forceinline__ double macro_cell(
    int64_t c, const Grid& g, const PhysF& ph, float* __restrict__ fpost, const float* __restrict__ gpost,
    const float* __restrict__ ux, const float* __restrict__ vy, const float* __restrict__ wz,
    const DevState* __restrict__ st, float* __restrict__ fei, float* __restrict__ rho,
    float* __restrict__ u, float* __restrict__ v, float* __restrict__ w,
    float* __restrict__ omega, float* __restrict__ src3)
{
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    float fv[Q], gv[Q];
    pull_slip(fpost, x, y, z, g, fv);
    pull_slip(gpost, x, y, z, g, gv);
    float dev = 0.f;
#pragma unroll
    for (int q = 0; q < Q; ++q) dev += fv[q];
    const float phi_old = fei[c];
    const float phi = phi_old + dev;
    float density;
    if (phi <= ph.fei_G) {
        density = ph.rho_G;
    } else if (phi >= ph.fei_L) {
        density = ph.rho_L;
    } else {
        // (rho_L - rho_G)/2 (sin(pi (phi - mean)/(phi_L - phi_G)) + 1) + rho_G
        // (CPU/Inamuro.cpp:313-327) rewritten with 1 + sin(x) = 2 sin^2(x/2 + pi/4)
        // as rho_G + (rho_L - rho_G) sin^2(pi/2 (phi - phi_G)/(phi_L - phi_G)).
        // Equal in exact arithmetic; in FP32 sin(x) + 1 cancels on the gas side
        // of the interface (sin ~ -1), which cost rho there up to 3e-5 relative.
        const float sn = sinf((phi - ph.fei_G) * ph.rho_arg);
        density = ph.rho_G + ph.drho * (sn * sn);
    }
    float uu = 0.f, vv = 0.f, ww = 0.f;
#pragma unroll
    for (int q = 1; q < Q; ++q) {
        if (CX[q] != 0) uu += CX[q] > 0 ? gv[q] : -gv[q];
        if (CY[q] != 0) vv += CY[q] > 0 ? gv[q] : -gv[q];
        if (CZ[q] != 0) ww += CZ[q] > 0 ? gv[q] : -gv[q];
    }
    fpost[c] = (phi_old - phi) + fv[0];  // f'_0 relative to the new phi (q = 0 is local)
    fei[c] = phi;
    rho[c] = density;
    u[c] = uu;
    v[c] = vv;
    w[c] = ww;
    // 1 + (1 - 1/tau_h) = 2 - 1/tau_h with tau_h = 1/rho + 1/2, i.e.
    // 4/(rho + 2): about 0.004 in the liquid, kept to FP32's relative
    // precision (1 - 1/tau_h itself is -0.996 there, and rounding it to FP32
    // would shift the small part that sets the pressure's diffusivity by up
    // to 1.5e-5 relative).
    omega[c] = 4.0f / (density + 2.0f);
    const float m = static_cast<float>(st->div_mean);
    const float sc = ((ux[c] - m) + vy[c]) + wz[c];
    src3[c] = sc * (1.0f / 3.0f);
    return static_cast<double>(sc);
}

__global__ __launch_bounds__(BLOCK) void k_macro(
    Grid g, PhysF ph, float* __restrict__ fpost, const float* __restrict__ gpost,
    const float* __restrict__ ux, const float* __restrict__ vy, const float* __restrict__ wz,
    const DevState* __restrict__ st, float* __restrict__ fei, float* __restrict__ rho,
    float* __restrict__ u, float* __restrict__ v, float* __restrict__ w,
    float* __restrict__ omega, float* __restrict__ src3, double* __restrict__ parts)
{
    __shared__ double lds[4];
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    double s_c = 0.0;
    if (c < g.n) s_c = macro_cell(c, g, ph, fpost, gpost, ux, vy, wz, st, fei, rho, u, v, w, omega, src3);
    // Block sums of the projected source (GPU/InamuroCUDA.cu:1897-1920).
    const double bs = block_sum(s_c, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = bs;
}

// Reconstructs the post-boundary FP64 pressure populations of cell c from the
// streamed deviations and the FP64 pressures of the cells they came from.
template <int NT = 0>
__device__ __forceinline__ void decode_h(const float* __restrict__ hin, const float* __restrict__ pbase,
                                         int64_t c, int x, int y, int z, const Grid& g, double (&H)[Q])
{
    const int xm = wrap(x, -1, g.lx), xp = wrap(x, 1, g.lx);
    const int ym = wrap(y, -1, g.ly), yp = wrap(y, 1, g.ly);
    const int zm = wrap(z, -1, g.lz), zp = wrap(z, 1, g.lz);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int sx = CX[q] == 1 ? xm : (CX[q] == -1 ? xp : x);
        const int sy = CY[q] == 1 ? ym : (CY[q] == -1 ? yp : y);
        const int sz = CZ[q] == 1 ? zm : (CZ[q] == -1 ? zp : z);
        // a value bounced at a z wall was encoded by this cell, with its base
        const int64_t src = z_bounced_in(q, z, g) ? c : lin(sx, sy, sz, g);
        const float hq = NT ? __builtin_nontemporal_load(&hin[q * g.n + c]) : hin[q * g.n + c];
        H[q] = EW[q] * static_cast<double>(pbase[src]) + static_cast<double>(hq);
    }
    slip(H, x, g.lx);
}

// One Poisson iteration in push form (see kernels_f64.hip for the FP64 twin).
// Every mode writes this cell's FP32 base b = float(p) for the next decode.
// NT: the streamed h is loaded with a non-temporal hint; PB: threads per
// block (the caller picks one of PUSH_CONFIGS).
template <int MODE, int NT, int PB>
__global__ __launch_bounds__(PB) void k_poisson_push(
    Grid g, const float* __restrict__ hin, float* __restrict__ hout,
    const float* __restrict__ bbase, const double* __restrict__ pown, float* __restrict__ bout,
    const float* __restrict__ omega, const float* __restrict__ src3,
    const DevState* __restrict__ st)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * PB + threadIdx.x;
    if (c >= g.n) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    float dev[Q];
    double pn;
    if (MODE == H_EQUILIBRIUM) {
        pn = pown[c];
#pragma unroll
        for (int q = 0; q < Q; ++q) dev[q] = 0.f;
    } else {
        double H[Q];
        if (MODE == H_EXPLICIT) {
#pragma unroll
            for (int q = 0; q < Q; ++q) H[q] = static_cast<double>(hin[q * g.n + c]);
            pn = pown[c];
        } else {
            decode_h<NT>(hin, bbase, c, x, y, z, g, H);
            if (MODE == H_GAUGED) {
                const double shift = st->shift;
#pragma unroll
                for (int q = 0; q < Q; ++q) H[q] -= EW[q] * shift;
                pn = pown[c];
            } else {
                pn = 0.0;
#pragma unroll
                for (int q = 0; q < Q; ++q) pn += H[q];
            }
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) dev[q] = static_cast<float>(H[q] - EW[q] * pn);
    }
    // Encoding base: STREAMED iterations keep the base set by the segment's
    // first iteration (after a check, or at the step's start), so the base
    // array is written once per check interval, not once per iteration.
    float b;
    if (MODE == H_STREAMED) {
        b = bbase[c];
    } else {
        b = static_cast<float>(pn);
        bout[c] = b;
    }
    const double rb = pn - static_cast<double>(b);  // exact in FP64
    const float om = omega[c];
    const float s3 = src3[c];
    const int xm = wrap(x, -1, g.lx), xp = wrap(x, 1, g.lx);
    const int ym = wrap(y, -1, g.ly), yp = wrap(y, 1, g.ly);
    const int zm = wrap(z, -1, g.lz), zp = wrap(z, 1, g.lz);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        // h* - E_q b = (h - E_q p)(1 - 1/tau_h) - E_q S/3 + E_q (p - b), with
        // d (1 - 1/tau_h) = d a - d formed in one rounding (a = 2 - 1/tau_h)
        const float o = fmaf(dev[q], om, -dev[q]) - ew(q, s3) + static_cast<float>(EW[q] * rb);
        const int dx = CX[q] == 1 ? xp : (CX[q] == -1 ? xm : x);
        const int dy = CY[q] == 1 ? yp : (CY[q] == -1 ? ym : y);
        const int dz = CZ[q] == 1 ? zp : (CZ[q] == -1 ? zm : z);
        if (z_bounced_out(q, z, g))
            hout[OPP[q] * g.n + c] = o;  // bounce-back at a z wall (E_opp = E_q)
        else
            hout[q * g.n + lin(dx, dy, dz, g)] = o;
    }
}

__global__ __launch_bounds__(BLOCK) void k_getp_sum(Grid g, const float* __restrict__ hin,
                                                    const float* __restrict__ pbase,
                                                    double* __restrict__ p,
                                                    double* __restrict__ parts)
{
    __shared__ double lds[4];
    double acc = 0.0;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x; c < g.n;
         c += static_cast<int64_t>(gridDim.x) * BLOCK) {
        int x, y, z;
        cell_xyz(c, g, x, y, z);
        double H[Q];
        decode_h(hin, pbase, c, x, y, z, g, H);
        double pc = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) pc += H[q];
        p[c] = pc;
        acc += pc;
    }
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// Dual residual (GPU/InamuroCUDA.cu:1114-1171) on the gauged mixed state.
__global__ __launch_bounds__(BLOCK) void k_fixed_point(
    Grid g, const float* __restrict__ hin, const float* __restrict__ pbase,
    const double* __restrict__ p, const float* __restrict__ omega,
    const float* __restrict__ src3, const DevState* __restrict__ st,
    double* __restrict__ parts)
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
            int qq = q;
            if ((sx == 0 && CX[q] == 1) || (sx == g.lx - 1 && CX[q] == -1)) {
#pragma unroll
                for (int r = 0; r < 5; ++r) {
                    if (q == SLIP_POS[r]) qq = SLIP_NEG[r];
                    if (q == SLIP_NEG[r]) qq = SLIP_POS[r];
                }
            }
            const int cqx = qq == q ? CX[q] : -CX[q];
            const int ssx = wrap(sx, -cqx, g.lx);
            const int ssy = wrap(sy, -CY[q], g.ly);
            const int ssz = wrap(sz, -CZ[q], g.lz);
            const int64_t s = lin(sx, sy, sz, g);
            const double Hq = EW[q] * static_cast<double>(pbase[lin(ssx, ssy, ssz, g)]) +
                              static_cast<double>(hin[qq * g.n + s]) - EW[q] * shift;
            const double ps = p[s];
            im[q] = EW[q] * ps + ((Hq - EW[q] * ps) * (static_cast<double>(omega[s]) - 1.0) -
                                  EW[q] * static_cast<double>(src3[s]));
        }
        slip(im, x, g.lx);
        double hv[Q];
        decode_h(hin, pbase, c, x, y, z, g, hv);
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

// correct_uvw, CPU/Inamuro.cpp:794-823: the FP64 pressure difference, then FP32.
__global__ __launch_bounds__(BLOCK) void k_correct(Grid g, const double* __restrict__ p,
                                                   const float* __restrict__ rho,
                                                   float* __restrict__ u, float* __restrict__ v,
                                                   float* __restrict__ w)
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
    const float inv_2rho = 0.5f / rho[c];
    u[c] -= static_cast<float>(p[lin(x_e, y, z, g)] - p[lin(x_w, y, z, g)]) * inv_2rho;
    v[c] -= static_cast<float>(p[lin(x, y_n, z, g)] - p[lin(x, y_s, z, g)]) * inv_2rho;
    w[c] -= static_cast<float>(p[lin(x, y, z_n, g)] - p[lin(x, y, z_s, g)]) * inv_2rho;
}

__global__ __launch_bounds__(BLOCK) void k_sum_f32(Grid g, const float* __restrict__ a,
                                                   double* __restrict__ parts)
{
    __shared__ double lds[4];
    double acc = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x; i < g.n;
         i += static_cast<int64_t>(gridDim.x) * BLOCK)
        acc += static_cast<double>(a[i]);
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

__global__ __launch_bounds__(BLOCK) void k_to_f64(int64_t n, const float* __restrict__ in,
                                                  double* __restrict__ out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (i < n) out[i] = static_cast<double>(in[i]);
}

__global__ __launch_bounds__(BLOCK) void k_to_f32(int64_t n, const double* __restrict__ in,
                                                  float* __restrict__ out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (i < n) out[i] = static_cast<float>(in[i]);
}

template <int PULLED, int IS_F>
__global__ __launch_bounds__(BLOCK) void k_fg_full(Grid g, const float* __restrict__ in,
                                                   const float* __restrict__ fei,
                                                   double* __restrict__ out)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    float v[Q];
    if (PULLED) {
        pull_slip(in, x, y, z, g, v);
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q] = in[q * g.n + c];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double o = static_cast<double>(v[q]);
        if (IS_F && q == 0) o += static_cast<double>(fei[c]);
        if (!IS_F) o += EW[q];
        out[q * g.n + c] = o;
    }
}

template <int IS_F>
__global__ __launch_bounds__(BLOCK) void k_fg_encode(Grid g, const double* __restrict__ in,
                                                     const float* __restrict__ fei,
                                                     float* __restrict__ out)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (c >= g.n) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double o = in[q * g.n + c];
        if (IS_F && q == 0) o -= static_cast<double>(fei[c]);
        if (!IS_F) o -= EW[q];
        out[q * g.n + c] = static_cast<float>(o);
    }
}

} // namespace

void collide(hipStream_t s, int mode, const Grid& g, const PhysF& ph, const float* fei,
             const float* rho, const float* u, const float* v, const float* w, const double* p,
             const float* fin, const float* gin, float* fout, float* gout, float* ux, float* vy,
             float* wz, double* partials)
{
    const int nb = stencil_blocks(g);
    switch (mode) {
    case FG_FIRST:
        k_collide<FG_FIRST><<<nb, BLOCK, 0, s>>>(g, ph, fei, rho, u, v, w, p, fin, gin, fout, gout, ux, vy, wz, partials);
        break;
    case FG_PULL:
        k_collide<FG_PULL><<<nb, BLOCK, 0, s>>>(g, ph, fei, rho, u, v, w, p, fin, gin, fout, gout, ux, vy, wz, partials);
        break;
    default:
        k_collide<FG_LOCAL><<<nb, BLOCK, 0, s>>>(g, ph, fei, rho, u, v, w, p, fin, gin, fout, gout, ux, vy, wz, partials);
        break;
    }
}

void macro(hipStream_t s, const Grid& g, const PhysF& ph, float* fpost, const float* gpost,
           const float* ux, const float* vy, const float* wz, DevState* st, float* fei,
           float* rho, float* u, float* v, float* w, float* omega, float* src3, double* partials)
{
    const int nb = grid_blocks(g.n);
    k_macro<<<nb, BLOCK, 0, s>>>(g, ph, fpost, gpost, ux, vy, wz, st, fei, rho, u, v, w, omega, src3,
                                 partials);
    f64::projected_finalize(s, partials, nb, st);
}

void poisson_iter(hipStream_t s, int mode, const Grid& g, const float* hin, float* hout,
                  const float* bbase, const double* pown, float* bout, const float* omega,
                  const float* src3, const DevState* st, int config)
{
    // PUSH_CONFIGS: 0 = 1024 threads with non-temporal loads, 1 = 512 with
    // non-temporal loads, 2 = 512 plain, 3 = 1024 plain.
#define LBM_MX_PUSH_AT(M, NT, PB)                                                                    \
    k_poisson_push<M, NT, PB><<<static_cast<int>((g.n + PB - 1) / PB), PB, 0, s>>>(                  \
        g, hin, hout, bbase, pown, bout, omega, src3, st)
#define LBM_MX_PUSH(M)                                                                         \
    do {                                                                                       \
        if (config == 1)                                                                       \
            LBM_MX_PUSH_AT(M, 1, 512);                                                         \
        else if (config == 2)                                                                  \
            LBM_MX_PUSH_AT(M, 0, 512);                                                         \
        else if (config == 3)                                                                  \
            LBM_MX_PUSH_AT(M, 0, 1024);                                                        \
        else                                                                                   \
            LBM_MX_PUSH_AT(M, 1, 1024);                                                        \
    } while (0)
    switch (mode) {
    case H_EQUILIBRIUM: LBM_MX_PUSH(H_EQUILIBRIUM); break;
    case H_GAUGED: LBM_MX_PUSH(H_GAUGED); break;
    case H_EXPLICIT: LBM_MX_PUSH(H_EXPLICIT); break;
    default: LBM_MX_PUSH(H_STREAMED); break;
    }
#undef LBM_MX_PUSH
#undef LBM_MX_PUSH_AT
}

void getp_sum(hipStream_t s, const Grid& g, const float* hin, const float* pbase, double* p,
              double* parts)
{
    k_getp_sum<<<red_blocks(g.n), BLOCK, 0, s>>>(g, hin, pbase, p, parts);
}

void fixed_point(hipStream_t s, const Grid& g, const float* hin, const float* pbase,
                 const double* p, const float* omega, const float* src3, const DevState* st,
                 double* parts)
{
    k_fixed_point<<<red_blocks(g.n), BLOCK, 0, s>>>(g, hin, pbase, p, omega, src3, st, parts);
}

void correct(hipStream_t s, const Grid& g, const double* p, const float* rho, float* u, float* v,
             float* w)
{
    k_correct<<<grid_blocks(g.n), BLOCK, 0, s>>>(g, p, rho, u, v, w);
}

void sum_f32(hipStream_t s, const Grid& g, const float* a, double* parts)
{
    k_sum_f32<<<red_blocks(g.n), BLOCK, 0, s>>>(g, a, parts);
}

void to_f64(hipStream_t s, int64_t n, const float* in, double* out)
{
    k_to_f64<<<grid_blocks(n), BLOCK, 0, s>>>(n, in, out);
}

void to_f32(hipStream_t s, int64_t n, const double* in, float* out)
{
    k_to_f32<<<grid_blocks(n), BLOCK, 0, s>>>(n, in, out);
}

void fg_full(hipStream_t s, const Grid& g, int pulled, int is_f, const float* in, const float* fei,
             double* out)
{
    const int nb = grid_blocks(g.n);
    if (pulled && is_f) k_fg_full<1, 1><<<nb, BLOCK, 0, s>>>(g, in, fei, out);
    else if (pulled) k_fg_full<1, 0><<<nb, BLOCK, 0, s>>>(g, in, fei, out);
    else if (is_f) k_fg_full<0, 1><<<nb, BLOCK, 0, s>>>(g, in, fei, out);
    else k_fg_full<0, 0><<<nb, BLOCK, 0, s>>>(g, in, fei, out);
}

void fg_encode(hipStream_t s, const Grid& g, int is_f, const double* in, const float* fei,
               float* out)
{
    const int nb = grid_blocks(g.n);
    if (is_f) k_fg_encode<1><<<nb, BLOCK, 0, s>>>(g, in, fei, out);
    else k_fg_encode<0><<<nb, BLOCK, 0, s>>>(g, in, fei, out);
}

} // namespace mx
} // namespace lbm
