// tpsph_internal.h -- definitions shared by the translation units of libtpsph.so: the
// device parameters, the views of the cell-sorted data, the device helpers of the pair
// terms and the Adami pass, the handle and the host-side helpers.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "tpsph.h"

namespace tpsph {

constexpr int kBlock = 256;  // 4 wavefronts of 64
constexpr int kCap = 32;     // neighbour-list entries per lane and phase
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;
constexpr int kBrickThreads = 1024;   // threads of the brick kernel (one block per CU), two
                                      // per target
constexpr int kBrickCap = 2048;       // LDS records of a brick: the union plus one far record
constexpr int kBrickCapGrad = 1664;   // the same with DensityDiffusionAntuono's density
                                      // gradients in the records (3D: 96 instead of 72 bytes)
constexpr double kFar = 1e100;        // coordinate of the far record (never within R)
constexpr int kBrickMaxRows = 256;    // union rows per brick (one scan block)
constexpr int kGateBlocks = 1024;     // grid-stride kernels behind the rebuild flag: a small
                                      // grid keeps the launch cheap when the flag is down
constexpr int kPad = 4;               // NaN entries after the SoA coordinate arrays

// message of the last failed tpsph_create (tpsph_last_error(NULL))
extern thread_local std::string g_create_error;

struct DevParams {
    int nd;
    int nf, nw;
    int use_visc, use_dd, clip_fluid, clip_wall;
    double h, h_inv, R, R2, two_h, az2;
    double kval_c;   // kernel normalisation sigma * h^-d
    double kgrad_c;  // -5 sigma h^-d h^-2: grad W = kgrad_c (1 - q/2)^3 r_ab
    double rho0, gamma, inv_gamma, p_bg, B, c;
    double alpha, beta, eps_h2;  // epsilon h^2
    double av_c;                 // alpha c h
    double dd2;                  // 2 delta h c (0 without density diffusion)
    double g[3];
    double p_offset;
    // grid
    double lo[3];
    double cs[3], inv_cs[3];  // cell edge per axis
    int n[3];
    int ncells;
    double hs_fluid;  // half skin + rounding margin (fluid cells)
    double hs_wall;   // rounding margin (static wall cells)
    double half_skin, half_skin2;
    // Verlet lists: squared list radii and the cell growth that makes the pencils reach them
    double rl2_fluid, rl2_wall, hs_list_fluid, hs_list_wall;
    // FP32 pre-test of the fluid list build (positions relative to lo inside the grid box
    // box[d]): a squared FP32 distance at most list_in2 is surely within the list radius,
    // one above list_out2 surely beyond it
    double box[3];
    float list_in2, list_out2;
    float wlist_in2, wlist_out2;  // the same bounds for the wall list radius
    // bricks of cells (processing order of the list kernels): cells per brick, bricks per axis
    int bc[3], nb[3];
    // union rows of a brick: cells beyond the brick along y and z, rows per brick, and the
    // radius (list radius + margin) and cell growth (margin) of the union
    int ur[3], maxr;
    double rl_union2, hs_union;
    // LDS records of a brick (kBrickCap, or kBrickCapGrad with DensityDiffusionAntuono in
    // 3D): a union of up to brick_cap - 1 particles stages, the last record is the far record
    int brick_cap;
    // StateEquationAdaptiveCole: {c, B, av_c, dd2} of the current kick in device memory
    // (k_sound_speed writes them), or nullptr for the constants above
    const double *dyn;
    double mach, c_min, c_max;
    // viscosity of the wall model (no-slip): 0 nothing, 1 ArtificialViscosityMonaghan,
    // 2 ViscosityAdami; its alpha c h, beta, epsilon h^2 and nu, and the fluid's own
    // kinematic viscosity alpha h c / (2 ndims + 4) (viscosity.jl:82-87)
    int visc_w;
    double w_alpha, w_ac, w_beta, w_eps_h2, w_nu, nu_kin;  // w_ac = w_alpha c
    // DensityDiffusionAntuono: the renormalised density gradient of every fluid particle
    // (cell-sorted, ND per particle, written by k_antuono_gradient each kick) or nullptr
    // for DensityDiffusionMolteniColagrossi; eps(h^2), the zero-distance bound of the
    // gradient-correction pass (smoothing_kernel_grad, abstract_system.jl:142-155)
    double *ddgrad;
    double eps_hh;
    // SummationDensity: every kick sums the density of each fluid particle over its fluid and
    // wall neighbours (k_sum_density) instead of integrating it; v and dv then hold the
    // velocity rows only.  vrows: rows of v and dv per fluid particle (ND + 1, or ND)
    int summation, vrows;
};

// The sound-speed dependent constants of this kick: P's own, or the adaptive ones.  Kernels
// that evaluate the state equation or the pair terms start with P = with_sound(P).
__device__ __forceinline__ DevParams with_sound(DevParams P)
{
    if (P.dyn) {
        P.c = P.dyn[0];
        P.B = P.dyn[1];
        P.av_c = P.dyn[2];
        P.dd2 = P.dyn[3];
        P.w_ac = P.w_alpha * P.c;
        P.nu_kin = P.alpha * P.h * P.c / (2 * P.nd + 4);
    }
    return P;
}

// Cell-sorted fluid data.  rec[2s], rec[2s+1] = {x, y, z, rho}, {vx, vy, vz, p} (one 64-byte
// line per particle, read by the pair phase), vol[s] = {m/rho, m}; x/y/z SoA for the radius
// test.
struct FluidView {
    const int *start;
    const int *perm;
    const double *x, *y, *z;
    const double4 *rec;
    const double2 *vol;
    const float4 *xf = nullptr;  // the FP32 positions of the list build (Verlet lists only)
};

struct WallView {
    const int *start;
    const double *x, *y, *z;
    const double4 *pos;  // {x, y, z, m}
    const double2 *pv;   // {p_w, m/rho_w} of the current kick
    const double *vel;   // wall velocity for the viscosity (3 per particle), or nullptr
    const float4 *xf = nullptr;  // the FP32 positions of the list build (Verlet lists only)
};
// ------------------------------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------------------------------
// Julia's max(x, 0) for Float64: NaN propagates, -0.0 becomes +0.0.
__device__ __forceinline__ double julia_max0(double x) { return (x != x) ? x : (x > 0.0 ? x : 0.0); }

template <int ND>
__device__ __forceinline__ double dist2(double a, double b, double c)
{
#pragma clang fp contract(off)
    double s = a * a + b * b;
    if constexpr (ND == 3) s = s + c * c;
    return s;
}

// The rebuild flag: one atomic per wavefront that has a particle out of its skin, skipped
// once the flag reads as raised (when most particles have moved, thousands of atomics on
// the one address would serialise)
__device__ __forceinline__ void raise_rebuild_flag(int *flag, bool moved)
{
    if (__any(moved) && ((int)threadIdx.x & 63) == 0 &&
        __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
        atomicOr(flag, 1);
}

__device__ __forceinline__ int cell_1d(double x, double lo, double inv, int n)
{
    double f = floor((x - lo) * inv);
    f = fmin(fmax(f, 0.0), (double)(n - 1));  // also maps NaN into the grid
    return (int)f;
}

__device__ __forceinline__ double state_equation(const DevParams &P, double rho)
{
    double p = P.B * (pow(rho / P.rho0, P.gamma) - 1.0) + P.p_bg;
    return P.clip_fluid ? julia_max0(p) : p;
}

// WendlandC2 kernel value with the r < 2h guard (smoothing_kernels.jl:30-34, 436-441).
__device__ __forceinline__ double wendland_c2(const DevParams &P, double r)
{
    double q = r * P.h_inv;
    double w = 1.0 - q * 0.5;
    double w2 = w * w;
    return r < P.two_h ? P.kval_c * (w2 * w2) * (2.0 * q + 1.0) : 0.0;
}

// Distance from coordinate y to the interval of cell c along axis d, grown by hs; border
// cells reach to infinity outwards.
__device__ __forceinline__ double transverse(const DevParams &P, int d, int c, double y, double hs)
{
    double lower = (c == 0) ? -INFINITY : P.lo[d] + c * P.cs[d] - hs;
    double upper = (c == P.n[d] - 1) ? INFINITY : P.lo[d] + (c + 1) * P.cs[d] + hs;
    return fmax(0.0, fmax(lower - y, y - upper));
}

// Byte-offset loads: base pointer in SGPRs, 32-bit offset in a VGPR (global_load saddr form).
template <typename T>
__device__ __forceinline__ T ldo(const T *__restrict__ base, uint32_t byte_off)
{
    return *(const T *)((const char *)base + byte_off);
}

// Enumerates the candidate ranges (pencils) around one point, z-outer, y-inner.
template <int ND>
struct Pencils {
    double xa, ya, za, hs, dz2;
    int y0, y1, z1, iy, iz, rowz;

    __device__ __forceinline__ void init(const DevParams &P, double x, double y, double z,
                                         double hs_)
    {
        xa = x;
        ya = y;
        za = z;
        hs = hs_;
        dz2 = 0.0;
        rowz = 0;
        const double reach = P.R + hs;
        y0 = cell_1d(ya - reach, P.lo[1], P.inv_cs[1], P.n[1]);
        y1 = cell_1d(ya + reach, P.lo[1], P.inv_cs[1], P.n[1]);
        if constexpr (ND == 3) {
            iz = cell_1d(za - reach, P.lo[2], P.inv_cs[2], P.n[2]) - 1;
            z1 = cell_1d(za + reach, P.lo[2], P.inv_cs[2], P.n[2]);
        } else {
            iz = -1;
            z1 = 0;
        }
        iy = y1 + 1;  // the first call advances to the first z plane
    }

    // next non-empty candidate range [jb, je); false when exhausted
    __device__ __forceinline__ bool next(const DevParams &P, const int *__restrict__ start,
                                         int &jb, int &je)
    {
        while (true) {
            if (iy > y1) {
                if (++iz > z1) return false;
                iy = y0;
                if constexpr (ND == 3) {
                    const double dz = transverse(P, 2, iz, za, hs);
                    dz2 = dz * dz;
                }
                rowz = P.n[0] * P.n[1] * iz;
            }
            const int cy = iy++;
            const double dy = transverse(P, 1, cy, ya, hs);
            const double r2 = P.R2 - fma(dy, dy, dz2);
            if (r2 < 0.0) continue;
            // conservative half-width: FP32 sqrt, enlarged well past its rounding error
            const double rx = (double)(__builtin_amdgcn_sqrtf((float)r2) * 1.000001f) + hs;
            const int x0 = cell_1d(xa - rx, P.lo[0], P.inv_cs[0], P.n[0]);
            const int x1 = cell_1d(xa + rx, P.lo[0], P.inv_cs[0], P.n[0]);
            const uint32_t row = (uint32_t)(rowz + P.n[0] * cy);
            jb = ldo(start, (row + (uint32_t)x0) * 4u);
            je = ldo(start, (row + (uint32_t)x1 + 1u) * 4u);
            if (jb < je) return true;
        }
    }
};

// Map block b to a tile so that each XCD (blocks are dealt round-robin over 8 XCDs) works
// on one contiguous part of the cell-sorted particles and its L2 keeps the neighbours.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    constexpr int kXcd = 8;
    const int per = nb / kXcd, extra = nb % kXcd;
    const int x = b % kXcd, k = b / kXcd;
    return (x < extra) ? x * (per + 1) + k : extra * (per + 1) + (x - extra) * per + k;
}

__device__ __forceinline__ bool skip(const int *flag) { return flag && *flag == 0; }

// ------------------------------------------------------------------------------------------
// Adami pressure extrapolation (dummy_particles.jl:489-569): one thread per wall particle in
// wall-cell order over its fluid neighbours within 2h, then the inverse state equation.
// ------------------------------------------------------------------------------------------
// Contribution of fluid particle j (sorted index) to the sums of the wall particle at xw.
// With a wall viscosity the kernel-weighted fluid velocity is summed into vs
// (interpolate_fluid_velocity!, dummy_particles.jl:716-731).
template <int ND>
__device__ __forceinline__ void adami_term(const DevParams &P, const FluidView &F,
                                           const double4 &xw, int j, double &pressure,
                                           double &volume, double *vs)
{
    const double4 a = F.rec[2 * (size_t)j];
    const double p0 = xw.x - a.x, p1 = xw.y - a.y;
    const double p2 = (ND == 3) ? xw.z - a.z : 0.0;
    const double d2 = dist2<ND>(p0, p1, p2);
    if (d2 > P.R2) return;
    const double W = wendland_c2(P, sqrt(d2));
    const double rho_f = a.w;
    const double4 b = F.rec[2 * (size_t)j + 1];
    const double p_f = b.w;
    double hydro = P.g[0] * (rho_f * p0) + P.g[1] * (rho_f * p1);
    if constexpr (ND == 3) hydro += P.g[2] * (rho_f * p2);
    pressure += ((P.p_offset + p_f) + hydro) * W;
    volume += W;
    if (P.visc_w) {
        vs[0] += W * b.x;
        vs[1] += W * b.y;
        if constexpr (ND == 3) vs[2] += W * b.z;
    }
}

// p_w = P_w / V_w (if V_w > eps), optional clipping, the inverse state equation; with a wall
// viscosity the no-slip wall velocity 2 v_wall - vs / V_w with v_wall = 0 (if V_w > eps,
// else vs as summed; compute_wall_velocity!, dummy_particles.jl:738-756).
__device__ __forceinline__ void adami_finish(const DevParams &P, int w, double mass,
                                             double pressure, double volume,
                                             double *__restrict__ pw, double *__restrict__ rhow,
                                             double2 *__restrict__ wpv, const double *vs,
                                             double *__restrict__ vel)
{
    if (P.visc_w) {
#pragma unroll
        for (int d = 0; d < 3; d++)
            vel[3 * (size_t)w + d] = (volume > DBL_EPSILON) ? 0.0 - vs[d] / volume : vs[d];
    }
    if (volume > DBL_EPSILON) pressure /= volume;
    if (P.clip_wall) pressure = julia_max0(pressure);
    const double rho = P.rho0 * pow((pressure - P.p_bg) / P.B + 1.0, P.inv_gamma);
    pw[w] = pressure;
    rhow[w] = rho;
    wpv[w] = make_double2(pressure, mass / rho);
}

// ------------------------------------------------------------------------------------------
// Fused WCSPH interaction (wcsph/rhs.jl:5-127) for fluid-fluid and fluid-wall neighbours plus
// gravity (semidiscretization.jl:720-728), one lane per fluid particle in cell-sorted order.
//
// With grad W = kdr r_ab, kdr = -5 sigma h^-2 (1 - q/2)^3, V_b = m_b / rho_b:
//   pressure   -m_b (p_a + p_b)/(rho_a rho_b) grad W = -V_b (p_a + p_b)/rho_a grad W
//   viscosity  (m_b alpha c mu + m_b beta mu^2)/rho_mean grad W  when v_ab.r_ab < 0,
//              mu = h v_ab.r_ab / (r^2 + eps h^2)                     (fluid neighbours)
//   continuity rho_a/rho_b m_b v_ab.grad W = rho_a V_b kdr v_ab.r_ab
//   diffusion  delta h c V_b psi.grad W = 2 delta h c V_b (rho_a - rho_b) kdr
//              (psi.grad W = 2 (rho_a - rho_b) kdr r_ab.r_ab / r^2)   (fluid neighbours)
// Wall neighbours: v_b = 0, p_b and rho_b from the Adami pass, no viscosity (the wall model's
// viscosity is `nothing`) and no density diffusion.  Per particle the fluid sum, the wall sum
// and gravity are added in this order, as the reference adds them into dv.
//
// sqrt and the viscosity division use the hardware rsq / rcp refined by Goldschmidt and
// Newton-Raphson steps (the refinement LLVM emits for IEEE f64 sqrt and division, without
// the range scaling: arguments here are normal and positive); results agree with the
// correctly rounded operations to an ulp or so.
// ------------------------------------------------------------------------------------------
// sqrt(x) and 1/x for normal positive x from the hardware approximations v_rsq_f64 and
// v_rcp_f64 (relative error about 5e-8, measured) refined by one step of Goldschmidt's
// square-root iteration and of Newton's reciprocal iteration: each step squares the
// relative error, so the results are within about 5e-15 (measured 4.2e-15 and 2.2e-15 over
// 2^20 inputs from 2^-41 to 1), far inside the 1e-12 parity bar of the pair terms.
__device__ __forceinline__ double sqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double g = x * y, hh = 0.5 * y;
    return fma(g, fma(-hh, g, 0.5), g);
}

__device__ __forceinline__ double rcp_pos(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

struct FluidNbr {
    double4 a, b;  // {x, y, z, rho}, {vx, vy, vz, p}
    double2 v;     // {V, m}
    int j;         // cell-sorted index (read only by the DensityDiffusionAntuono term)
};

struct WallNbr {
    double4 a;  // {x, y, z, m}
    double2 pv; // {p, V}
    double v[3];  // velocity for the viscosity (only with a wall viscosity)
};

__device__ __forceinline__ FluidNbr load_fluid(const FluidView &F, int j)
{
    const uint32_t o = (uint32_t)j;
    return {ldo(F.rec, o * 64u), ldo(F.rec, o * 64u + 32u), ldo(F.vol, o * 16u), j};
}

__device__ __forceinline__ WallNbr load_wall(const WallView &Wv, int j)
{
    const uint32_t o = (uint32_t)j;
    WallNbr n{ldo(Wv.pos, o * 32u), ldo(Wv.pv, o * 16u), {0.0, 0.0, 0.0}};
    if (Wv.vel) {
#pragma unroll
        for (int d = 0; d < 3; d++) n.v[d] = Wv.vel[3 * (size_t)j + d];
    }
    return n;
}

// ga: the target's renormalised density gradient under DensityDiffusionAntuono, whose
// psi_ab = 2 (rho_a - rho_b - (grad rho_a + grad rho_b).r_ab / 2) r_ab / r^2
// (density_diffusion.jl:155-174) replaces the Molteni-Colagrossi 2 (rho_a - rho_b);
// nullptr for Molteni-Colagrossi (or no density diffusion).
// SUM: SummationDensity's pressure acceleration -m_b (p_a/rho_a^2 + p_b/rho_b^2) grad W
// (pressure_acceleration.jl:8-15) with p_a = the target's p/rho^2 passed in the p_a slot;
// the density rate is not used then (no continuity equation, fluid.jl:133-139).
template <int ND, bool SUM = false>
__device__ __forceinline__ void pair_fluid(const DevParams &P, const FluidNbr &n,
                                           const double *xa, const double *va, double rho_a,
                                           double p_a, double inv_rho_a, double *acc,
                                           double &drho, const double *ga = nullptr)
{
    const double p0 = xa[0] - n.a.x, p1 = xa[1] - n.a.y, p2 = (ND == 3) ? xa[2] - n.a.z : 0.0;
    const double d2 = dist2<ND>(p0, p1, p2);
    const double r = sqrt_pos(d2);
    const double w = fma(-0.5 * P.h_inv, r, 1.0);
    const double kdr = P.kgrad_c * (w * w * w);
    const double v0 = va[0] - n.b.x, v1 = va[1] - n.b.y, v2 = (ND == 3) ? va[2] - n.b.z : 0.0;
    double vr = v0 * p0 + v1 * p1;
    if constexpr (ND == 3) vr += v2 * p2;
    const double rho_b = n.a.w;
    double fac = SUM ? -n.v.y * (p_a + n.b.w / (rho_b * rho_b)) : -n.v.x * inv_rho_a * (p_a + n.b.w);
    if (P.use_visc && vr < 0.0) {
        // (m_b alpha c mu + m_b beta mu^2) / rho_mean, rho_mean = (rho_a + rho_b) / 2
        const double q = d2 + P.eps_h2;
        const double sum_rho = rho_a + rho_b;
        fac = fma(2.0 * P.av_c * n.v.y * vr, rcp_pos(q * sum_rho), fac);
        if (P.beta != 0.0) {
            const double mu = P.h * vr * rcp_pos(q);
            fac += 2.0 * n.v.y * P.beta * (mu * mu) * rcp_pos(sum_rho);
        }
    }
    fac *= kdr;
    acc[0] = fma(fac, p0, acc[0]);
    acc[1] = fma(fac, p1, acc[1]);
    if constexpr (ND == 3) acc[2] = fma(fac, p2, acc[2]);
    double drr = rho_a - rho_b;
    if (ga) {
        const double *gb = P.ddgrad + (size_t)n.j * ND;
        double gr = (ga[0] + gb[0]) * p0 + (ga[1] + gb[1]) * p1;
        if constexpr (ND == 3) gr += (ga[2] + gb[2]) * p2;
        drr = fma(-0.5, gr, drr);
    }
    drho = fma(n.v.x * kdr, fma(P.dd2, drr, rho_a * vr), drho);
}

// The wall viscosity (none without one) acts along r_ab for ArtificialViscosityMonaghan
// (viscosity.jl:89-132, approaching pairs only) and along v_ab for ViscosityAdami
// (viscosity.jl:219-280, with eta_a = nu_a rho_a from the fluid's kinematic viscosity and
// eta_b = nu rho_b, viscosity_model dispatch of wall_boundary/system.jl:286-289 and
// fluid.jl:109-112); v_b is the wall's no-slip velocity and rho_b = m_b / V_b.
// SUM: as in pair_fluid, with the wall's rho_b = m_b / V_b of the Adami pass.
template <int ND, bool SUM = false>
__device__ __forceinline__ void pair_wall(const DevParams &P, const WallNbr &n,
                                          const double *xa, const double *va, double rho_a,
                                          double p_a, double inv_rho_a, double *acc,
                                          double &drho, double m_a)
{
    const double p0 = xa[0] - n.a.x, p1 = xa[1] - n.a.y, p2 = (ND == 3) ? xa[2] - n.a.z : 0.0;
    const double d2 = dist2<ND>(p0, p1, p2);
    const double r = sqrt_pos(d2);
    const double w = fma(-0.5 * P.h_inv, r, 1.0);
    const double kdr = P.kgrad_c * (w * w * w);
    double vr = va[0] * p0 + va[1] * p1;
    if constexpr (ND == 3) vr += va[2] * p2;
    double fac;
    if constexpr (SUM) {
        const double rho_w = n.a.w / n.pv.y;
        fac = -n.a.w * (p_a + n.pv.x / (rho_w * rho_w)) * kdr;
    } else {
        fac = -n.pv.y * inv_rho_a * (p_a + n.pv.x) * kdr;
    }
    if (P.visc_w) {
        const double v0 = va[0] - n.v[0], v1 = va[1] - n.v[1];
        const double v2 = (ND == 3) ? va[2] - n.v[2] : 0.0;
        const double m_b = n.a.w, V_b = n.pv.y, rho_b = m_b / V_b;
        if (P.visc_w == 1) {
            double vrw = v0 * p0 + v1 * p1;
            if constexpr (ND == 3) vrw += v2 * p2;
            if (vrw < 0.0) {
                const double mu = P.h * vrw / (d2 + P.w_eps_h2);
                fac += (m_b * P.w_ac * mu + m_b * P.w_beta * (mu * mu)) /
                       (0.5 * (rho_a + rho_b)) * kdr;
            }
        } else {
            const double eta_a = P.nu_kin * rho_a, eta_b = P.w_nu * rho_b;
            const double V_a = m_a / rho_a;
            const double tmp = 2.0 * eta_a * eta_b / ((eta_a + eta_b) * (d2 + P.w_eps_h2) * m_a);
            const double visc = (V_a * V_a + V_b * V_b) * (kdr * d2) * tmp;
            acc[0] = fma(visc, v0, acc[0]);
            acc[1] = fma(visc, v1, acc[1]);
            if constexpr (ND == 3) acc[2] = fma(visc, v2, acc[2]);
        }
    }
    acc[0] = fma(fac, p0, acc[0]);
    acc[1] = fma(fac, p1, acc[1]);
    if constexpr (ND == 3) acc[2] = fma(fac, p2, acc[2]);
    drho = fma(n.pv.y * kdr, rho_a * vr, drho);
}

// The viscosity form of the fluid-fluid pairs as a compile-time constant of the branchless
// pair loops: none, ArtificialViscosityMonaghan with beta = 0, or with beta != 0.
constexpr int kViscNone = 0, kViscAlpha = 1, kViscAlphaBeta = 2;
inline int visc_mode(const DevParams &P)
{
    return P.use_visc ? (P.beta != 0.0 ? kViscAlphaBeta : kViscAlpha) : kViscNone;
}

// pair_fluid without branches, for candidates evaluated in groups: an out-of-range (or
// invalid) candidate gets kdr = 0 by a select, so its terms are exact zeros whatever the
// square root of its distance gives (NaN for d2 = 0), and the viscosity acts through
// min(v_ab.r_ab, 0), which is v_ab.r_ab for approaching pairs and zero otherwise (a zero
// term added to fac leaves it unchanged).  VISC is visc_mode(P).  For an in-range candidate
// every operation is the one of pair_fluid, so the sums agree bit for bit (up to the sign
// of a zero).  Every intermediate of an out-of-range candidate stays finite or is selected
// away: the far record (coordinates 1e100, zero volume) gives d2 about 1e200, whose
// reciprocals are normal numbers.
//
// ga, gb (DensityDiffusionAntuono only): the renormalised density gradients of the target and
// the candidate, whose term is pair_fluid's.
// SUM (SummationDensity): the pressure term -m_b (p_a/rho_a^2 + p_b/rho_b^2) grad W with the
// target's p/rho^2 in p_a and the candidate's in n.b.w (the brick kernel stages p/rho^2 in
// the pressure slot), and no density rate.
template <int ND, int VISC, bool ANT = false, bool SUM = false>
__device__ __forceinline__ void pair_fluid_masked(const DevParams &P, const FluidNbr &n,
                                                  bool valid, const double *xa,
                                                  const double *va, double rho_a, double p_a,
                                                  double inv_rho_a, double *acc, double &drho,
                                                  const double *ga = nullptr,
                                                  const double *gb = nullptr)
{
    const double p0 = xa[0] - n.a.x, p1 = xa[1] - n.a.y, p2 = (ND == 3) ? xa[2] - n.a.z : 0.0;
    const double d2 = dist2<ND>(p0, p1, p2);
    const bool in = valid & (d2 <= P.R2) & (d2 >= P.az2);
    const double r = sqrt_pos(d2);
    const double w = fma(-0.5 * P.h_inv, r, 1.0);
    const double kdr = in ? P.kgrad_c * (w * w * w) : 0.0;
    const double v0 = va[0] - n.b.x, v1 = va[1] - n.b.y, v2 = (ND == 3) ? va[2] - n.b.z : 0.0;
    double vr = v0 * p0 + v1 * p1;
    if constexpr (ND == 3) vr += v2 * p2;
    const double rho_b = n.a.w;
    double fac = SUM ? -n.v.y * (p_a + n.b.w) : -n.v.x * inv_rho_a * (p_a + n.b.w);
    if constexpr (VISC != kViscNone) {
        const double q = d2 + P.eps_h2;
        const double sum_rho = rho_a + rho_b;
        const double vn = fmin(vr, 0.0);
        fac = fma(2.0 * P.av_c * n.v.y * vn, rcp_pos(q * sum_rho), fac);
        if constexpr (VISC == kViscAlphaBeta) {
            const double mu = P.h * vn * rcp_pos(q);
            fac += 2.0 * n.v.y * P.beta * (mu * mu) * rcp_pos(sum_rho);
        }
    }
    fac *= kdr;
    acc[0] = fma(fac, p0, acc[0]);
    acc[1] = fma(fac, p1, acc[1]);
    if constexpr (ND == 3) acc[2] = fma(fac, p2, acc[2]);
    if constexpr (!SUM) {
        double drr = rho_a - rho_b;
        if constexpr (ANT) {
            double gr = (ga[0] + gb[0]) * p0 + (ga[1] + gb[1]) * p1;
            if constexpr (ND == 3) gr += (ga[2] + gb[2]) * p2;
            drr = fma(-0.5, gr, drr);
        }
        drho = fma(n.v.x * kdr, fma(P.dd2, drr, rho_a * vr), drho);
    }
}

// Two-phase sweep over the neighbours of one target: phase 1 runs the exact radius test over
// the candidates of the pencils and appends hits to a per-lane list in LDS, phase 2 evaluates
// the pair terms on the dense list; the phases alternate until every lane has exhausted its
// pencils.  lst points at
// this lane's column of the block's LDS lists (stride kBlock).  Phase 1 tests four
// candidates per iteration without branches: the arrays cx/cy/cz carry kPad NaN entries
// past their end, so reads beyond a range stay in bounds and never pass the test, and every
// candidate is written to the list while the count only advances on a hit.
template <int ND, bool WALL, int CAP = kCap, int STRIDE = kBlock, bool SUM = false>
__device__ __forceinline__ void neighbour_sweep(const DevParams &P, const int *__restrict__ start,
                                                const double *__restrict__ cx,
                                                const double *__restrict__ cy,
                                                const double *__restrict__ cz, double hs,
                                                bool active, const FluidView &F,
                                                const WallView &Wv, int *lst, const double *xa,
                                                const double *va, double rho_a, double p_a,
                                                double inv_rho_a, double *acc, double &drho,
                                                double m_a = 0.0, const double *ga = nullptr)
{
    Pencils<ND> it;
    it.init(P, xa[0], xa[1], (ND == 3) ? xa[2] : 0.0, hs);
    bool done = !active;
    int j = 0, je = 0;
    while (true) {
        int cnt = 0;
        // phase 1: exact radius test
        while (!done && cnt <= CAP - 4) {
            if (j >= je) {
                if (!it.next(P, start, j, je)) done = true;
                continue;
            }
            const uint32_t o = (uint32_t)j * 8u;
            double d2[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const double p0 = xa[0] - ldo(cx, o + 8u * k);
                const double p1 = xa[1] - ldo(cy, o + 8u * k);
                const double p2 = (ND == 3) ? xa[2] - ldo(cz, o + 8u * k) : 0.0;
                d2[k] = dist2<ND>(p0, p1, p2);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const bool hit = (j + k < je) & (d2[k] <= P.R2) & (d2[k] >= P.az2);
                lst[cnt * STRIDE] = j + k;
                cnt += hit ? 1 : 0;
            }
            j += 4;
        }
        // phase 2: pair terms on the dense list, loads one pair ahead
        if (cnt > 0) {
            if constexpr (WALL) {
                WallNbr cur = load_wall(Wv, lst[0]);
                for (int k = 0; k < cnt; k++) {
                    const WallNbr nxt = load_wall(Wv, lst[min(k + 1, cnt - 1) * STRIDE]);
                    pair_wall<ND, SUM>(P, cur, xa, va, rho_a, p_a, inv_rho_a, acc, drho, m_a);
                    cur = nxt;
                }
            } else {
                FluidNbr cur = load_fluid(F, lst[0]);
                for (int k = 0; k < cnt; k++) {
                    const FluidNbr nxt = load_fluid(F, lst[min(k + 1, cnt - 1) * STRIDE]);
                    pair_fluid<ND, SUM>(P, cur, xa, va, rho_a, p_a, inv_rho_a, acc, drho, ga);
                    cur = nxt;
                }
            }
        }
        if (!__any(!done)) break;
    }
}

// largest r in [0, n) with pre[r] <= i (pre ascending, pre[0] = 0)
__device__ __forceinline__ int bsearch_le(const int *pre, int n, int i)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// bounds of cell c along axis d grown by hs; border cells reach to infinity outwards
__device__ __forceinline__ double cell_lo(const DevParams &P, int d, int c, double hs)
{
    return (c == 0) ? -INFINITY : P.lo[d] + c * P.cs[d] - hs;
}
__device__ __forceinline__ double cell_hi(const DevParams &P, int d, int c, double hs)
{
    return (c == P.n[d] - 1) ? INFINITY : P.lo[d] + (c + 1) * P.cs[d] + hs;
}


inline int nblocks(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }
inline int gate_blocks(int64_t n) { return std::max(1, std::min(nblocks(n), kGateBlocks)); }

}  // namespace tpsph

namespace tpsph {
// ------------------------------------------------------------------------------------------
// Counting-sort building blocks shared by the cell lists of the library's translation units
// (each unit holds its own copy: they live in an unnamed namespace).  Every kernel is gated
// by a rebuild flag (nullptr = always runs).
// ------------------------------------------------------------------------------------------
namespace {

// Exclusive scan over the block: an inclusive scan within each wavefront by shuffles, then
// the wavefront totals through LDS (s holds at least kBlock / 64 ints)
__device__ __forceinline__ int block_exclusive_scan(int value, int *s, int &total)
{
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    int x = value;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(x, off);
        if (lane >= off) x += t;
    }
    if (lane == 63) s[wave] = x;
    __syncthreads();
    int below = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) {
        const int t = s[w];
        below += w < wave ? t : 0;
        total += t;
    }
    __syncthreads();  // s is free again for the next call
    return below + x - value;
}

// kScanItems consecutive ints from p + base (16-byte loads where p is 16-byte aligned and
// the items lie within n), zeros past n
__device__ __forceinline__ void scan_load(const int *p, int base, int n, bool vec, int *v)
{
    static_assert(kScanItems == 8, "two int4 per thread");
    if (vec && base + kScanItems <= n) {
        const int4 a = *reinterpret_cast<const int4 *>(p + base);
        const int4 b = *reinterpret_cast<const int4 *>(p + base + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; k++) v[k] = (base + k < n) ? p[base + k] : 0;
    }
}

__device__ __forceinline__ void scan_store(int *p, int base, int n, bool vec, const int *r)
{
    if (vec && base + kScanItems <= n) {
        *reinterpret_cast<int4 *>(p + base) = make_int4(r[0], r[1], r[2], r[3]);
        *reinterpret_cast<int4 *>(p + base + 4) = make_int4(r[4], r[5], r[6], r[7]);
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; k++)
            if (base + k < n) p[base + k] = r[k];
    }
}

__device__ __forceinline__ bool aligned16(const void *p)
{
    return ((uintptr_t)p & 15u) == 0;
}

[[maybe_unused]] __global__ void __launch_bounds__(kBlock)
    k_scan_tiles(const int *flag, const int *in, int *out, int *tile_sums, int n)
{
    if (skip(flag)) return;
    __shared__ int s[kBlock];
    const bool vec = aligned16(in) && aligned16(out);
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    scan_load(in, base, n, vec, v);
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) sum += v[k];
    int total;
    int run = block_exclusive_scan(sum, s, total);
    int r[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        r[k] = run;
        run += v[k];
    }
    scan_store(out, base, n, vec, r);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

[[maybe_unused]] __global__ void __launch_bounds__(kBlock) k_scan_single(const int *flag, int *data, int n)
{
    if (skip(flag)) return;
    __shared__ int s[kBlock];
    int carry = 0;
    for (int chunk = 0; chunk < n; chunk += kScanTile) {
        int base = chunk + threadIdx.x * kScanItems;
        int v[kScanItems];
        int sum = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            v[k] = (base + k < n) ? data[base + k] : 0;
            sum += v[k];
        }
        int total;
        int run = carry + block_exclusive_scan(sum, s, total);
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            if (base + k < n) data[base + k] = run;
            run += v[k];
        }
        carry += total;
    }
}

[[maybe_unused]] __global__ void k_scan_add(const int *flag, int *out, const int *tile_sums, int n)
{
    if (skip(flag)) return;
    int i = blockIdx.x * kScanTile + threadIdx.x;
    int add = tile_sums[blockIdx.x];
    for (int k = 0; k < kScanItems; k++, i += kBlock)
        if (i < n) out[i] += add;
}

// Also clears the histogram again: the counts are zero between builds, so the next build
// needs no separate clearing pass.
[[maybe_unused]] __global__ void k_scatter(const int *flag, int n, const int *__restrict__ cell,
                          const int *__restrict__ rank, const int *__restrict__ start,
                          int *__restrict__ perm, int *__restrict__ count)
{
    if (skip(flag)) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c = cell[i];
        perm[start[c] + rank[i]] = i;
        count[c] = 0;
    }
}

// Order every cell's members by particle index: the result is the stable counting sort,
// independent of the order in which the histogram atomics were served.
[[maybe_unused]] __global__ void k_sort_cells(const int *flag, int ncells, const int *__restrict__ start,
                             int *__restrict__ perm)
{
    if (skip(flag)) return;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
        int b = start[c], e = start[c + 1];
        for (int k = b + 1; k < e; k++) {
            int key = perm[k];
            int j = k - 1;
            while (j >= b && perm[j] > key) {
                perm[j + 1] = perm[j];
                j--;
            }
            perm[j + 1] = key;
        }
    }
}

[[maybe_unused]] __global__ void k_set_i32(int *p, int value) { *p = value; }

}  // namespace

}  // namespace tpsph

// ==========================================================================================
// handle
// ==========================================================================================
struct tpsph_handle {
    tpsph_params params{};
    tpsph::DevParams P{};
    int nd = 0;
    int nf = 0, nw = 0;
    hipStream_t stream = nullptr;
    std::string last_error;

    int *flag = nullptr;  // rebuild of the fluid cell list pending
    // fluid cell list + cell-sorted state
    int *f_cell = nullptr, *f_rank = nullptr, *f_count = nullptr, *f_start = nullptr,
        *f_perm = nullptr;
    int *scan_tiles = nullptr;
    double *f_mass = nullptr;  // caller order
    double *xs = nullptr, *ys = nullptr, *zs = nullptr, *snap = nullptr;
    float4 *xf = nullptr;  // Verlet lists: FP32 positions at the last rebuild (+ kPad NaN)
    float4 *wxf = nullptr;  // Verlet lists: FP32 positions of the sorted walls (+ kPad NaN)
    int *iperm = nullptr;  // sorted index of every caller particle (last rebuild)
    double4 *rec = nullptr;
    double2 *vol = nullptr;
    // walls (static, cell-sorted) + per kick
    int *w_start = nullptr, *w_perm = nullptr;
    double *wx = nullptr, *wy = nullptr, *wz = nullptr, *pw = nullptr, *rhow = nullptr;
    double4 *wpos = nullptr;
    double2 *wpv = nullptr;
    unsigned char *wall_near = nullptr;  // per cell: a wall particle lies within reach
    int search = TPSPH_SEARCH_CELLS;  // resolved pair search
    // Verlet lists (slot-major, ls slots per row) and their counts
    int *nbr = nullptr, *cnt = nullptr, *wnbr = nullptr, *wcnt = nullptr;
    int ls = 0, cap_f = 0, cap_w = 0;
    // processing order of the list kernels (bricks of cells) and the bricks in LDS
    int *ord = nullptr, *bcount = nullptr, *bstart = nullptr;
    int nbricks = 0;
    bool bricks = false;
    int *brow_gb = nullptr, *brow_lds = nullptr, *bntot = nullptr, *nel = nullptr, *bwork = nullptr;
    int *bfall = nullptr;  // mixed precision: the bricks left to the FP64 pass
    unsigned short *nbr16 = nullptr;
    double *dyn = nullptr;       // adaptive sound speed: {c, B, av_c, dd2}
    double *wvel = nullptr;      // no-slip wall velocities (3 per wall particle)
    double *ddgrad = nullptr;    // DensityDiffusionAntuono: renormalised density gradients
    double *vmax_part = nullptr; // per-block maxima of v.v (kGateBlocks)
    double *sum_rho = nullptr;   // SummationDensity: summed densities (cell-sorted)
    // Verlet lists of the wall particles over the fluid (Adami pass), active wall particles
    int *anbr = nullptr, *acnt = nullptr, *wactive = nullptr;
    int lw = 0, cap_a = 0;
    // counters rebuilt with the lists: [0] non-empty bricks, [1] active wall particles,
    // [2] fluid particles with wall candidates (wtargets)
    int *counters = nullptr, *wtargets = nullptr;
    int n_cu = 256;  // compute units of the device (grid of the brick kernel)
    // host-mode staging and solver scratch
    double *st_v = nullptr, *st_u = nullptr, *st_dv = nullptr, *st_du = nullptr;
    double *sol_dv = nullptr, *sol_tv = nullptr, *sol_tu = nullptr;
    bool have_kick = false;
    bool mixed = false;  // TPSPH_PRECISION_MIXED: FP32 pair terms in the brick kernel
    double *integ = nullptr;  // tpsph_integrals: block partials and the result
    int *nrebuild = nullptr;  // cell-list rebuilds since create (tpsph_rebuild_count)
    // reverse mode (tpsph_adjoint.hip): cell-sorted adjoints of the rates {dv, drho}, the
    // Adami coefficients of the walls, per-particle parameter partials (alpha, delta, g in the
    // caller's order; g per wall particle), reduction partials, the reverse sweep's adjoints
    // of (v, u) and of the stage registers, the checkpoints and one segment's stage states
    double4 *adj_kb = nullptr;
    double2 *adj_coef = nullptr;
    double *adj_th = nullptr, *adj_gw = nullptr, *adj_red = nullptr;
    double *adj_vb = nullptr, *adj_ub = nullptr, *adj_tvb = nullptr, *adj_tub = nullptr;
    double *adj_ck = nullptr, *adj_st = nullptr;
    size_t adj_ck_cap = 0, adj_st_cap = 0;  // doubles allocated
    int64_t adj_nsteps = 0, adj_every = 0;
    double adj_dt = 0.0;
    bool adj_ready = false;  // tpsph_adjoint_forward has stored a trajectory

    bool timing = false;
    struct Rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> event_pool;
    size_t events_used = 0;
    std::vector<Rec> records;

    tpsph::FluidView fluid_view() const { return {f_start, f_perm, xs, ys, zs, rec, vol, xf}; }
    tpsph::WallView wall_view() const { return {w_start, wx, wy, wz, wpos, wpv, wvel, wxf}; }
};


namespace tpsph {

inline int fail(tpsph_handle *h, int code, const std::string &msg)
{
    if (h)
        h->last_error = msg;
    else
        g_create_error = msg;
    return code;
}

#define HIP_TRY(h, expr)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((h), TPSPH_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

#define TRY(expr)                        \
    do {                                 \
        int rc_ = (expr);                \
        if (rc_ != TPSPH_OK) return rc_; \
    } while (0)

template <typename T>
inline int dalloc(tpsph_handle *h, T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    HIP_TRY(h, hipMalloc((void **)p, count * sizeof(T)));
    return TPSPH_OK;
}

// all bits set: a quiet NaN in every double
inline int hip_fill(tpsph_handle *h, double *p, size_t count)
{
    HIP_TRY(h, hipMemset(p, 0xff, count * sizeof(double)));
    return TPSPH_OK;
}

inline int launch_check(tpsph_handle *h, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(h, TPSPH_ERR_HIP, std::string("launch of ") + what + ": " + hipGetErrorString(e));
    return TPSPH_OK;
}

struct TimedScope {
    tpsph_handle *h;
    hipEvent_t b = nullptr;
    TimedScope(tpsph_handle *h_, int kind) : h(h_)
    {
        if (!h->timing) return;
        if (h->events_used + 2 > h->event_pool.size()) {
            for (int k = 0; k < 64; k++) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return;
                h->event_pool.push_back(e);
            }
        }
        hipEvent_t a = h->event_pool[h->events_used++];
        b = h->event_pool[h->events_used++];
        (void)hipEventRecord(a, h->stream);
        h->records.push_back({kind, a, b});
    }
    ~TimedScope()
    {
        if (b) (void)hipEventRecord(b, h->stream);
    }
};

// Argument checks of tpsph_create (the reference's ArgumentError checks); failures go to the
// create error (tpsph_last_error(NULL)).
int validate(const tpsph_params *p, int64_t nf, int64_t nw, const double *fm, const double *wx,
             const double *wm);

// The device parameters of params for nf fluid and nw wall particles on the grid of the box
// [lo, hi]: kernel and state-equation constants, cells, skin radii and bricks.
int derive_params(const tpsph_params *params, int nf, int nw, const double *lo,
                  const double *hi, DevParams &P, int &nbricks);

// the fluid-fluid pass of TPSPH_PRECISION_MIXED (tpsph_mixed.hip), in place of k_fluid_brick
int launch_fluid_brick_mixed(tpsph_handle *h, double *dv);

// Exclusive scan of n ints from in to out on the handle's stream (tile sums in h->scan_tiles).
inline int scan(tpsph_handle *h, const int *flag, const int *in, int *out, int n)
{
    int tiles = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL((k_scan_tiles), dim3(tiles), dim3(kBlock), 0, h->stream, flag, in, out, h->scan_tiles, n);
    hipLaunchKernelGGL((k_scan_single), dim3(1), dim3(kBlock), 0, h->stream, flag, h->scan_tiles, tiles);
    hipLaunchKernelGGL((k_scan_add), dim3(tiles), dim3(kBlock), 0, h->stream, flag, out, h->scan_tiles, n);
    return launch_check(h, "scan");
}

// gathered: the cell-sorted records already hold (v, u) (written by k_ck_stage_gather)
int kick_dispatch(tpsph_handle *h, double *dv, const double *v, const double *u,
                  bool gathered = false);

// A kick without its fluid pass: the records of (v, u) with the fluid pressures, the rebuild
// of the cell order and the lists (when a particle left its skin), the Adami pass.
// keep_order: the caller knows that no particle of (v, u) has left its skin since the last
// rebuild, so the rebuild launches are left out and the flag is lowered first.
int prepare_dispatch(tpsph_handle *h, const double *v, const double *u, bool keep_order = false);

// CarpenterKennedy2N54 (williamson_condition = false) coefficients A (A[0] unused) and B
extern const double kCkA[5], kCkB[5];

// nsteps CK2N54 steps of the device state (v, u) in place.  With store_v / store_u the state
// every stage's kick sees is copied there first, stage k = 5 step + i at offset k v_len / u_len,
// and the run stops after storing the last one ((v, u) is then left at that stage state).
int run_ck2n54(tpsph_handle *h, double *v, double *u, double dt, int64_t nsteps,
               double *store_v, double *store_u);

inline bool is_host(const tpsph_handle *h) { return h->params.memory == TPSPH_MEM_HOST; }
inline size_t v_len(const tpsph_handle *h) { return (size_t)h->nf * h->P.vrows; }
inline size_t u_len(const tpsph_handle *h) { return (size_t)h->nf * h->nd; }

inline int ensure_staging(tpsph_handle *h)
{
    if (h->st_v) return TPSPH_OK;
    TRY(dalloc(h, &h->st_v, v_len(h)));
    TRY(dalloc(h, &h->st_u, u_len(h)));
    TRY(dalloc(h, &h->st_dv, v_len(h)));
    TRY(dalloc(h, &h->st_du, u_len(h)));
    return TPSPH_OK;
}

}  // namespace tpsph
