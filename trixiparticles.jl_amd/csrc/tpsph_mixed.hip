// tpsph_mixed.hip -- the fluid-fluid pass of the mixed-precision right-hand side
// (tpsph_params.precision == TPSPH_PRECISION_MIXED).
//
// The same bricks, unions and 16-bit LDS lists as k_fluid_brick (tpsph.hip), with the pair
// arithmetic in FP32:
//   - positions are staged relative to the centre of the brick (FP64 difference, rounded
//     once to FP32), so the pair differences carry the FP32 error of coordinates of a few R
//     instead of the domain's extent;
//   - densities are staged as rho - rho0 (FP64 difference, rounded once), so the
//     Molteni-Colagrossi term rho_a - rho_b keeps its relative precision;
//   - the density rate is summed with Kahan's compensated summation (one compensated add per
//     group of four list slots), and the two lanes of a target combine their (sum,
//     compensation) pairs in FP64;
//   - FP64 stays where it is: the state, the state equation (pressures come from the FP64
//     density of the stage kernel), the wall sums (k_fluid_walls), the Adami pass, the
//     stage update and the written dv / drho.
// Bricks that do not stage (a union beyond the LDS records, an overflowed list) run the FP64
// terms from global memory in k_fluid_brick<ND, true>.
//
// LDS: 36 bytes per record (two float4 and one float) instead of 72, so two blocks of 1024
// threads share a CU and one block's staging overlaps the other's pair terms.
#include "tpsph_internal.h"

using namespace tpsph;

namespace {

constexpr float kFar32 = 1e18f;  // coordinate of the far record: d^2 = 1e36 stays finite

template <int ND>
struct BrickSmem32 {
    float4 pa[kBrickCap];  // {x, y, z, rho - rho0} relative to the brick centre
    float4 pb[kBrickCap];  // {vx, vy, vz, p}
    float vol[kBrickCap];  // V = m / rho
    int row_gb[kBrickMaxRows], row_lds[kBrickMaxRows + 1];
    int item;
    double org[3];         // the brick centre
};

// What the FP32 pass needs, packed small so that it stays in scalar registers: the pair
// constants in FP32 and the brick geometry in FP64.  Under StateEquationAdaptiveCole the
// viscosity and diffusion constants of this kick are read from dyn ({c, B, alpha c h,
// 2 delta h c}, written by k_sound_speed).
struct Pair32 {
    float R2, az2, h_inv_half, kgrad_c, eps_h2, two_av_c, dd2, two_rho0, rho0;
    int use_visc, maxr, nw;
    int nb0, nb1, bc[3];
    double lo[3], cs[3], g[3], rho0_64;
    const double *dyn;
};

Pair32 pair32(const DevParams &P)
{
    Pair32 q;
    q.R2 = (float)P.R2;
    q.az2 = (float)P.az2;
    q.h_inv_half = (float)(0.5 * P.h_inv);
    q.kgrad_c = (float)P.kgrad_c;
    q.eps_h2 = (float)P.eps_h2;
    q.two_av_c = (float)(2.0 * P.av_c);
    q.dd2 = (float)P.dd2;
    q.two_rho0 = (float)(2.0 * P.rho0);
    q.rho0 = (float)P.rho0;
    q.use_visc = P.use_visc;
    q.maxr = P.maxr;
    q.nw = P.nw;
    q.nb0 = P.nb[0];
    q.nb1 = P.nb[1];
    for (int d = 0; d < 3; d++) {
        q.bc[d] = P.bc[d];
        q.lo[d] = P.lo[d];
        q.cs[d] = P.cs[d];
        q.g[d] = P.g[d];
    }
    q.rho0_64 = P.rho0;
    q.dyn = P.dyn;
    return q;
}

// Kahan's compensated sum: s + c carries the running total, c the negated rounding error.
struct Kahan {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x)
    {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// One candidate in FP32 (pair_fluid_masked's terms).  Out-of-range candidates (and the far
// record) get kdr = 0: their terms are exact zeros.  The range test selects no branch: d2 is
// clamped into [az2, R2] for the kernel factor, which the mask then zeroes.  The density-rate
// term is added to drho_part, the plain FP32 sum of one list group (at most four terms) that
// the caller adds to the compensated total.
template <int ND>
__device__ __forceinline__ void pair32_term(const Pair32 &q, const float4 &a, const float4 &b,
                                            float V_b, const float *xa, const float *va,
                                            float drho_a, float rho_a, float p_a,
                                            float inv_rho_a, float *acc, float &drho_part)
{
    const float p0 = xa[0] - a.x, p1 = xa[1] - a.y, p2 = (ND == 3) ? xa[2] - a.z : 0.0f;
    float d2 = p0 * p0 + p1 * p1;
    if constexpr (ND == 3) d2 = d2 + p2 * p2;
    const bool in = (d2 <= q.R2) & (d2 >= q.az2);
    const float d2s = __builtin_fminf(__builtin_fmaxf(d2, q.az2), q.R2);
    const float r = d2s * __builtin_amdgcn_rsqf(d2s);
    const float w = 1.0f - q.h_inv_half * r;
    const float kdr = q.kgrad_c * (w * w * w) * (in ? 1.0f : 0.0f);
    const float v0 = va[0] - b.x, v1 = va[1] - b.y, v2 = (ND == 3) ? va[2] - b.z : 0.0f;
    float vr = v0 * p0 + v1 * p1;
    if constexpr (ND == 3) vr = vr + v2 * p2;
    float fac = -V_b * inv_rho_a * (p_a + b.w);
    if (q.use_visc) {
        // (m_b alpha c mu) / rho_mean with rho_mean = (rho_a + rho_b) / 2, m_b = V_b rho_b
        const float sum_rho = q.two_rho0 + (drho_a + a.w);
        const float m_b = V_b * (q.rho0 + a.w);
        const float fv = fac + q.two_av_c * m_b * vr * __builtin_amdgcn_rcpf((d2s + q.eps_h2) * sum_rho);
        fac = vr < 0.0f ? fv : fac;
    }
    fac = fac * kdr;
    acc[0] += fac * p0;
    acc[1] += fac * p1;
    if constexpr (ND == 3) acc[2] += fac * p2;
    drho_part += V_b * kdr * (q.dd2 * (drho_a - a.w) + rho_a * vr);
}

// Fluid-fluid pass over the staged bricks, one block of kBrickThreads per brick taken from
// the work counter work[0] (two blocks per CU).  Bricks that do not stage are listed in bfall
// (work[1] of them) for k_fluid_brick<ND, true> (FP64, global memory), launched after this
// kernel.
template <int ND>
__global__ void __launch_bounds__(kBrickThreads, 8)
    k_fluid_brick32(int *flag, Pair32 q, FluidView F, const int *__restrict__ ord,
                    const int *__restrict__ bstart, const int *__restrict__ nel,
                    const int *__restrict__ nne, int *__restrict__ work,
                    const int *__restrict__ brow_gb, const int *__restrict__ brow_lds,
                    const int *__restrict__ bntot, const unsigned short *__restrict__ nbr16,
                    const int *__restrict__ cnt, int ls, const int *__restrict__ wcnt,
                    double *__restrict__ dv, int *__restrict__ bfall)
{
    if (q.dyn) {  // the sound speed of this kick
        q.two_av_c = (float)(2.0 * q.dyn[2]);
        q.dd2 = (float)q.dyn[3];
    }
    __shared__ BrickSmem32<ND> S;
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && tid == 0) *flag = 0;  // the rebuild (if any) is done
    const int ne = *nne;
    while (true) {
        if (tid == 0) S.item = atomicAdd(work, 1);
        __syncthreads();
        const int item = S.item;
        if (item >= ne) break;
        const int b = nel[item];
        const int ntot = bntot[b];
        if (ntot > kBrickCap - 1) {  // left to the FP64 pass (uniform per block)
            if (tid == 0) bfall[atomicAdd(&work[1], 1)] = b;
        } else {
            const int t0 = bstart[b], t1 = bstart[b + 1];
            for (int r = tid; r <= q.maxr; r += kBrickThreads) {
                if (r < q.maxr) S.row_gb[r] = brow_gb[(size_t)b * q.maxr + r];
                S.row_lds[r] = brow_lds[(size_t)b * (q.maxr + 1) + r];
            }
            if (tid == 0) {
                const int bx = b % q.nb0, by = (b / q.nb0) % q.nb1, bz = b / (q.nb0 * q.nb1);
                const int bc[3] = {bx, by, bz};
#pragma unroll
                for (int d = 0; d < 3; d++)
                    S.org[d] = d < ND ? q.lo[d] + (bc[d] * q.bc[d] + 0.5 * q.bc[d]) * q.cs[d] : 0.0;
                constexpr int f = kBrickCap - 1;  // the far record
                S.pa[f] = make_float4(kFar32, kFar32, kFar32, 0.0f);
                S.pb[f] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                S.vol[f] = 0.0f;
            }
            __syncthreads();
            const double o0 = S.org[0], o1 = S.org[1], o2 = S.org[2];
            for (int i = tid; i < ntot; i += kBrickThreads) {
                const int r = bsearch_le(S.row_lds, q.maxr + 1, i);
                const int g = S.row_gb[r] + (i - S.row_lds[r]);
                const double4 a = F.rec[2 * (size_t)g];
                const double4 c = F.rec[2 * (size_t)g + 1];
                S.pa[i] = make_float4((float)(a.x - o0), (float)(a.y - o1),
                                      ND == 3 ? (float)(a.z - o2) : 0.0f, (float)(a.w - q.rho0_64));
                S.pb[i] = make_float4((float)c.x, (float)c.y, ND == 3 ? (float)c.z : 0.0f, (float)c.w);
                S.vol[i] = (float)F.vol[g].x;
            }
            __syncthreads();
            // two lanes per target take alternate groups of four list slots
            const int half = tid & 1;
            for (int base = t0; base < t1; base += kBrickThreads / 2) {
                const int t = base + (tid >> 1);
                const bool active = t < t1;
                const int col = active ? t : t0;
                const int s = ord[col];
                double4 a64 = make_double4(0.0, 0.0, 0.0, 1.0), c64 = make_double4(0.0, 0.0, 0.0, 0.0);
                int n = 0, nw = 0;
                if (active) {
                    a64 = F.rec[2 * (size_t)s];
                    c64 = F.rec[2 * (size_t)s + 1];
                    n = cnt[col];
                    if (q.nw > 0) nw = wcnt[col];
                }
                const float xa[3] = {(float)(a64.x - o0), (float)(a64.y - o1),
                                     ND == 3 ? (float)(a64.z - o2) : 0.0f};
                const float va[3] = {(float)c64.x, (float)c64.y, ND == 3 ? (float)c64.z : 0.0f};
                const float drho_a = (float)(a64.w - q.rho0_64);
                const float rho_a = (float)a64.w, p_a = (float)c64.w;
                const float inv_rho_a = (float)(1.0 / a64.w);
                float acc[3] = {0.0f, 0.0f, 0.0f};
                Kahan drho;
                const ushort4 *l4 = reinterpret_cast<const ushort4 *>(nbr16);
                const int ng = (max(n, 0) + 3) >> 2;
                // list rows clamped into the list rather than selected (a select made the
                // compiler issue flat loads, whose waits hold up the LDS reads)
                ushort4 cur4 = l4[(size_t)min(half, max(ng - 1, 0)) * ls + col];
                for (int g = half; g < ng; g += 2) {
                    const ushort4 nxt4 = l4[(size_t)min(g + 2, ng - 1) * ls + col];
                    // one candidate at a time: eight waves per SIMD hide the LDS latency
                    const unsigned short e4[4] = {cur4.x, cur4.y, cur4.z, cur4.w};
                    float part = 0.0f;
#pragma unroll 1
                    for (int k = 0; k < 4; k++) {
                        const int e = e4[k];
                        pair32_term<ND>(q, S.pa[e], S.pb[e], S.vol[e], xa, va, drho_a, rho_a, p_a,
                                        inv_rho_a, acc, part);
                    }
                    drho.add(part);
                    cur4 = nxt4;
                }
                // the partner's sums in FP64, the compensation folded in
                double acc64[3] = {(double)acc[0], (double)acc[1], (double)acc[2]};
                double drho64 = (double)drho.s - (double)drho.c;
#pragma unroll
                for (int d = 0; d < 3; d++) acc64[d] += __shfl_xor(acc64[d], 1);
                drho64 += __shfl_xor(drho64, 1);
                if (active && half == 0) {
                    const int i = F.perm[s];
                    const bool walls = nw != 0;  // k_fluid_walls adds the wall sum, then gravity
#pragma unroll
                    for (int d = 0; d < ND; d++)
                        dv[(size_t)i * (ND + 1) + d] = walls ? acc64[d] : acc64[d] + q.g[d];
                    dv[(size_t)i * (ND + 1) + ND] = drho64;
                }
            }
        }
        __syncthreads();  // the LDS is staged again for the next brick
    }
}

}  // namespace

int tpsph::launch_fluid_brick_mixed(tpsph_handle *h, double *dv)
{
    const int grid = std::max(1, std::min(h->nbricks, 2 * h->n_cu));
    if (h->nd == 2)
        hipLaunchKernelGGL((k_fluid_brick32<2>), dim3(grid), dim3(kBrickThreads), 0, h->stream,
                           h->flag, pair32(h->P), h->fluid_view(), h->ord, h->bstart, h->nel,
                           h->counters, h->bwork, h->brow_gb, h->brow_lds, h->bntot, h->nbr16,
                           h->cnt, h->ls, h->wcnt, dv, h->bfall);
    else
        hipLaunchKernelGGL((k_fluid_brick32<3>), dim3(grid), dim3(kBrickThreads), 0, h->stream,
                           h->flag, pair32(h->P), h->fluid_view(), h->ord, h->bstart, h->nel,
                           h->counters, h->bwork, h->brow_gb, h->brow_lds, h->bntot, h->nbr16,
                           h->cnt, h->ls, h->wcnt, dv, h->bfall);
    return launch_check(h, "k_fluid_brick32");
}
