// tpsph.hip -- MI355X (gfx950) WCSPH right-hand side behind the C ABI of include/tpsph.h.
//
// One kick! (semidiscretization.jl:589-612) runs, on the handle's stream:
//   1. k_gather<0>   gather the fluid state into cell-sorted order with the current
//                    permutation, fluid state equation, and flag a rebuild if any particle
//                    moved by more than half the skin since the last cell-list build;
//   2. rebuild, only when flagged (every kernel reads the flag and returns otherwise; the
//      fluid kernel of step 4 lowers the flag again):
//      k_cells + k_scan_* + k_scatter + k_sort_cells (a stable counting sort on the linear
//      cell index), k_gather<1> (gather again, snapshot of the build positions) and, with
//      Verlet lists, k_lists (fluid and wall candidates within R + skin);
//   3. k_wall        Adami pressure extrapolation + inverse state equation per wall particle;
//   4. the fused fluid-fluid + fluid-wall interaction + gravity, results scattered back to
//      the caller's particle order: k_fluid_brick + k_fluid_walls (Verlet lists with the
//      candidates staged brick by brick in LDS), k_fluid_list (Verlet lists) or k_fluid
//      (cell walk).
// The wall cell list is static and built once at create.
//
// Neighbour search.  Cells have edge R/k (k = cell_divisions, R = compact support 2h).  For
// a target at x_a the candidate cells are enumerated as x-rows ("pencils") of cells: for
// every (y, z) cell pair the transverse distance d_t from x_a to the pencil (cell interval
// grown by half the skin) prunes the pencil when d_t > R, and otherwise only the x-cells
// within sqrt(R^2 - d_t^2) + skin/2 of x_a are visited -- one contiguous range of the
// cell-sorted arrays.  Border cells extend to infinity outwards, so clamped particles
// (outside the grid box) are always visited.  Every candidate then passes the exact test
// d^2 = (dx*dx + dy*dy) + dz*dz <= R^2 (FP contraction off, the oracle's evaluation order),
// so the neighbour relation is exactly the reference's.
//
// Pair evaluation.  k_fluid is two-phase per wavefront: phase 1 runs the exact radius test
// over the candidates and appends hits to a per-lane list in LDS (kCap entries per lane);
// phase 2 evaluates the pair terms on the dense lists.  Phases alternate until every lane
// has exhausted its pencils, so any neighbour count is handled.  k_fluid_list streams the
// Verlet lists instead (about 1.1x the neighbour count per particle, read as contiguous
// 256-byte rows per wavefront) and applies the same exact test to every entry.
#include "tpsph_internal.h"

#include <mutex>

namespace tpsph {
thread_local std::string g_create_error;
}  // namespace tpsph

using namespace tpsph;

namespace {


// ------------------------------------------------------------------------------------------
// cell list (counting sort), every kernel gated by the rebuild flag (nullptr = always)
// ------------------------------------------------------------------------------------------

template <int ND>
__global__ void k_cells(const int *flag, DevParams P, int n, const double *__restrict__ x,
                        int *__restrict__ cell, int *__restrict__ rank, int *__restrict__ count)
{
    if (skip(flag)) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double *xi = x + (size_t)i * ND;
        int c0 = cell_1d(xi[0], P.lo[0], P.inv_cs[0], P.n[0]);
        int c1 = cell_1d(xi[1], P.lo[1], P.inv_cs[1], P.n[1]);
        int c2 = (ND == 3) ? cell_1d(xi[2], P.lo[2], P.inv_cs[2], P.n[2]) : 0;
        int lin = c0 + P.n[0] * (c1 + P.n[1] * c2);
        cell[i] = lin;
        rank[i] = atomicAdd(&count[lin], 1);
    }
}

// Processing order of the list kernels: the cell-sorted particles grouped by bricks of cells
// (counted by k_gather<1>), so the targets of a block are a compact cluster and the union of
// their neighbourhoods is small.  The order inside a brick follows the histogram atomics; it
// only decides which lane handles which particle, never a result.
// ord[t] = sorted index of the t-th particle in brick order; clears the brick histogram.
__global__ void k_brick_scatter(const int *flag, int n, const int *__restrict__ bid,
                                const int *__restrict__ rank, const int *__restrict__ bstart,
                                int *__restrict__ ord, int *__restrict__ bcount)
{
    if (skip(flag)) return;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const int b = bid[s];
        ord[bstart[b] + rank[s]] = s;
        bcount[b] = 0;
    }
}

// ------------------------------------------------------------------------------------------
// gather into cell-sorted order + fluid state equation (wcsph/system.jl:426-438)
//   MODE 0: runs when no rebuild is pending; flags one if a particle left its skin.  A block
//           that starts after another block raised the flag returns early: the rebuild that
//           follows gathers everything again.
//   MODE 1: runs after a rebuild; stores the build positions.
// ------------------------------------------------------------------------------------------
// MODE 1 with bcount also counts the particles of every brick of cells (the processing
// order of the list kernels: see k_brick_scatter), recording each particle's brick and rank,
// and zeroes the `nzero` counters of the list build.
// FP32 position for the pre-test of the list build: relative to the grid origin, NaN
// outside the grid box (the list build then leaves the particle to the FP64 test)
template <int ND>
__device__ __forceinline__ float4 list_f32(const DevParams &P, double x0, double x1, double x2)
{
    const double x[3] = {x0, x1, x2};
    bool inbox = true;
    float c[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < ND; d++) {
        const double r = x[d] - P.lo[d];
        inbox = inbox && r >= 0.0 && r <= P.box[d];
        c[d] = (float)r;
    }
    const float nan = __builtin_nanf("");
    return inbox ? make_float4(c[0], c[1], c[2], 0.f) : make_float4(nan, nan, nan, nan);
}

template <int ND>
__global__ void k_wall_f32(DevParams P, int nw, const double4 *__restrict__ wpos,
                           float4 *__restrict__ wxf)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const double4 x = wpos[w];
    wxf[w] = list_f32<ND>(P, x.x, x.y, x.z);
}

template <int ND, int MODE>
__global__ void k_gather(int *flag, int *work, DevParams P, const int *__restrict__ perm,
                         const double *__restrict__ u, const double *__restrict__ v,
                         const double *__restrict__ m, double *__restrict__ xs,
                         double *__restrict__ ys, double *__restrict__ zs,
                         double4 *__restrict__ rec, double2 *__restrict__ vol,
                         double *__restrict__ snap, int *__restrict__ iperm,
                         int *__restrict__ bid, int *__restrict__ rank,
                         int *__restrict__ bcount, int *__restrict__ zero, int nzero,
                         const int *__restrict__ noverflow, int *__restrict__ nrebuild,
                         float4 *__restrict__ xf)
{
    P = with_sound(P);
    const int f = *flag;
    // MODE 0 runs first in every kick: it resets the work counter of the brick kernel
    if (MODE == 0 && work && blockIdx.x == 0 && threadIdx.x == 0) work[0] = work[1] = work[2] = 0;
    if ((MODE == 0) == (f != 0)) return;
    const size_t nf = (size_t)P.nf;
    // with Verlet lists the SoA coordinates are read between rebuilds only by the cell walk
    // of overflowed lists (noverflow counts them; nullptr = always needed)
    const bool soa = !noverflow || *noverflow != 0;
    // the loop bound is uniform per wavefront (blockDim is a multiple of 64): the brick
    // counting below works on whole wavefronts
    for (int s0 = blockIdx.x * blockDim.x; s0 < P.nf; s0 += gridDim.x * blockDim.x) {
        const int s = s0 + (int)threadIdx.x;
        const bool in = s < P.nf;
        int b = -1;          // brick of the particle (MODE 1 with bcount)
        bool moved = false;  // left its skin (MODE 0)
        if (in) {
            const int i = perm[s];
            double x[3] = {0.0, 0.0, 0.0}, vv[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int d = 0; d < ND; d++) {
                x[d] = u[(size_t)i * ND + d];
                vv[d] = v[(size_t)i * P.vrows + d];
            }
            // SummationDensity: density and pressure are written by k_sum_finish
            const double rho = P.summation ? 1.0 : v[(size_t)i * (ND + 1) + ND];
            const double mass = m[i];
            if (MODE == 1 || soa) {
                xs[s] = x[0];
                ys[s] = x[1];
                if constexpr (ND == 3) zs[s] = x[2];
            }
            rec[2 * (size_t)s] = make_double4(x[0], x[1], x[2], rho);
            rec[2 * (size_t)s + 1] = make_double4(vv[0], vv[1], vv[2], state_equation(P, rho));
            vol[s] = make_double2(mass / rho, mass);
            if constexpr (MODE == 0) {
                // Euclidean displacement since the build (bounds the per-axis drift the grown
                // cells allow for, and the drift the Verlet lists allow for)
                double e2 = 0.0;
#pragma unroll
                for (int d = 0; d < ND; d++) {
                    const double e = x[d] - snap[d * nf + s];
                    e2 = fma(e, e, e2);
                }
                moved = !(e2 <= P.half_skin2);
            } else {
#pragma unroll
                for (int d = 0; d < ND; d++) snap[d * nf + s] = x[d];
                iperm[i] = s;
                if (xf) xf[s] = list_f32<ND>(P, x[0], x[1], x[2]);
                if (bcount) {
                    const int c0 = cell_1d(x[0], P.lo[0], P.inv_cs[0], P.n[0]);
                    const int c1 = cell_1d(x[1], P.lo[1], P.inv_cs[1], P.n[1]);
                    const int c2 = (ND == 3) ? cell_1d(x[2], P.lo[2], P.inv_cs[2], P.n[2]) : 0;
                    b = c0 / P.bc[0] + P.nb[0] * (c1 / P.bc[1] + P.nb[1] * (c2 / P.bc[2]));
                }
            }
        }
        if constexpr (MODE == 0) raise_rebuild_flag(flag, moved);
        if (MODE == 1 && bcount) {
            // a wavefront holds consecutive cell-sorted particles, so equal bricks come in runs:
            // the first lane of a run takes ranks for the whole run with one atomic, and lane l
            // of the run gets base + l (a rank only decides which lane handles which particle)
            const int lane = (int)threadIdx.x & 63;
            const int prev = __shfl_up(b, 1);
            const bool head = lane == 0 || b != prev;
            const unsigned long long heads = __ballot(head);
            const unsigned long long upto = (2ull << lane) - 1;  // lanes 0..lane (all for 63)
            const int first = 63 - __clzll((long long)(heads & upto));
            const unsigned long long after = heads & ~upto;
            const int end = after ? __ffsll((long long)after) - 1 : 64;
            int base = 0;
            if (head && b >= 0) base = atomicAdd(&bcount[b], end - first);
            base = __shfl(base, first);
            if (b >= 0) {
                bid[s] = b;
                rank[s] = base + (lane - first);
            }
        }
    }
    if (MODE == 1 && zero && blockIdx.x == 0 && (int)threadIdx.x < nzero) zero[threadIdx.x] = 0;
    if (MODE == 1 && nrebuild && blockIdx.x == 0 && threadIdx.x == 0) *nrebuild += 1;
}


// alist / acnt: the wall particles' Verlet lists of fluid candidates (slot-major, row
// length lw), or nullptr for the cell walk (every wall particle).  With lists only the
// active wall particles (a non-empty list; active / nactive, built at the rebuild) are
// visited: the others keep the values the rebuild gave them.  LANES lanes share a wall
// particle (lane q takes list entries q, q + LANES, ...; the partial sums are added at the
// end).  A count of -1 (overflow) walks the cells on the first lane.
template <int ND, int LANES>
__global__ void __launch_bounds__(kBlock)
    k_wall(DevParams P, FluidView F, const double4 *__restrict__ wpos, double *__restrict__ pw,
           double *__restrict__ rhow, double2 *__restrict__ wpv, const int *__restrict__ alist,
           const int *__restrict__ acnt, int lw, const int *__restrict__ active,
           const int *__restrict__ nactive, double *__restrict__ wvel)
{
    P = with_sound(P);
    const int na = alist ? *nactive : P.nw;
    const int q = threadIdx.x % LANES;
    // the loop bound is uniform per LANES lanes, so the partners shuffle together
    for (int g = (blockIdx.x * blockDim.x + threadIdx.x) / LANES; g < na;
         g += gridDim.x * blockDim.x / LANES) {
        const int w = alist ? active[g] : g;
        const double4 xw = wpos[w];
        double pressure = 0.0, volume = 0.0, vs[3] = {0.0, 0.0, 0.0};
        const int n = alist ? acnt[w] : -1;
        if (n >= 0) {
            // the next list entry is loaded before the current one is evaluated
            int j = q < n ? alist[(size_t)q * lw + w] : 0;
            for (int k = q; k < n; k += LANES) {
                const int jn = alist[(size_t)min(k + LANES, n - 1) * lw + w];
                adami_term<ND>(P, F, xw, j, pressure, volume, vs);
                j = jn;
            }
        } else if (q == 0) {
            Pencils<ND> it;
            it.init(P, xw.x, xw.y, xw.z, P.hs_fluid);
            int jb, je;
            while (it.next(P, F.start, jb, je))
                for (int j = jb; j < je; j++) adami_term<ND>(P, F, xw, j, pressure, volume, vs);
        }
#pragma unroll
        for (int o = 1; o < LANES; o <<= 1) {
            pressure += __shfl_xor(pressure, o);
            volume += __shfl_xor(volume, o);
            if (P.visc_w) {
#pragma unroll
                for (int d = 0; d < ND; d++) vs[d] += __shfl_xor(vs[d], o);
            }
        }
        if (q == 0) adami_finish(P, w, xw.w, pressure, volume, pw, rhow, wpv, vs, wvel);
    }
}


// SUM: SummationDensity (the pressure acceleration of pair_fluid<ND, true>, velocity rows
// of dv only).
template <int ND, bool SUM = false>
__global__ void __launch_bounds__(kBlock)
    k_fluid(int *flag, DevParams P, FluidView F, WallView Wv,
            const unsigned char *__restrict__ wall_near, double *__restrict__ dv)
{
    P = with_sound(P);
    if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;  // the rebuild (if any) is done
    __shared__ int lists[kCap * kBlock];
    int *lst = lists + threadIdx.x;
    const int s = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const bool active = s < P.nf;
    double xa[3] = {0.0, 0.0, 0.0}, va[3] = {0.0, 0.0, 0.0}, rho_a = 1.0, p_a = 0.0;
    if (active) {
        const double4 a = F.rec[2 * (size_t)s];
        const double4 b = F.rec[2 * (size_t)s + 1];
        xa[0] = a.x;
        xa[1] = a.y;
        xa[2] = a.z;
        rho_a = a.w;
        va[0] = b.x;
        va[1] = b.y;
        va[2] = b.z;
        p_a = b.w;
    }
    const double inv_rho_a = 1.0 / rho_a;
    if constexpr (SUM) p_a = p_a / (rho_a * rho_a);  // the target's p / rho^2
    double ga[3] = {0.0, 0.0, 0.0};  // DensityDiffusionAntuono: this particle's gradient
    if (!SUM && P.ddgrad && active) {
#pragma unroll
        for (int d = 0; d < ND; d++) ga[d] = P.ddgrad[(size_t)s * ND + d];
    }

    double acc[3] = {0.0, 0.0, 0.0}, drho = 0.0;
    neighbour_sweep<ND, false, kCap, kBlock, SUM>(P, F.start, F.x, F.y, F.z, P.hs_fluid, active,
                                                  F, Wv, lst, xa, va, rho_a, p_a, inv_rho_a,
                                                  acc, drho, 0.0,
                                                  (!SUM && P.ddgrad) ? ga : nullptr);
    double accw[3] = {0.0, 0.0, 0.0}, drhow = 0.0;
    if (P.nw > 0) {
        // walls are static: only cells flagged at create can have wall neighbours
        bool near = false;
        if (active) {
            const int c0 = cell_1d(xa[0], P.lo[0], P.inv_cs[0], P.n[0]);
            const int c1 = cell_1d(xa[1], P.lo[1], P.inv_cs[1], P.n[1]);
            const int c2 = (ND == 3) ? cell_1d(xa[2], P.lo[2], P.inv_cs[2], P.n[2]) : 0;
            near = wall_near[c0 + P.n[0] * (c1 + P.n[1] * c2)] != 0;
        }
        if (__any(near)) {
            const double m_a = (P.visc_w == 2 && active) ? F.vol[s].y : 1.0;
            neighbour_sweep<ND, true, kCap, kBlock, SUM>(P, Wv.start, Wv.x, Wv.y, Wv.z,
                                                         P.hs_wall, near, F, Wv, lst, xa, va,
                                                         rho_a, p_a, inv_rho_a, accw, drhow, m_a);
        }
    }
    if (!active) return;
    const int i = F.perm[s];
    constexpr int rows = SUM ? ND : ND + 1;
#pragma unroll
    for (int d = 0; d < ND; d++) dv[(size_t)i * rows + d] = (acc[d] + accw[d]) + P.g[d];
    if constexpr (!SUM) dv[(size_t)i * rows + ND] = drho + drhow;
}

// SummationDensity (summation_density!, density_calculators.jl:26-48): the density of every
// cell-sorted fluid particle s, sum_b m_b W(|x_s - x_b|) over the fluid particles within the
// compact support (itself included, W(0)) and then over the wall particles, the neighbour
// systems in the semidiscretization's order.  Written to rho (cell-sorted); k_sum_finish
// then sets the records, so that no lane reads a record another lane writes.
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_sum_density(DevParams P, FluidView F, WallView Wv,
                  const unsigned char *__restrict__ wall_near, double *__restrict__ rho)
{
    const int s = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (s >= P.nf) return;
    const double x0 = F.x[s], x1 = F.y[s], x2 = (ND == 3) ? F.z[s] : 0.0;
    double sum = 0.0;
    Pencils<ND> it;
    it.init(P, x0, x1, x2, P.hs_fluid);
    int jb, je;
    while (it.next(P, F.start, jb, je))
        for (int j = jb; j < je; j++) {
            const double d2 = dist2<ND>(x0 - F.x[j], x1 - F.y[j], (ND == 3) ? x2 - F.z[j] : 0.0);
            if (d2 <= P.R2) sum += F.vol[j].y * wendland_c2(P, sqrt(d2));
        }
    if (P.nw > 0) {
        const int c0 = cell_1d(x0, P.lo[0], P.inv_cs[0], P.n[0]);
        const int c1 = cell_1d(x1, P.lo[1], P.inv_cs[1], P.n[1]);
        const int c2 = (ND == 3) ? cell_1d(x2, P.lo[2], P.inv_cs[2], P.n[2]) : 0;
        if (wall_near[c0 + P.n[0] * (c1 + P.n[1] * c2)]) {
            Pencils<ND> iw;
            iw.init(P, x0, x1, x2, P.hs_wall);
            while (iw.next(P, Wv.start, jb, je))
                for (int w = jb; w < je; w++) {
                    const double4 xw = Wv.pos[w];
                    const double d2 = dist2<ND>(x0 - xw.x, x1 - xw.y, (ND == 3) ? x2 - xw.z : 0.0);
                    if (d2 <= P.R2) sum += xw.w * wendland_c2(P, sqrt(d2));
                }
        }
    }
    rho[s] = sum;
}

// The same sums over the Verlet lists (group t of four lanes takes the particle ord[t]): the
// particle itself (m W(0)), its fluid list and its wall list, each within the compact
// support; an overflowed list walks the cells instead (the SoA coordinates are current
// while some list overflows; the walls are static).  The four lanes' partial sums are added
// in a fixed order.
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_sum_density_lists(DevParams P, FluidView F, WallView Wv, const int *__restrict__ ord,
                        const int *__restrict__ nbr, const int *__restrict__ cnt,
                        const int *__restrict__ wnbr, const int *__restrict__ wcnt, int ls,
                        double *__restrict__ rho)
{
    const int q = threadIdx.x % 4;
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) / 4;
    const bool active = t < P.nf;
    const int tc = active ? t : P.nf - 1;  // an in-bounds particle for the idle lanes
    const int s = ord[tc];
    const double4 a = F.rec[2 * (size_t)s];
    const double x0 = a.x, x1 = a.y, x2 = (ND == 3) ? a.z : 0.0;
    double sum = (q == 0) ? F.vol[s].y * wendland_c2(P, 0.0) : 0.0;
    const int n = active ? cnt[tc] : 0;
    if (n >= 0) {
        for (int k = q; k < n; k += 4) {
            const int j = nbr[(size_t)k * ls + tc];
            const double4 b = F.rec[2 * (size_t)j];
            const double d2 = dist2<ND>(x0 - b.x, x1 - b.y, (ND == 3) ? x2 - b.z : 0.0);
            if (d2 <= P.R2) sum += F.vol[j].y * wendland_c2(P, sqrt(d2));
        }
    } else {
        Pencils<ND> it;
        it.init(P, x0, x1, x2, P.hs_fluid);
        int jb, je;
        while (it.next(P, F.start, jb, je))
            for (int j = jb + q; j < je; j += 4) {
                const double d2 = dist2<ND>(x0 - F.x[j], x1 - F.y[j], (ND == 3) ? x2 - F.z[j] : 0.0);
                if (d2 <= P.R2 && j != s) sum += F.vol[j].y * wendland_c2(P, sqrt(d2));
            }
    }
    const int nw = (active && P.nw > 0) ? wcnt[tc] : 0;
    if (nw > 0) {
        for (int k = q; k < nw; k += 4) {
            const double4 xw = Wv.pos[wnbr[(size_t)k * ls + tc]];
            const double d2 = dist2<ND>(x0 - xw.x, x1 - xw.y, (ND == 3) ? x2 - xw.z : 0.0);
            if (d2 <= P.R2) sum += xw.w * wendland_c2(P, sqrt(d2));
        }
    } else if (nw < 0) {
        Pencils<ND> iw;
        iw.init(P, x0, x1, x2, P.hs_wall);
        int jb, je;
        while (iw.next(P, Wv.start, jb, je))
            for (int w = jb + q; w < je; w += 4) {
                const double4 xw = Wv.pos[w];
                const double d2 = dist2<ND>(x0 - xw.x, x1 - xw.y, (ND == 3) ? x2 - xw.z : 0.0);
                if (d2 <= P.R2) sum += xw.w * wendland_c2(P, sqrt(d2));
            }
    }
    sum += __shfl_xor(sum, 1);
    sum += __shfl_xor(sum, 2);
    if (active && q == 0) rho[s] = sum;
}

// The fluid part of the sums brick by brick: the union staged in LDS as {x, y} and {z, m}
// fields (the far record: far away, zero mass), each target summing m W over its list's LDS
// indices, four lanes per target; the particle itself (m W(0)) and the wall part as in
// k_sum_density_lists.  Bricks that do not stage take the global lists.  Blocks take bricks
// from work[2].
template <int ND>
struct SumSmem {
    double2 p0[kBrickCap], p1[kBrickCap];  // {x, y}, {z, m} (2D: {0, m})
    int row_gb[kBrickMaxRows], row_lds[kBrickMaxRows + 1];
    int item;
};

template <int ND>
__global__ void __launch_bounds__(kBrickThreads)
    k_sum_density_brick(DevParams P, FluidView F, WallView Wv, const int *__restrict__ ord,
                        const int *__restrict__ bstart, const int *__restrict__ nel,
                        const int *__restrict__ nne, int *__restrict__ work,
                        const int *__restrict__ brow_gb, const int *__restrict__ brow_lds,
                        const int *__restrict__ bntot, const unsigned short *__restrict__ nbr16,
                        const int *__restrict__ nbr, const int *__restrict__ cnt,
                        const int *__restrict__ wnbr, const int *__restrict__ wcnt, int ls,
                        double *__restrict__ rho)
{
    __shared__ SumSmem<ND> S;
    const int tid = threadIdx.x;
    const int ne = *nne;
    const int cap = P.brick_cap;
    while (true) {
        if (tid == 0) S.item = atomicAdd(&work[2], 1);
        __syncthreads();
        const int item = S.item;
        if (item >= ne) break;
        const int b = nel[item];
        const int t0 = bstart[b], t1 = bstart[b + 1];
        const int ntot = bntot[b];
        const bool staged = ntot <= cap - 1;
        if (staged) {
            for (int r = tid; r <= P.maxr; r += kBrickThreads) {
                if (r < P.maxr) S.row_gb[r] = brow_gb[(size_t)b * P.maxr + r];
                S.row_lds[r] = brow_lds[(size_t)b * (P.maxr + 1) + r];
            }
            if (tid == 0) {  // the far record
                S.p0[cap - 1] = make_double2(kFar, kFar);
                S.p1[cap - 1] = make_double2(kFar, 0.0);
            }
            __syncthreads();
            for (int i = tid; i < ntot; i += kBrickThreads) {
                const int r = bsearch_le(S.row_lds, P.maxr + 1, i);
                const int g = S.row_gb[r] + (i - S.row_lds[r]);
                const double4 a = F.rec[2 * (size_t)g];
                S.p0[i] = make_double2(a.x, a.y);
                S.p1[i] = make_double2(ND == 3 ? a.z : 0.0, F.vol[g].y);
            }
            __syncthreads();
        }
        const int q = tid & 3;
        for (int base = t0; base < t1; base += kBrickThreads / 4) {
            const int t = base + tid / 4;
            const bool active = t < t1;
            const int col = active ? t : t0;
            const int s = ord[col];
            const double4 a = F.rec[2 * (size_t)s];
            const double x0 = a.x, x1 = a.y, x2 = (ND == 3) ? a.z : 0.0;
            double sum = (q == 0) ? F.vol[s].y * wendland_c2(P, 0.0) : 0.0;
            const int n = active ? cnt[col] : 0;
            if (staged) {
                const ushort4 *l4 = reinterpret_cast<const ushort4 *>(nbr16);
                const int ng = (max(n, 0) + 3) >> 2;
                // rows clamped into the list rather than a select between a load and a
                // register (which made the compiler issue flat loads)
                ushort4 cur4 = l4[(size_t)min(q, max(ng - 1, 0)) * ls + col];
                for (int g = q; g < ng; g += 4) {
                    const ushort4 nxt4 = l4[(size_t)min(g + 4, ng - 1) * ls + col];
                    const unsigned short e4[4] = {cur4.x, cur4.y, cur4.z, cur4.w};
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const double2 b0 = S.p0[e4[k]], b1 = S.p1[e4[k]];
                        const double d2 = dist2<ND>(x0 - b0.x, x1 - b0.y, (ND == 3) ? x2 - b1.x : 0.0);
                        // the far record (and the slots past the list) add zero mass
                        if (d2 <= P.R2) sum += b1.y * wendland_c2(P, sqrt(d2));
                    }
                    cur4 = nxt4;
                }
            } else if (n >= 0) {
                for (int k = q; k < n; k += 4) {
                    const int j = nbr[(size_t)k * ls + col];
                    const double4 c = F.rec[2 * (size_t)j];
                    const double d2 = dist2<ND>(x0 - c.x, x1 - c.y, (ND == 3) ? x2 - c.z : 0.0);
                    if (d2 <= P.R2) sum += F.vol[j].y * wendland_c2(P, sqrt(d2));
                }
            } else {
                Pencils<ND> it;
                it.init(P, x0, x1, x2, P.hs_fluid);
                int jb, je;
                while (it.next(P, F.start, jb, je))
                    for (int j = jb + q; j < je; j += 4) {
                        const double d2 = dist2<ND>(x0 - F.x[j], x1 - F.y[j], (ND == 3) ? x2 - F.z[j] : 0.0);
                        if (d2 <= P.R2 && j != s) sum += F.vol[j].y * wendland_c2(P, sqrt(d2));
                    }
            }
            const int nw = (active && P.nw > 0) ? wcnt[col] : 0;
            if (nw > 0) {
                for (int k = q; k < nw; k += 4) {
                    const double4 xw = Wv.pos[wnbr[(size_t)k * ls + col]];
                    const double d2 = dist2<ND>(x0 - xw.x, x1 - xw.y, (ND == 3) ? x2 - xw.z : 0.0);
                    if (d2 <= P.R2) sum += xw.w * wendland_c2(P, sqrt(d2));
                }
            } else if (nw < 0) {
                Pencils<ND> iw;
                iw.init(P, x0, x1, x2, P.hs_wall);
                int jb, je;
                while (iw.next(P, Wv.start, jb, je))
                    for (int w = jb + q; w < je; w += 4) {
                        const double4 xw = Wv.pos[w];
                        const double d2 = dist2<ND>(x0 - xw.x, x1 - xw.y, (ND == 3) ? x2 - xw.z : 0.0);
                        if (d2 <= P.R2) sum += xw.w * wendland_c2(P, sqrt(d2));
                    }
            }
            sum += __shfl_xor(sum, 1);
            sum += __shfl_xor(sum, 2);
            if (active && q == 0) rho[s] = sum;
        }
        __syncthreads();  // the LDS is staged again for the next brick
    }
}

// The records of the summed densities: rho, the state equation's pressure and V = m / rho.
__global__ void __launch_bounds__(kBlock)
    k_sum_finish(DevParams P, int nf, const double *__restrict__ rho, double4 *__restrict__ rec,
                 double2 *__restrict__ vol)
{
    P = with_sound(P);
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= nf) return;
    const double r = rho[s];
    rec[2 * (size_t)s].w = r;
    rec[2 * (size_t)s + 1].w = state_equation(P, r);
    const double m = vol[s].y;
    vol[s] = make_double2(m / r, m);
}

// Flag every cell that has a wall particle within the wall list radius of some point of the
// cell (static).
template <int ND>
__global__ void k_wall_near(DevParams P, int nw, const double4 *__restrict__ wpos,
                            unsigned char *__restrict__ near)
{
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const double4 x = wpos[w];
    // the wall list radius R + skin/2 (and margins) of the Verlet lists, which covers the
    // radius R of the cell walk
    const double reach = P.R + P.hs_wall + P.hs_list_wall;
    const int x0 = cell_1d(x.x - reach, P.lo[0], P.inv_cs[0], P.n[0]);
    const int x1 = cell_1d(x.x + reach, P.lo[0], P.inv_cs[0], P.n[0]);
    const int y0 = cell_1d(x.y - reach, P.lo[1], P.inv_cs[1], P.n[1]);
    const int y1 = cell_1d(x.y + reach, P.lo[1], P.inv_cs[1], P.n[1]);
    const int z0 = (ND == 3) ? cell_1d(x.z - reach, P.lo[2], P.inv_cs[2], P.n[2]) : 0;
    const int z1 = (ND == 3) ? cell_1d(x.z + reach, P.lo[2], P.inv_cs[2], P.n[2]) : 0;
    for (int cz = z0; cz <= z1; cz++)
        for (int cy = y0; cy <= y1; cy++)
            for (int cx = x0; cx <= x1; cx++) near[cx + P.n[0] * (cy + P.n[1] * cz)] = 1;
}

// ------------------------------------------------------------------------------------------
// Verlet lists, rebuilt together with the cell list.  Until the next rebuild no particle
// moves by more than half the skin, so a fluid pair within R now was within R + skin at the
// build and a fluid-wall pair within R + skin/2 (walls do not move).  The lists hold those
// candidates; every kick still applies the exact radius test, so the neighbour sets are the
// same as those of a full search.  Entry k of sorted particle s sits at k * ls + s: the 64
// lanes of a wavefront read one contiguous 256-byte row per slot.  A particle with more
// candidates than the capacity stores count -1 and walks the cells in the kick instead.
// ------------------------------------------------------------------------------------------
// The union rows of the brick of a target, for the LDS indices of its list (the records the
// brick kernel stages): row r of the union covers the cells (uy0 + r % nuy, uz0 + r / nuy)
// and holds the cell-sorted particles gb[r] ... in LDS slots lds[r] .. lds[r + 1] - 1.
struct BrickRows {
    const int *gb, *lds;
    int uy0, uz0, nuy;
};

// Verlet list of fluid particle s (sorted index) in column col: the candidates within the
// list radius, in the pencils' order.  BRICK: also the 16-bit LDS indices of the entries in
// the union of the target's brick, four per 8-byte group (the slots past the end of the list
// point at the far record); ok turns false if an entry lies outside the union (the brick
// then does not stage).
template <int ND, bool WALL, bool BRICK = false>
__device__ __forceinline__ void build_list(const DevParams &P, const FluidView &F,
                                           const WallView &Wv, int s, int col, int ls, int cap,
                                           int *__restrict__ nbr, int *__restrict__ cnt,
                                           const BrickRows &br = BrickRows{},
                                           unsigned long long *__restrict__ nbr16 = nullptr,
                                           bool *ok = nullptr)
{
    const double xa = F.x[s], ya = F.y[s], za = (ND == 3) ? F.z[s] : 0.0;
    const int *start = WALL ? Wv.start : F.start;
    const double *cx = WALL ? Wv.x : F.x;
    const double *cy = WALL ? Wv.y : F.y;
    const double *cz = WALL ? Wv.z : F.z;
    const double rl2 = WALL ? P.rl2_wall : P.rl2_fluid;
    // an FP32 test on the FP32 positions first; only a distance within a few FP32 ulps of
    // the box size of the list radius (or a particle outside the grid box, NaN there) takes
    // the FP64 test, so the lists are those of the FP64 test
    const float4 xaf = F.xf[s];
    const float4 *cf = WALL ? Wv.xf : F.xf;
    const float in2 = WALL ? P.wlist_in2 : P.list_in2, out2 = WALL ? P.wlist_out2 : P.list_out2;
    Pencils<ND> it;
    it.init(P, xa, ya, za, WALL ? P.hs_list_wall : P.hs_list_fluid);
    int n = 0, j, je;
    int rg = 0, rl = 0, rl1 = 0;      // BRICK: the union row of the current pencil
    bool row_ok = false, good = true;
    unsigned long long pack = 0ull;    // BRICK: the LDS indices of the open group
    const unsigned long long kFarSlot = (unsigned long long)(P.brick_cap - 1);
    while (it.next(P, start, j, je)) {
        if constexpr (BRICK) {
            // the pencil is the cell row (iy - 1, iz) of the fluid grid
            const int ry = (it.iy - 1) - br.uy0, rz = ((ND == 3) ? it.iz : 0) - br.uz0;
            const int r = ry + br.nuy * rz;
            row_ok = ry >= 0 && ry < br.nuy && rz >= 0 && r < P.maxr;
            if (row_ok) {
                rg = br.gb[r];
                rl = br.lds[r];
                rl1 = br.lds[r + 1];
            }
        }
        // four candidates per step, their loads in flight together: the SoA arrays carry
        // kPad entries past their end, so the loads stay in bounds, and j + k < je masks the
        // candidates past the range (the list order is the candidates' order)
        for (; j < je; j += 4) {
            bool hit[4];
            float4 q[4];
#pragma unroll
            for (int k = 0; k < 4; k++) q[k] = cf[j + k];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float ex = xaf.x - q[k].x, ey = xaf.y - q[k].y;
                const float ez = (ND == 3) ? xaf.z - q[k].z : 0.f;
                const float d2f = ex * ex + ey * ey + ez * ez;
                hit[k] = d2f <= in2;
                if (!hit[k] && !(d2f > out2) && j + k < je)
                    hit[k] = dist2<ND>(xa - cx[j + k], ya - cy[j + k],
                                       (ND == 3) ? za - cz[j + k] : 0.0) <= rl2;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (j + k < je && hit[k] && (WALL || j + k != s)) {
                    if (n < cap) {
                        nbr[(size_t)n * ls + col] = j + k;
                        if constexpr (BRICK) {
                            const int e = rl + (j + k - rg);
                            const bool in = row_ok && j + k >= rg && e < rl1 && e < P.brick_cap - 1;
                            good = good && in;
                            pack |= (in ? (unsigned long long)e : kFarSlot) << (16 * (n & 3));
                            if ((n & 3) == 3) {
                                nbr16[(size_t)(n >> 2) * ls + col] = pack;
                                pack = 0ull;
                            }
                        }
                    }
                    n++;
                }
            }
        }
    }
    if constexpr (BRICK) {
        if (n <= cap && (n & 3) != 0) {  // pad the last group with the far record
            for (int k = n & 3; k < 4; k++) pack |= kFarSlot << (16 * k);
            nbr16[(size_t)(n >> 2) * ls + col] = pack;
        }
        *ok = good && n <= cap;
    }
    cnt[col] = n <= cap ? n : -1;
}

// Fluid and wall lists of every fluid particle, stored in column t for the particle ord[t].
// The wall lists only for particles in cells flagged near a wall (wall_near, whose reach
// covers the wall list radius).  With bricks (nbr16 != nullptr) the fluid lists also get
// their LDS indices in the union of the particle's brick (brick_union_body ran before), and
// a brick with an entry outside its union, or with an overflowed list, is marked not to
// stage (bntot = INT_MAX).
template <int ND>
__device__ __forceinline__ void
    lists_body(int bx, int gx, const int *flag, DevParams P, FluidView F, WallView Wv, const int *__restrict__ ord,
            int ls, int cap_f, int *__restrict__ nbr, int *__restrict__ cnt, int cap_w,
            int *__restrict__ wnbr, int *__restrict__ wcnt, int *__restrict__ wtargets,
            int *__restrict__ nwtargets, int *__restrict__ noverflow,
            const unsigned char *__restrict__ wall_near, const int *__restrict__ brow_gb,
            const int *__restrict__ brow_lds, int *__restrict__ bntot,
            unsigned long long *__restrict__ nbr16)
{
    if (skip(flag)) return;
    for (int t = bx * kBlock + threadIdx.x; t < P.nf; t += gx * kBlock) {
        const int s = ord[t];
        const int c0 = cell_1d(F.x[s], P.lo[0], P.inv_cs[0], P.n[0]);
        const int c1 = cell_1d(F.y[s], P.lo[1], P.inv_cs[1], P.n[1]);
        const int c2 = (ND == 3) ? cell_1d(F.z[s], P.lo[2], P.inv_cs[2], P.n[2]) : 0;
        int b = -1;
        if (nbr16) {
            const int bxi = c0 / P.bc[0], byi = c1 / P.bc[1], bzi = c2 / P.bc[2];
            b = bxi + P.nb[0] * (byi + P.nb[1] * bzi);
            if (bntot[b] > P.brick_cap - 1) b = -1;  // the brick's union does not stage
        }
        if (b >= 0) {
            const int byi = c1 / P.bc[1], bzi = c2 / P.bc[2];
            const BrickRows br{brow_gb + (size_t)b * P.maxr, brow_lds + (size_t)b * (P.maxr + 1),
                               byi * P.bc[1] - P.ur[1], (ND == 3) ? bzi * P.bc[2] - P.ur[2] : 0,
                               P.bc[1] + 2 * P.ur[1]};
            bool ok = true;
            build_list<ND, false, true>(P, F, Wv, s, t, ls, cap_f, nbr, cnt, br, nbr16, &ok);
            if (!ok) bntot[b] = INT_MAX;
        } else {
            build_list<ND, false>(P, F, Wv, s, t, ls, cap_f, nbr, cnt);
        }
        // an overflowed list walks the cells, which need the SoA coordinates every kick
        if (cnt[t] < 0) atomicAdd(noverflow, 1);
        if (P.nw > 0) {
            if (wall_near[c0 + P.n[0] * (c1 + P.n[1] * c2)]) {
                build_list<ND, true>(P, F, Wv, s, t, ls, cap_w, wnbr, wcnt);
                if (wtargets && wcnt[t] != 0) wtargets[atomicAdd(nwtargets, 1)] = t;
            } else {
                wcnt[t] = 0;
            }
        }
    }
}

// Verlet lists of the wall particles (wall-sorted order w) over the fluid: candidates within
// R + skin/2 of the static wall particle, for the Adami pass.
template <int ND>
__device__ __forceinline__ void
    wall_lists_body(int bx, int gx, const int *flag, DevParams P, FluidView F, const double4 *__restrict__ wpos,
                 int lw, int cap, int *__restrict__ nbr, int *__restrict__ cnt,
                 int *__restrict__ active, int *__restrict__ nactive, double *__restrict__ pw,
                 double *__restrict__ rhow, double2 *__restrict__ wpv, double *__restrict__ wvel)
{
    P = with_sound(P);
    if (skip(flag)) return;
    for (int w = bx * kBlock + threadIdx.x; w < P.nw; w += gx * kBlock) {
        const double4 xw = wpos[w];
        const float4 xwf = list_f32<ND>(P, xw.x, xw.y, xw.z);
        Pencils<ND> it;
        it.init(P, xw.x, xw.y, xw.z, P.hs_list_wall);
        int n = 0, j, je;
        while (it.next(P, F.start, j, je)) {
            for (; j < je; j++) {
                // the FP32 pre-test of the list build, FP64 in its band (and for NaN)
                const float4 q = F.xf[j];
                const float ex = xwf.x - q.x, ey = xwf.y - q.y, ez = (ND == 3) ? xwf.z - q.z : 0.f;
                const float d2f = ex * ex + ey * ey + ez * ez;
                bool hit = d2f <= P.wlist_in2;
                if (!hit && !(d2f > P.wlist_out2))
                    hit = dist2<ND>(xw.x - F.x[j], xw.y - F.y[j], (ND == 3) ? xw.z - F.z[j] : 0.0) <= P.rl2_wall;
                if (hit) {
                    if (n < cap) nbr[(size_t)n * lw + w] = j;
                    n++;
                }
            }
        }
        cnt[w] = n <= cap ? n : -1;
        if (n == 0) {  // no fluid within reach until the next rebuild: the values of empty sums
            const double vs0[3] = {0.0, 0.0, 0.0};
            adami_finish(P, w, xw.w, 0.0, 0.0, pw, rhow, wpv, vs0, wvel);
        }
        else
            active[atomicAdd(nactive, 1)] = w;
    }
}

// Broadcast of lane I of each quad (four consecutive lanes) to the whole quad (DPP).
template <int I>
__device__ __forceinline__ int quad_bcast(int v)
{
    return __builtin_amdgcn_mov_dpp(v, I * 0x55, 0xF, 0xF, false);
}

// Fluid kernel over the Verlet lists.
//
// Loads are quad-cooperative.  In slot k every lane needs the 64-byte record rec[2j], rec[2j+1]
// of its own candidate j.  The four lanes of a quad fetch the four records of the quad
// together: one 16-byte chunk per lane and record, so each load instruction reads 16 whole
// records (half cache lines) instead of 64 scattered 16-byte pieces.  The chunks pass
// through a wavefront-private LDS stage (80-byte stride: the writes and the reads are both
// free of bank conflicts), from which each lane reads its own record.  The candidate
// indices are read two slots ahead, the records one slot ahead of the pair terms.  The exact
// radius test (and the zero-distance skip) selects the pairs.
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_fluid_list(int *flag, DevParams P, FluidView F, WallView Wv, const int *__restrict__ ord,
                 const int *__restrict__ nbr, const int *__restrict__ cnt,
                 const int *__restrict__ wnbr, const int *__restrict__ wcnt, int ls,
                 double *__restrict__ dv)
{
    P = with_sound(P);
    if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;  // the rebuild (if any) is done
    // 8 KiB per wavefront: the staging area (64 records x 80 bytes) while the lists are
    // streamed, then the int lists (kCap x 64 lanes) of the cell-walk fallback
    constexpr int kWaveUnits = 64 * kCap / 4;  // 16-byte units
    __shared__ double2 smem[kBlock / 64][kWaveUnits];
    const int lane = threadIdx.x & 63;
    double2 *stg = smem[threadIdx.x >> 6];
    int *lst = reinterpret_cast<int *>(stg) + lane;  // fallback lists, stride 64
    // t: position in the processing order (the list column), s: the cell-sorted particle
    const int t = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const bool active = t < P.nf;
    const int col = active ? t : 0;  // in-bounds column for the inactive lanes' loads
    const int s = ord[col];
    const int sl = s;
    double xa[3] = {0.0, 0.0, 0.0}, va[3] = {0.0, 0.0, 0.0}, rho_a = 1.0, p_a = 0.0;
    int n = 0, nw = 0;
    if (active) {
        const double4 a = F.rec[2 * (size_t)s];
        const double4 b = F.rec[2 * (size_t)s + 1];
        xa[0] = a.x;
        xa[1] = a.y;
        xa[2] = a.z;
        rho_a = a.w;
        va[0] = b.x;
        va[1] = b.y;
        va[2] = b.z;
        p_a = b.w;
        n = cnt[col];
        if (P.nw > 0) nw = wcnt[col];
    }
    const double inv_rho_a = 1.0 / rho_a;
    double ga[3] = {0.0, 0.0, 0.0};  // DensityDiffusionAntuono: this particle's gradient
    if (P.ddgrad && active) {
#pragma unroll
        for (int d = 0; d < ND; d++) ga[d] = P.ddgrad[(size_t)s * ND + d];
    }
    const double *gap = P.ddgrad ? ga : nullptr;

    double acc[3] = {0.0, 0.0, 0.0}, drho = 0.0;
    {
        const int m = max(n, 0);
        int mw = m;  // longest list of the wavefront
        for (int o = 32; o > 0; o >>= 1) mw = max(mw, __shfl_xor(mw, o));
        const double2 *rec2 = reinterpret_cast<const double2 *>(F.rec);
        const uint32_t r16 = 16u * (uint32_t)(lane & 3);
        double2 *mine = stg + 5 * (lane & ~3) + (lane & 3);  // chunk (lane & 3) of the quad's records
        const double2 *own = stg + 5 * lane;
        double2 c0, c1, c2, c3, vb;
        auto issue = [&](int j) {
            c0 = ldo(rec2, (uint32_t)quad_bcast<0>(j) * 64u + r16);
            c1 = ldo(rec2, (uint32_t)quad_bcast<1>(j) * 64u + r16);
            c2 = ldo(rec2, (uint32_t)quad_bcast<2>(j) * 64u + r16);
            c3 = ldo(rec2, (uint32_t)quad_bcast<3>(j) * 64u + r16);
            vb = ldo(F.vol, (uint32_t)j * 16u);
        };
        int jc = (m > 0) ? nbr[col] : sl;  // candidate whose record is in flight
        if (mw > 0) issue(jc);
        int jn = (1 < m) ? nbr[(size_t)ls + col] : sl;  // candidate of the next slot
        for (int k = 0; k < mw; k++) {
            mine[0] = c0;
            mine[5] = c1;
            mine[10] = c2;
            mine[15] = c3;
            __builtin_amdgcn_wave_barrier();
            const double2 a0 = own[0], a1 = own[1], b0 = own[2], b1 = own[3];
            const FluidNbr cur = {make_double4(a0.x, a0.y, a1.x, a1.y),
                                  make_double4(b0.x, b0.y, b1.x, b1.y), vb, jc};
            __builtin_amdgcn_wave_barrier();
            if (k + 1 < mw) issue(jn);
            jc = jn;
            jn = (k + 2 < m) ? nbr[(size_t)(k + 2) * ls + col] : sl;
            if (k < m) {
                const double p0 = xa[0] - cur.a.x, p1 = xa[1] - cur.a.y;
                const double p2 = (ND == 3) ? xa[2] - cur.a.z : 0.0;
                const double d2 = dist2<ND>(p0, p1, p2);
                if (d2 <= P.R2 && d2 >= P.az2)
                    pair_fluid<ND>(P, cur, xa, va, rho_a, p_a, inv_rho_a, acc, drho, gap);
            }
        }
    }
    if (__any(n < 0))
        neighbour_sweep<ND, false, kCap, 64>(P, F.start, F.x, F.y, F.z, P.hs_fluid, n < 0, F,
                                             Wv, lst, xa, va, rho_a, p_a, inv_rho_a, acc, drho,
                                             0.0, gap);

    double accw[3] = {0.0, 0.0, 0.0}, drhow = 0.0;
    if (__any(nw != 0)) {
        const double m_a = P.visc_w == 2 ? F.vol[s].y : 1.0;
        const int m = max(nw, 0);
        int j = (m > 0) ? wnbr[col] : 0;
        WallNbr cur = load_wall(Wv, j);
        for (int k = 0; k < m; k++) {
            const int jn = (k + 1 < m) ? wnbr[(size_t)(k + 1) * ls + col] : j;
            const WallNbr nxt = load_wall(Wv, jn);
            const double p0 = xa[0] - cur.a.x, p1 = xa[1] - cur.a.y;
            const double p2 = (ND == 3) ? xa[2] - cur.a.z : 0.0;
            const double d2 = dist2<ND>(p0, p1, p2);
            if (d2 <= P.R2 && d2 >= P.az2)
                pair_wall<ND>(P, cur, xa, va, rho_a, p_a, inv_rho_a, accw, drhow, m_a);
            cur = nxt;
            j = jn;
        }
        if (__any(nw < 0))
            neighbour_sweep<ND, true, kCap, 64>(P, Wv.start, Wv.x, Wv.y, Wv.z, P.hs_wall, nw < 0,
                                                F, Wv, lst, xa, va, rho_a, p_a, inv_rho_a, accw,
                                                drhow, m_a);
    }
    if (!active) return;
    const int i = F.perm[s];
#pragma unroll
    for (int d = 0; d < ND; d++) dv[(size_t)i * (ND + 1) + d] = (acc[d] + accw[d]) + P.g[d];
    dv[(size_t)i * (ND + 1) + ND] = drho + drhow;
}


// ------------------------------------------------------------------------------------------
// Bricks in LDS.  A brick is a block of bc[0] x bc[1] x bc[2] cells; its targets are the
// fluid particles in those cells (ord[bstart[b] .. bstart[b+1])).  Its union is every fluid
// particle in a cell row (y, z) near the brick, trimmed in x: the cells within the list radius
// R + skin of the brick box, as contiguous ranges of the cell-sorted arrays.  Every Verlet list
// entry of a target lies in the union of its brick, so after one coalesced staging pass of the
// union into LDS the pair terms read every candidate from LDS.  The lists also store the
// candidate's LDS index (u16, four slots per ushort4; the slots past the end of a list hold
// the index of the far record, kBrickCap - 1, which no target reaches).  A brick whose union
// does not leave room for the far record, or with a target whose list overflowed, is marked
// (bntot = INT_MAX) and its targets read their candidates from global memory.
// ------------------------------------------------------------------------------------------

// Union rows of every brick (one block per brick, grid-stride) and, in the first block, the
// list of non-empty bricks.  Row r = (cy - (by bc[1] - ur[1])) + nuy (cz - (bz bc[2] - ur[2])).
template <int ND>
__device__ __forceinline__ void
    brick_union_body(int bx, int gx, const int *flag, DevParams P, const int *__restrict__ start,
                  const int *__restrict__ bstart, int nbricks, int *__restrict__ brow_gb,
                  int *__restrict__ brow_lds, int *__restrict__ bntot, int *__restrict__ nel,
                  int *__restrict__ nne)
{
    if (skip(flag)) return;
    __shared__ int scan_buf[kBlock];
    const int tid = threadIdx.x;
    if (bx == 0) {
        // the non-empty bricks in brick order (the order the brick kernel takes them in;
        // neighbouring bricks in flight share their unions in L2)
        int carry = 0;
        for (int c = 0; c < nbricks; c += kScanTile) {
            const int b0 = c + tid * kScanItems;
            int f[kScanItems], n = 0;
#pragma unroll
            for (int k = 0; k < kScanItems; k++) {
                const int b = b0 + k;
                f[k] = (b < nbricks && bstart[b + 1] > bstart[b]) ? 1 : 0;
                n += f[k];
            }
            int tot;
            int pos = carry + block_exclusive_scan(n, scan_buf, tot);
#pragma unroll
            for (int k = 0; k < kScanItems; k++)
                if (f[k]) nel[pos++] = b0 + k;
            carry += tot;
        }
        if (tid == 0) *nne = carry;
    }
    const int nuy = P.bc[1] + 2 * P.ur[1];
    for (int b = bx; b < nbricks; b += gx) {
        // only the bricks with targets are staged or looked up (most of a tank's bricks hold
        // no fluid); the condition is uniform over the block
        if (bstart[b + 1] == bstart[b]) continue;
        const int bx = b % P.nb[0], by = (b / P.nb[0]) % P.nb[1], bz = b / (P.nb[0] * P.nb[1]);
        const int X0 = bx * P.bc[0], X1 = min(X0 + P.bc[0], P.n[0]) - 1;
        const int Y0 = by * P.bc[1], Y1 = min(Y0 + P.bc[1], P.n[1]) - 1;
        const int Z0 = (ND == 3) ? bz * P.bc[2] : 0;
        const int Z1 = (ND == 3) ? min(Z0 + P.bc[2], P.n[2]) - 1 : 0;
        const double m = P.hs_union;
        const double bxlo = cell_lo(P, 0, X0, m), bxhi = cell_hi(P, 0, X1, m);
        const double bylo = cell_lo(P, 1, Y0, m), byhi = cell_hi(P, 1, Y1, m);
        const double bzlo = (ND == 3) ? cell_lo(P, 2, Z0, m) : 0.0;
        const double bzhi = (ND == 3) ? cell_hi(P, 2, Z1, m) : 0.0;
        int cntr = 0, gb = 0;
        if (tid < P.maxr) {
            const int cy = Y0 - P.ur[1] + tid % nuy;
            const int cz = (ND == 3) ? Z0 - P.ur[2] + tid / nuy : 0;
            if (cy >= 0 && cy < P.n[1] && cz >= 0 && cz < P.n[2]) {
                const double dy = fmax(0.0, fmax(cell_lo(P, 1, cy, m) - byhi, bylo - cell_hi(P, 1, cy, m)));
                double dz = 0.0;
                if constexpr (ND == 3)
                    dz = fmax(0.0, fmax(cell_lo(P, 2, cz, m) - bzhi, bzlo - cell_hi(P, 2, cz, m)));
                const double r2 = P.rl_union2 - (dy * dy + dz * dz);
                if (r2 >= 0.0) {
                    const double rx = sqrt(r2) * (1.0 + 1e-9) + m;
                    const int x0 = cell_1d(bxlo - rx, P.lo[0], P.inv_cs[0], P.n[0]);
                    const int x1 = cell_1d(bxhi + rx, P.lo[0], P.inv_cs[0], P.n[0]);
                    const int row = P.n[0] * (cy + P.n[1] * cz);
                    gb = start[row + x0];
                    cntr = start[row + x1 + 1] - gb;
                }
            }
            brow_gb[(size_t)b * P.maxr + tid] = gb;
        }
        int tot;
        const int lds0 = block_exclusive_scan(cntr, scan_buf, tot);
        if (tid < P.maxr) brow_lds[(size_t)b * (P.maxr + 1) + tid] = lds0;
        if (tid == 0) {
            brow_lds[(size_t)b * (P.maxr + 1) + P.maxr] = tot;
            bntot[b] = tot;
        }
    }
}

struct ListsArgs {
    const int *ord;
    int ls, cap_f;
    int *nbr, *cnt;
    int cap_w;
    int *wnbr, *wcnt, *wtargets, *nwtargets, *noverflow;
    const unsigned char *wall_near;
    const int *brow_gb, *brow_lds;
    int *bntot;
    unsigned long long *nbr16;  // nullptr without bricks
};
struct UnionArgs {
    const int *start, *bstart;
    int nbricks;
    int *brow_gb, *brow_lds, *bntot, *nel, *nne;
};
struct WallListsArgs {
    const double4 *wpos;
    int lw, cap;
    int *nbr, *cnt, *active, *nactive;
    double *pw, *rhow;
    double2 *wpv;
    double *wvel;
};

// The rebuild's union tables of the bricks (first nblocks_union blocks) beside the Verlet
// lists of the wall particles (the Adami pass): both read only the new cell-sorted order.
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_brick_union_and_walls(const int *flag, DevParams P, FluidView F, UnionArgs b,
                            WallListsArgs w, int nblocks_union)
{
    if ((int)blockIdx.x < nblocks_union)
        brick_union_body<ND>(blockIdx.x, nblocks_union, flag, P, b.start, b.bstart, b.nbricks,
                             b.brow_gb, b.brow_lds, b.bntot, b.nel, b.nne);
    else
        wall_lists_body<ND>(blockIdx.x - nblocks_union, gridDim.x - nblocks_union, flag, P, F,
                            w.wpos, w.lw, w.cap, w.nbr, w.cnt, w.active, w.nactive, w.pw, w.rhow,
                            w.wpv, w.wvel);
}

// The fluid particles' Verlet lists (with their LDS indices when the bricks stage) and wall
// lists, after the union tables.  Its candidate loads wait on L2 latency: a grid of one
// particle per thread with registers bounded for six waves per SIMD (80 VGPRs, a few
// spilled outside the candidate loop) builds the lists 9% faster than four waves per SIMD
// at 92 VGPRs; eight waves (64 VGPRs) spill inside the loop and are slower.
template <int ND>
__global__ void __launch_bounds__(kBlock, 6)
    k_lists(const int *flag, DevParams P, FluidView F, WallView Wv, ListsArgs a)
{
    lists_body<ND>(blockIdx.x, gridDim.x, flag, P, F, Wv, a.ord, a.ls, a.cap_f, a.nbr, a.cnt,
                   a.cap_w, a.wnbr, a.wcnt, a.wtargets, a.nwtargets, a.noverflow, a.wall_near,
                   a.brow_gb, a.brow_lds, a.bntot, a.nbr16);
}

// Staged union: one record per particle as four 16-byte fields and one 8-byte field, each in
// an array of its own, read with one ds_read_b128 (ds_read_b64) per field.  The 16-byte
// stride spreads random candidates over all banks; 72- or 64-byte records made the lanes of
// a read collide on fewer bank groups (measured 2% and 12% slower).  The last record is the
// far record that pads the list groups (finite values, zero volume).  With a global brick
// the same bytes hold the lanes' int lists.
// ANT (DensityDiffusionAntuono): also the renormalised density gradient, g0 {gx, gy} and in
// 3D g1 {gz}, in CAP = kBrickCapGrad records there (96 bytes each: the LDS holds fewer).
template <int ND, bool ANT = false>
struct BrickSmem {
    static constexpr int CAP = (ANT && ND == 3) ? kBrickCapGrad : kBrickCap;
    // 3D: p0 {x, y}, p1 {z, rho}, p2 {vx, vy}, p3 {vz, p};  2D: p0 {x, y}, p1 {rho, vx},
    // p2 {vy, p} (p3 unused); vol: V = m / rho
    double2 p0[CAP], p1[CAP], p2[CAP], p3[ND == 3 ? CAP : 1];
    double vol[CAP];
    double2 g0[ANT ? CAP : 1];
    double g1[(ANT && ND == 3) ? CAP : 1];
    int row_gb[kBrickMaxRows], row_lds[kBrickMaxRows + 1];
    int item;
};

template <int ND, bool ANT>
__device__ __forceinline__ FluidNbr lds_fluid(const BrickSmem<ND, ANT> &S, int e)
{
    FluidNbr n;
    const double2 a = S.p0[e], b = S.p1[e], c = S.p2[e];
    const double V = S.vol[e];
    if constexpr (ND == 3) {
        const double2 d = S.p3[e];
        n.a = make_double4(a.x, a.y, b.x, b.y);
        n.b = make_double4(c.x, c.y, d.x, d.y);
        n.v = make_double2(V, V * b.y);
    } else {
        n.a = make_double4(a.x, a.y, 0.0, b.x);
        n.b = make_double4(b.y, c.x, 0.0, c.y);
        n.v = make_double2(V, V * b.x);
    }
    return n;
}

// The density gradient of staged record e (DensityDiffusionAntuono).
template <int ND, bool ANT>
__device__ __forceinline__ void lds_grad(const BrickSmem<ND, ANT> &S, int e, double *g)
{
    const double2 a = S.g0[e];
    g[0] = a.x;
    g[1] = a.y;
    g[2] = (ND == 3) ? S.g1[e] : 0.0;
}

// Fluid-fluid kernel over the bricks (one block of kBrickThreads per non-empty brick, taken
// from a work counter that k_gather<0> resets).  Targets with wall candidates are completed
// by k_fluid_walls.
// GLOBAL_ONLY: only the bricks that do not stage, which the mixed-precision pass
// (tpsph_mixed.hip) leaves to this kernel: the work[1] bricks it listed in bfall, taken by
// blocks in turn.  VISC: visc_mode(P), the viscosity form of the staged pair loop.  ANT:
// DensityDiffusionAntuono (P.ddgrad), whose gradients are staged beside the records.  SUM:
// SummationDensity (pressure slot staged as p / rho^2, no density rate, dv of ND rows).
template <int ND, bool GLOBAL_ONLY = false, int VISC = kViscAlpha, bool ANT = false,
          bool SUM = false>
__global__ void __launch_bounds__(kBrickThreads)
    k_fluid_brick(int *flag, DevParams P, FluidView F, WallView Wv, const int *__restrict__ ord,
                  const int *__restrict__ bstart, const int *__restrict__ nel,
                  const int *__restrict__ nne, int *__restrict__ work,
                  const int *__restrict__ brow_gb, const int *__restrict__ brow_lds,
                  const int *__restrict__ bntot, const unsigned short *__restrict__ nbr16,
                  const int *__restrict__ nbr, const int *__restrict__ cnt,
                  const int *__restrict__ wnbr, const int *__restrict__ wcnt, int ls,
                  double *__restrict__ dv, const int *__restrict__ bfall = nullptr)
{
    P = with_sound(P);
    // cell-walk fallback lists of a global brick: as many entries per lane as the staging
    // area holds (at most kCap)
    using Smem = BrickSmem<ND, ANT>;
    constexpr int kCapB = Smem::CAP;
    constexpr int kFallbackCap = std::min<int>(kCap, sizeof(Smem) / (sizeof(int) * kBrickThreads));
    static_assert(kFallbackCap >= 8, "the fallback lists must fit the staging area");
    __shared__ Smem S;
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && tid == 0) *flag = 0;  // the rebuild (if any) is done
    int *lst = reinterpret_cast<int *>(&S) + tid;  // fallback lists, stride kBrickThreads
    const int ne = GLOBAL_ONLY ? work[1] : *nne;
    for (int turn = blockIdx.x;; turn += gridDim.x) {
        int item = turn;
        if (!GLOBAL_ONLY) {
            if (tid == 0) S.item = atomicAdd(work, 1);
            __syncthreads();
            item = S.item;
        }
        if (item >= ne) break;
        const int b = GLOBAL_ONLY ? bfall[item] : nel[item];
        const int t0 = bstart[b], t1 = bstart[b + 1];
        const int ntot = bntot[b];
        const bool staged = ntot <= kCapB - 1;
        if (GLOBAL_ONLY && staged) continue;  // uniform per block; no LDS is touched
        if (staged) {
            for (int r = tid; r <= P.maxr; r += kBrickThreads) {
                if (r < P.maxr) S.row_gb[r] = brow_gb[(size_t)b * P.maxr + r];
                S.row_lds[r] = brow_lds[(size_t)b * (P.maxr + 1) + r];
            }
            if (tid == 0) {  // the far record: coordinates kFar, rho 1, the rest 0
                constexpr int f = kCapB - 1;
                if constexpr (ANT) {
                    S.g0[f] = make_double2(0.0, 0.0);
                    if constexpr (ND == 3) S.g1[f] = 0.0;
                }
                S.p0[f] = make_double2(kFar, kFar);
                if constexpr (ND == 3) {
                    S.p1[f] = make_double2(kFar, 1.0);
                    S.p2[f] = make_double2(0.0, 0.0);
                    S.p3[f] = make_double2(0.0, 0.0);
                } else {
                    S.p1[f] = make_double2(1.0, 0.0);
                    S.p2[f] = make_double2(0.0, 0.0);
                }
                S.vol[f] = 0.0;
            }
            __syncthreads();
            for (int i = tid; i < ntot; i += kBrickThreads) {
                const int r = bsearch_le(S.row_lds, P.maxr + 1, i);
                const int g = S.row_gb[r] + (i - S.row_lds[r]);
                const double4 a = F.rec[2 * (size_t)g];
                const double4 c = F.rec[2 * (size_t)g + 1];
                const double pslot = SUM ? c.w / (a.w * a.w) : c.w;
                if constexpr (ND == 3) {
                    S.p0[i] = make_double2(a.x, a.y);
                    S.p1[i] = make_double2(a.z, a.w);
                    S.p2[i] = make_double2(c.x, c.y);
                    S.p3[i] = make_double2(c.z, pslot);
                } else {
                    S.p0[i] = make_double2(a.x, a.y);
                    S.p1[i] = make_double2(a.w, c.x);
                    S.p2[i] = make_double2(c.y, pslot);
                }
                S.vol[i] = F.vol[g].x;
                if constexpr (ANT) {
                    const double *gg = P.ddgrad + (size_t)g * ND;
                    S.g0[i] = make_double2(gg[0], gg[1]);
                    if constexpr (ND == 3) S.g1[i] = gg[2];
                }
            }
            __syncthreads();
        }
        // Four lanes per target: lane q of target i takes every fourth group of four list
        // slots, starting at group q (every fourth slot of the global lists); the four
        // partial sums are added across the lanes (a fixed butterfly) and lane 0 writes.
        // With four lanes a brick's targets (about 385) take two passes of 256 whose waves
        // spread evenly over the four SIMDs; two lanes left a quarter of the waves idle and
        // one SIMD with the most work (8% slower).
        constexpr int kL = 4;
        const int half = tid & (kL - 1);
        for (int base = t0; base < t1; base += kBrickThreads / kL) {
            const int t = base + (tid / kL);
            const bool active = t < t1;
            const int col = active ? t : t0;
            const int s = ord[col];
            double xa[3] = {0.0, 0.0, 0.0}, va[3] = {0.0, 0.0, 0.0}, rho_a = 1.0, p_a = 0.0;
            int n = 0, nw = 0;
            if (active) {
                const double4 a = F.rec[2 * (size_t)s];
                const double4 c = F.rec[2 * (size_t)s + 1];
                xa[0] = a.x;
                xa[1] = a.y;
                xa[2] = a.z;
                rho_a = a.w;
                va[0] = c.x;
                va[1] = c.y;
                va[2] = c.z;
                p_a = c.w;
                n = cnt[col];
                if (P.nw > 0) nw = wcnt[col];
            }
            const double inv_rho_a = 1.0 / rho_a;
            if constexpr (SUM) p_a = p_a / (rho_a * rho_a);  // the target's p / rho^2
            double acc[3] = {0.0, 0.0, 0.0}, drho = 0.0;
            double ga[3] = {0.0, 0.0, 0.0};  // DensityDiffusionAntuono: the target's gradient
            if constexpr (ANT) {
                if (active) {
#pragma unroll
                    for (int d = 0; d < ND; d++) ga[d] = P.ddgrad[(size_t)s * ND + d];
                }
            }
            const double *gap = ANT ? ga : nullptr;
            const int m = max(n, 0);
            if (staged) {
                // four LDS indices per 8-byte load, one group ahead of the pair terms
                const ushort4 *l4 = reinterpret_cast<const ushort4 *>(nbr16);
                const int ng = (m + 3) >> 2;
                // rows clamped into the list (row 0 always exists) rather than a select between
                // a load and a register: the select made the compiler keep cur4 in scratch
                // and issue a flat load, whose wait also held up every LDS read
                ushort4 cur4 = l4[(size_t)min(half, max(ng - 1, 0)) * ls + col];
                for (int g = half; g < ng; g += kL) {
                    const ushort4 nxt4 = l4[(size_t)min(g + kL, ng - 1) * ls + col];
                    // the candidates are independent chains without branches: two at a time
                    // interleave within the register budget of four waves per SIMD
#pragma unroll 1
                    for (int h2 = 0; h2 < 2; h2++) {
                        // slots past the end of the list read the far record: exact zeros
                        const int ea = h2 ? cur4.z : cur4.x, eb = h2 ? cur4.w : cur4.y;
                        const FluidNbr na = lds_fluid<ND>(S, ea);
                        const FluidNbr nb = lds_fluid<ND>(S, eb);
                        if constexpr (ANT) {
                            double gna[3], gnb[3];
                            lds_grad<ND>(S, ea, gna);
                            lds_grad<ND>(S, eb, gnb);
                            pair_fluid_masked<ND, VISC, true>(P, na, true, xa, va, rho_a, p_a,
                                                              inv_rho_a, acc, drho, ga, gna);
                            pair_fluid_masked<ND, VISC, true>(P, nb, true, xa, va, rho_a, p_a,
                                                              inv_rho_a, acc, drho, ga, gnb);
                        } else {
                            pair_fluid_masked<ND, VISC, false, SUM>(P, na, true, xa, va, rho_a, p_a,
                                                                    inv_rho_a, acc, drho);
                            pair_fluid_masked<ND, VISC, false, SUM>(P, nb, true, xa, va, rho_a, p_a,
                                                                    inv_rho_a, acc, drho);
                        }
                    }
                    cur4 = nxt4;
                }
            } else {
                for (int k = half; k < m; k += kL) {
                    const FluidNbr nb = load_fluid(F, nbr[(size_t)k * ls + col]);
                    const double p0 = xa[0] - nb.a.x, p1 = xa[1] - nb.a.y;
                    const double p2 = (ND == 3) ? xa[2] - nb.a.z : 0.0;
                    const double d2 = dist2<ND>(p0, p1, p2);
                    if (d2 <= P.R2 && d2 >= P.az2)
                        pair_fluid<ND, SUM>(P, nb, xa, va, rho_a, p_a, inv_rho_a, acc, drho, gap);
                }
                if (__any(n < 0))
                    neighbour_sweep<ND, false, kFallbackCap, kBrickThreads, SUM>(
                        P, F.start, F.x, F.y, F.z, P.hs_fluid, n < 0 && half == 0, F, Wv, lst, xa,
                        va, rho_a, p_a, inv_rho_a, acc, drho, 0.0, gap);
            }
#pragma unroll
            for (int o = 1; o < kL; o <<= 1) {
#pragma unroll
                for (int d = 0; d < 3; d++) acc[d] += __shfl_xor(acc[d], o);
                drho += __shfl_xor(drho, o);
            }
            if (active && half == 0) {
                // with wall candidates k_fluid_walls adds the wall sum and then gravity
                constexpr int rows = SUM ? ND : ND + 1;
                const int i = F.perm[s];
                const bool walls = nw != 0;
#pragma unroll
                for (int d = 0; d < ND; d++)
                    dv[(size_t)i * rows + d] = walls ? acc[d] : acc[d] + P.g[d];
                if constexpr (!SUM) dv[(size_t)i * rows + ND] = drho;
            }
        }
        __syncthreads();  // the LDS is staged again for the next brick
    }
}

// Wall sums of the targets with wall candidates (wtargets, built at the rebuild) after
// k_fluid_brick: dv = (dv + wall sum) + gravity, the reference's order of the sums.
// Four lanes share a target: lane q takes list entries q, q + 4, ...; the sums are added
// across the four lanes before the write.
// SUM: SummationDensity (the target's p / rho^2 in the pressure slot, dv of ND rows).
template <int ND, bool SUM = false>
__global__ void __launch_bounds__(kBlock)
    k_fluid_walls(DevParams P, FluidView F, WallView Wv, const int *__restrict__ ord,
                  const int *__restrict__ wtargets, const int *__restrict__ nwtargets,
                  const int *__restrict__ wnbr, const int *__restrict__ wcnt, int ls,
                  double *__restrict__ dv)
{
    P = with_sound(P);
    constexpr int kLanes = 4;
    __shared__ int lists[kCap * kBlock];  // only for particles whose wall list overflowed
    int *lst = lists + threadIdx.x;
    const int nt = *nwtargets;
    const int q = threadIdx.x % kLanes;
    for (int base = blockIdx.x * (kBlock / kLanes); base < nt; base += gridDim.x * (kBlock / kLanes)) {
        const int idx = base + (int)threadIdx.x / kLanes;
        const bool active = idx < nt;
        const int t = wtargets[active ? idx : base];
        const int s = ord[t];
        const double4 a = F.rec[2 * (size_t)s];
        const double4 c = F.rec[2 * (size_t)s + 1];
        const double xa[3] = {a.x, a.y, a.z}, va[3] = {c.x, c.y, c.z};
        const double rho_a = a.w, inv_rho_a = 1.0 / rho_a;
        const double p_a = SUM ? c.w / (rho_a * rho_a) : c.w;
        const double m_a = P.visc_w == 2 ? F.vol[s].y : 1.0;
        const int nw = active ? wcnt[t] : 0;
        double accw[3] = {0.0, 0.0, 0.0}, drhow = 0.0;
        const int m = max(nw, 0);
        // the next list entry is loaded before the current one is evaluated
        int jw = q < m ? wnbr[(size_t)q * ls + t] : 0;
        for (int k = q; k < m; k += kLanes) {
            const WallNbr cur = load_wall(Wv, jw);
            jw = wnbr[(size_t)min(k + kLanes, m - 1) * ls + t];
            const double p0 = xa[0] - cur.a.x, p1 = xa[1] - cur.a.y;
            const double p2 = (ND == 3) ? xa[2] - cur.a.z : 0.0;
            const double d2 = dist2<ND>(p0, p1, p2);
            if (d2 <= P.R2 && d2 >= P.az2)
                pair_wall<ND, SUM>(P, cur, xa, va, rho_a, p_a, inv_rho_a, accw, drhow, m_a);
        }
        if (__any(nw < 0))
            neighbour_sweep<ND, true, kCap, kBlock, SUM>(P, Wv.start, Wv.x, Wv.y, Wv.z, P.hs_wall,
                                                         nw < 0 && q == 0, F, Wv, lst, xa, va,
                                                         rho_a, p_a, inv_rho_a, accw, drhow, m_a);
#pragma unroll
        for (int o = 1; o < kLanes; o <<= 1) {
#pragma unroll
            for (int d = 0; d < 3; d++) accw[d] += __shfl_xor(accw[d], o);
            drhow += __shfl_xor(drhow, o);
        }
        if (active && q == 0) {
            double *o = dv + (size_t)F.perm[s] * (SUM ? ND : ND + 1);
#pragma unroll
            for (int d = 0; d < ND; d++) o[d] = (o[d] + accw[d]) + P.g[d];
            if constexpr (!SUM) o[ND] = o[ND] + drhow;
        }
    }
}

// DensityDiffusionAntuono's update! (density_diffusion.jl:176-211), over the fluid
// neighbours b of fluid particle a:
//   M_a = -sum_b V_b grad W_ab (x) r_ab        (compute_gradient_correction_matrix!,
//                                               corrections.jl:329-354)
//   L_a = M_a^-1, or the identity where |det M_a| < 1f-9
//                                              (correction_matrix_inversion_step!,
//                                               corrections.jl:425-457)
//   grad rho_a = -sum_b (rho_a - rho_b) V_b L_a grad W_ab = -L_a s_a,
//   s_a = sum_b (rho_a - rho_b) V_b grad W_ab  (L_a is the same for every pair of a, so it
//                                               multiplies the sum once)
// with V_b = m_b / rho_b and grad W zero at distance >= R and at distance^2 < eps(h^2)
// (smoothing_kernel_grad, abstract_system.jl:142-155).  M_a is symmetric: its upper
// triangle is summed.  Four lanes share a particle; their partial sums are added with
// shuffles.
struct GradSums {
    double m[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // M_00 M_01 M_02 M_11 M_12 M_22
    double sv[3] = {0.0, 0.0, 0.0};
};

// One pair's terms; vk = V_b |grad W| / r (zero for a pair outside the kernel's support).
template <int ND>
__device__ __forceinline__ void grad_add(GradSums &g, double vk, double drr, double p0,
                                         double p1, double p2)
{
    const double f = drr * vk;
    g.m[0] -= vk * p0 * p0;
    g.m[1] -= vk * p0 * p1;
    g.m[3] -= vk * p1 * p1;
    g.sv[0] += f * p0;
    g.sv[1] += f * p1;
    if constexpr (ND == 3) {
        g.m[2] -= vk * p0 * p2;
        g.m[4] -= vk * p1 * p2;
        g.m[5] -= vk * p2 * p2;
        g.sv[2] += f * p2;
    }
}

// V_b |grad W| / r of a pair at squared distance d2, zero outside [eps(h^2), R^2) without a
// branch (the square root of an excluded d2 may be NaN: it is selected away).
__device__ __forceinline__ double grad_vk(const DevParams &P, double V_b, double d2)
{
    const bool in = (d2 < P.R2) & (d2 >= P.eps_hh);
    const double r = sqrt_pos(d2);
    const double w = fma(-0.5 * P.h_inv, r, 1.0);
    return in ? V_b * (P.kgrad_c * (w * w * w)) : 0.0;
}

// Sums of the four lanes of a quad (a fixed butterfly), then the inversion and the gradient
// of particle s (lane 0 of an active quad writes it).
template <int ND>
__device__ __forceinline__ void grad_finish(const DevParams &P, GradSums &g, bool write, int s)
{
#pragma unroll
    for (int o = 1; o < 4; o <<= 1) {
#pragma unroll
        for (int i = 0; i < 6; i++)
            if (ND == 3 || i == 0 || i == 1 || i == 3) g.m[i] += __shfl_xor(g.m[i], o);
#pragma unroll
        for (int i = 0; i < ND; i++) g.sv[i] += __shfl_xor(g.sv[i], o);
    }
    if (!write) return;
    const double *m = g.m, *sv = g.sv;
    constexpr double kSingular = (double)1.0e-9f;  // the Float32 literal 1.0f-9
    double gr[3];
    if constexpr (ND == 2) {
        const double det = m[0] * m[3] - m[1] * m[1];
        if (fabs(det) < kSingular) {
            gr[0] = -sv[0];
            gr[1] = -sv[1];
        } else {
            const double id = 1.0 / det;
            gr[0] = -(m[3] * sv[0] - m[1] * sv[1]) * id;
            gr[1] = -(m[0] * sv[1] - m[1] * sv[0]) * id;
        }
    } else {
        // cofactors of the symmetric M (its adjugate is symmetric too)
        const double c00 = m[3] * m[5] - m[4] * m[4];
        const double c01 = m[2] * m[4] - m[1] * m[5];
        const double c02 = m[1] * m[4] - m[2] * m[3];
        const double c11 = m[0] * m[5] - m[2] * m[2];
        const double c12 = m[1] * m[2] - m[0] * m[4];
        const double c22 = m[0] * m[3] - m[1] * m[1];
        const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
        if (fabs(det) < kSingular) {
            gr[0] = -sv[0];
            gr[1] = -sv[1];
            gr[2] = -sv[2];
        } else {
            const double id = 1.0 / det;
            gr[0] = -(c00 * sv[0] + c01 * sv[1] + c02 * sv[2]) * id;
            gr[1] = -(c01 * sv[0] + c11 * sv[1] + c12 * sv[2]) * id;
            gr[2] = -(c02 * sv[0] + c12 * sv[1] + c22 * sv[2]) * id;
        }
    }
#pragma unroll
    for (int d = 0; d < ND; d++) P.ddgrad[(size_t)s * ND + d] = gr[d];
}

// The sums of particle s from global memory: its Verlet list (n >= 0 entries in column t,
// lane q takes entries q, q + 4, ...) or, for an overflowed list or the cell search
// (n < 0), the cell walk over the SoA coordinates, which are current whenever they are
// walked.
template <int ND>
__device__ __forceinline__ void grad_sums_global(const DevParams &P, const FluidView &F,
                                                 GradSums &g, const double *xa, double rho_a,
                                                 int n, const int *__restrict__ nbr, int t,
                                                 int ls, int q)
{
    if (n >= 0) {
        for (int k = q; k < n; k += 4) {
            const int j = nbr[(size_t)k * ls + t];
            const double4 b = F.rec[2 * (size_t)j];
            const double p0 = xa[0] - b.x, p1 = xa[1] - b.y;
            const double p2 = (ND == 3) ? xa[2] - b.z : 0.0;
            const double d2 = dist2<ND>(p0, p1, p2);
            if (d2 < P.R2 && d2 >= P.eps_hh)
                grad_add<ND>(g, grad_vk(P, F.vol[j].x, d2), rho_a - b.w, p0, p1, p2);
        }
    } else {
        Pencils<ND> it;
        it.init(P, xa[0], xa[1], xa[2], P.hs_fluid);
        int jb, je;
        while (it.next(P, F.start, jb, je)) {
            for (int j = jb + q; j < je; j += 4) {
                const double p0 = xa[0] - F.x[j], p1 = xa[1] - F.y[j];
                const double p2 = (ND == 3) ? xa[2] - F.z[j] : 0.0;
                const double d2 = dist2<ND>(p0, p1, p2);
                if (d2 < P.R2 && d2 >= P.eps_hh)
                    grad_add<ND>(g, grad_vk(P, F.vol[j].x, d2), rho_a - F.rec[2 * (size_t)j].w,
                                 p0, p1, p2);
            }
        }
    }
}

// The gradients over the Verlet lists (group t takes the particle ord[t] and its list) or,
// without lists (nbr == nullptr), the cell walk of every particle.
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_antuono_gradient(DevParams P, FluidView F, const int *__restrict__ ord,
                       const int *__restrict__ nbr, const int *__restrict__ cnt, int ls)
{
    const int q = threadIdx.x % 4;
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) / 4;
    const bool active = t < P.nf;
    const int tc = active ? t : P.nf - 1;  // an in-bounds particle for the idle lanes
    const int s = nbr ? ord[tc] : tc;
    const double4 a = F.rec[2 * (size_t)s];
    const double xa[3] = {a.x, a.y, a.z};
    GradSums g;
    const int n = !active ? 0 : (nbr ? cnt[t] : -1);
    grad_sums_global<ND>(P, F, g, xa, a.w, n, nbr, t, ls, q);
    grad_finish<ND>(P, g, active && q == 0, s);
}

// The gradients brick by brick (the lists' LDS indices, the union staged in LDS as
// {x, y} and {z, rho} fields and V, the far record padding the list groups), as
// k_fluid_brick takes the fluid pass; bricks that do not stage read global memory.  Blocks
// take bricks from work[2].
template <int ND>
struct GradSmem {
    double2 p0[kBrickCap], p1[kBrickCap];  // {x, y}, {z, rho} (2D: {0, rho})
    double vol[kBrickCap];
    int row_gb[kBrickMaxRows], row_lds[kBrickMaxRows + 1];
    int item;
};

template <int ND>
__global__ void __launch_bounds__(kBrickThreads)
    k_antuono_gradient_brick(DevParams P, FluidView F, const int *__restrict__ ord,
                             const int *__restrict__ bstart, const int *__restrict__ nel,
                             const int *__restrict__ nne, int *__restrict__ work,
                             const int *__restrict__ brow_gb, const int *__restrict__ brow_lds,
                             const int *__restrict__ bntot,
                             const unsigned short *__restrict__ nbr16,
                             const int *__restrict__ nbr, const int *__restrict__ cnt, int ls)
{
    __shared__ GradSmem<ND> S;
    const int tid = threadIdx.x;
    const int ne = *nne;
    const int cap = P.brick_cap;
    while (true) {
        if (tid == 0) S.item = atomicAdd(&work[2], 1);
        __syncthreads();
        const int item = S.item;
        if (item >= ne) break;
        const int b = nel[item];
        const int t0 = bstart[b], t1 = bstart[b + 1];
        const int ntot = bntot[b];
        const bool staged = ntot <= cap - 1;
        if (staged) {
            for (int r = tid; r <= P.maxr; r += kBrickThreads) {
                if (r < P.maxr) S.row_gb[r] = brow_gb[(size_t)b * P.maxr + r];
                S.row_lds[r] = brow_lds[(size_t)b * (P.maxr + 1) + r];
            }
            if (tid == 0) {  // the far record: far away, zero volume
                S.p0[cap - 1] = make_double2(kFar, kFar);
                S.p1[cap - 1] = make_double2(kFar, 1.0);
                S.vol[cap - 1] = 0.0;
            }
            __syncthreads();
            for (int i = tid; i < ntot; i += kBrickThreads) {
                const int r = bsearch_le(S.row_lds, P.maxr + 1, i);
                const int g = S.row_gb[r] + (i - S.row_lds[r]);
                const double4 a = F.rec[2 * (size_t)g];
                S.p0[i] = make_double2(a.x, a.y);
                S.p1[i] = make_double2(ND == 3 ? a.z : 0.0, a.w);
                S.vol[i] = F.vol[g].x;
            }
            __syncthreads();
        }
        const int q = tid & 3;
        for (int base = t0; base < t1; base += kBrickThreads / 4) {
            const int t = base + tid / 4;
            const bool active = t < t1;
            const int col = active ? t : t0;
            const int s = ord[col];
            const double4 a = F.rec[2 * (size_t)s];
            const double xa[3] = {a.x, a.y, ND == 3 ? a.z : 0.0};
            const int n = active ? cnt[col] : 0;
            GradSums gs;
            if (staged) {
                const ushort4 *l4 = reinterpret_cast<const ushort4 *>(nbr16);
                const int ng = (max(n, 0) + 3) >> 2;
                // rows clamped into the list rather than a select between a load and a
                // register (which made the compiler issue flat loads)
                ushort4 cur4 = l4[(size_t)min(q, max(ng - 1, 0)) * ls + col];
                for (int g = q; g < ng; g += 4) {
                    const ushort4 nxt4 = l4[(size_t)min(g + 4, ng - 1) * ls + col];
                    const unsigned short e4[4] = {cur4.x, cur4.y, cur4.z, cur4.w};
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int e = e4[k];
                        const double2 b0 = S.p0[e], b1 = S.p1[e];
                        const double p0 = xa[0] - b0.x, p1 = xa[1] - b0.y;
                        const double p2 = (ND == 3) ? xa[2] - b1.x : 0.0;
                        const double d2 = dist2<ND>(p0, p1, p2);
                        grad_add<ND>(gs, grad_vk(P, S.vol[e], d2), a.w - b1.y, p0, p1, p2);
                    }
                    cur4 = nxt4;
                }
            } else {
                grad_sums_global<ND>(P, F, gs, xa, a.w, n, nbr, col, ls, q);
            }
            grad_finish<ND>(P, gs, active && q == 0, s);
        }
        __syncthreads();  // the LDS is staged again for the next brick
    }
}

// drift!: du = v (semidiscretization.jl:561-571; delta_v = 0 without shifting)
// (vrows rows of v per particle: ND + 1, or ND under SummationDensity)
template <int ND>
__global__ void k_drift(int nf, int vrows, const double *__restrict__ v, double *__restrict__ du)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
#pragma unroll
    for (int d = 0; d < ND; d++) du[(size_t)i * ND + d] = v[(size_t)i * vrows + d];
}

// One 2N-storage stage of the partition (v, u) with slopes (dv, v): the drift slope is read
// before v is updated, as the integrator evaluates kick! and drift! at the same stage state.
// A block owns kBlock particles and walks their u and then their v entries as flat,
// contiguous ranges, so every load and store of a wavefront covers whole cache lines.
// Then, what k_gather<0> does for the next kick: the cell-sorted records of the new state
// (positions, velocity, density, pressure by the state equation, volume), the skin check
// and the reset of the brick work counter.  iperm[i] is the sorted index of particle i of
// the last rebuild.
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_ck_stage_gather(int *flag, int *work, DevParams P, int nf, double *__restrict__ v,
                      double *__restrict__ u, double *__restrict__ tv, double *__restrict__ tu,
                      const double *__restrict__ dv, double a, double b, double dt, int first,
                      const int *__restrict__ iperm, const double *__restrict__ m,
                      double *__restrict__ xs, double *__restrict__ ys, double *__restrict__ zs,
                      double4 *__restrict__ rec, double2 *__restrict__ vol,
                      const double *__restrict__ snap, const int *__restrict__ noverflow)
{
    P = with_sound(P);
    const int i0 = blockIdx.x * kBlock;
    const int np = min(kBlock, nf - i0);
    const size_t u0 = (size_t)i0 * ND, v0 = (size_t)i0 * (ND + 1);
    for (int e = threadIdx.x; e < np * ND; e += kBlock) {
        const int p = e / ND, d = e - p * ND;
        const double slope = v[v0 + (size_t)p * (ND + 1) + d];
        const size_t k = u0 + e;
        const double t = first ? dt * slope : fma(dt, slope, a * tu[k]);
        tu[k] = t;
        u[k] = fma(b, t, u[k]);
    }
    __syncthreads();  // the v entries read above are written below
    for (int e = threadIdx.x; e < np * (ND + 1); e += kBlock) {
        const size_t k = v0 + e;
        const double t = first ? dt * dv[k] : fma(dt, dv[k], a * tv[k]);
        tv[k] = t;
        v[k] = fma(b, t, v[k]);
    }
    __syncthreads();  // the block's new u and v are complete
    if (work && blockIdx.x == 0 && threadIdx.x == 0) work[0] = work[1] = work[2] = 0;
    const int i = i0 + (int)threadIdx.x;
    bool moved = false;  // left its skin
    if (i < nf) {
        const size_t nfs = (size_t)nf;
        const int s = iperm[i];
        double x[3] = {0.0, 0.0, 0.0}, vv[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < ND; d++) {
            x[d] = u[(size_t)i * ND + d];
            vv[d] = v[(size_t)i * (ND + 1) + d];
        }
        const double rho = v[(size_t)i * (ND + 1) + ND];
        const double mass = m[i];
        if (!noverflow || *noverflow != 0) {  // see k_gather
            xs[s] = x[0];
            ys[s] = x[1];
            if constexpr (ND == 3) zs[s] = x[2];
        }
        rec[2 * (size_t)s] = make_double4(x[0], x[1], x[2], rho);
        rec[2 * (size_t)s + 1] = make_double4(vv[0], vv[1], vv[2], state_equation(P, rho));
        vol[s] = make_double2(mass / rho, mass);
        double e2 = 0.0;
#pragma unroll
        for (int d = 0; d < ND; d++) {
            const double e = x[d] - snap[d * nfs + s];
            e2 = fma(e, e, e2);
        }
        moved = !(e2 <= P.half_skin2);
    }
    raise_rebuild_flag(flag, moved);
}

// ------------------------------------------------------------------------------------------
// StateEquationAdaptiveCole: update_speed_of_sound! (wcsph/system.jl:307-320) at the start
// of every kick.  k_vmax: per-block maxima of v.v over the fluid's current velocities (the
// caller-order v_ode); k_sound_speed: c = min(c_max, max(c_min, sqrt(max v.v) / Ma)) and the
// constants that depend on it, computed as tpsph_create computes them from a fixed c.
// ------------------------------------------------------------------------------------------
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_vmax(int nf, int vrows, const double *__restrict__ v, double *__restrict__ part)
{
    __shared__ double s[kBlock];
    double m = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < nf; i += gridDim.x * kBlock) {
        double d = v[(size_t)i * vrows] * v[(size_t)i * vrows];
#pragma unroll
        for (int k = 1; k < ND; k++) d += v[(size_t)i * vrows + k] * v[(size_t)i * vrows + k];
        m = fmax(m, d);
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kBlock)
    k_sound_speed(DevParams P, int nparts, const double *__restrict__ part, double *__restrict__ dyn,
                  double alpha, double delta)
{
    __shared__ double s[kBlock];
    double m = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kBlock) m = fmax(m, part[i]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double c = fmin(P.c_max, fmax(P.c_min, sqrt(s[0]) / P.mach));
        dyn[0] = c;
        dyn[1] = P.rho0 * (c * c) / P.gamma;
        dyn[2] = alpha * c * P.h;
        dyn[3] = P.use_dd ? 2.0 * delta * P.h * c : 0.0;
    }
}

// ------------------------------------------------------------------------------------------
// FP64 integrals of the fluid state (tpsph_integrals): per-block partial sums over a fixed
// assignment of particles to blocks and a fixed tree inside each block, then one block sums
// the partials in block order -- the result does not depend on scheduling.
// ------------------------------------------------------------------------------------------
constexpr int kIntegrals = 8;
constexpr int kIntegralBlocks = 1024;

template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_integrals(int nf, const double *__restrict__ v, const double *__restrict__ u,
                const double *__restrict__ m, DevParams P, double *__restrict__ part)
{
    __shared__ double s[kIntegrals][kBlock];
    double acc[kIntegrals] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < nf; i += gridDim.x * kBlock) {
        const double mi = m[i];
        const double rho = v[(size_t)i * (ND + 1) + ND];
        double v2 = 0.0, gx = 0.0;
#pragma unroll
        for (int d = 0; d < ND; d++) {
            const double vd = v[(size_t)i * (ND + 1) + d];
            acc[1 + d] += mi * vd;
            v2 += vd * vd;
            gx += P.g[d] * u[(size_t)i * ND + d];
        }
        acc[0] += mi;
        acc[4] += 0.5 * mi * v2;
        acc[5] -= mi * gx;
        acc[6] += mi / rho;
        acc[7] += rho;
    }
#pragma unroll
    for (int k = 0; k < kIntegrals; k++) s[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
#pragma unroll
            for (int k = 0; k < kIntegrals; k++) s[k][threadIdx.x] += s[k][threadIdx.x + o];
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < kIntegrals) part[(size_t)blockIdx.x * kIntegrals + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(kBlock) k_integrals_final(int nparts, const double *__restrict__ part,
                                                            double *__restrict__ out)
{
    if ((int)threadIdx.x >= kIntegrals) return;
    double acc = 0.0;
    for (int b = 0; b < nparts; b++) acc += part[(size_t)b * kIntegrals + threadIdx.x];
    out[threadIdx.x] = acc;
}

__global__ void k_rk_stage(int64_t n, double *__restrict__ y, double *__restrict__ tmp,
                           const double *__restrict__ k, double a, double b, double dt,
                           int first)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double t = first ? dt * k[i] : fma(dt, k[i], a * tmp[i]);
    tmp[i] = t;
    y[i] = fma(b, t, y[i]);
}

// y = a y + b x, the broadcasts of a low-storage Runge-Kutta step one at a time: b == 0 reads
// only y (y = a y), a == 0 only x (y = b x), both zero write zeros; otherwise fma(b, x, a y)
// (the rounding of muladd(b, x, a y)).
__global__ void k_axpby(int64_t n, double *__restrict__ y, double a, const double *x, double b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (b == 0.0)
        y[i] = a == 0.0 ? 0.0 : a * y[i];
    else if (a == 0.0)
        y[i] = b * x[i];
    else
        y[i] = fma(b, x[i], a * y[i]);
}

// ------------------------------------------------------------------------------------------
// neighbour CSR export (parity tests): rows = cell-sorted points (rx, ry, rz; row_perm to the
// caller's order), columns = cell-sorted points with cell offsets col_start
// ------------------------------------------------------------------------------------------
template <int ND, bool FILL>
__global__ void k_nbr(DevParams P, int nrows, const double *__restrict__ rx,
                      const double *__restrict__ ry, const double *__restrict__ rz,
                      const int *__restrict__ row_perm, const int *__restrict__ col_start,
                      const double *__restrict__ cx, const double *__restrict__ cy,
                      const double *__restrict__ cz, const int *__restrict__ col_perm,
                      double hs, int *__restrict__ counts, const int64_t *__restrict__ offsets,
                      int *__restrict__ idx)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nrows) return;
    const double xa = rx[s], ya = ry[s], za = (ND == 3) ? rz[s] : 0.0;
    Pencils<ND> it;
    it.init(P, xa, ya, za, hs);
    const int row = row_perm[s];
    int64_t o = FILL ? offsets[row] : 0;
    const int64_t b = o;
    int cnt = 0, jb, je;
    while (it.next(P, col_start, jb, je)) {
        for (int j = jb; j < je; j++) {
            const double d2 = dist2<ND>(xa - cx[j], ya - cy[j], (ND == 3) ? za - cz[j] : 0.0);
            if (d2 <= P.R2) {
                if (FILL) idx[o++] = col_perm[j];
                cnt++;
            }
        }
    }
    if (!FILL) {
        counts[row] = cnt;
        return;
    }
    for (int64_t k = b + 1; k < o; k++) {  // ascending columns
        int key = idx[k];
        int64_t j = k - 1;
        while (j >= b && idx[j] > key) {
            idx[j + 1] = idx[j];
            j--;
        }
        idx[j + 1] = key;
    }
}

// Pairs that the last kick's kernels evaluated (parity tests of the pair search): for every
// row particle, the candidates its kernel read -- the brick list (LDS indices mapped back to
// sorted indices through the brick's row tables, as the staging pass maps them), the int
// Verlet list, or the cell walk of an overflowed list or of the cell search -- that pass the
// kernel's exact test on the kick's positions: az2 <= d^2 <= R^2 for the fluid-fluid and
// fluid-wall sums (the pair terms), d^2 <= R^2 for the Adami sums of the walls.
struct EvalPairsArgs {
    int pair;           // TPSPH_PAIR_*
    int lists, bricks;  // the kick's pair search
    const int *ord, *bstart, *bntot, *brow_gb, *brow_lds;
    int nbricks;
    const unsigned short *nbr16;
    const int *nbr, *cnt, *wnbr, *wcnt;
    int ls;
    const int *anbr, *acnt;
    int lw;
    const int *w_perm;
};

template <int ND, bool FILL>
__global__ void k_eval_pairs(DevParams P, FluidView F, WallView Wv, EvalPairsArgs a,
                             int *__restrict__ counts, const int64_t *__restrict__ offsets,
                             int *__restrict__ idx)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool wall_rows = a.pair == TPSPH_PAIR_WALL_FLUID;
    const int nrows = wall_rows ? P.nw : P.nf;
    if (r >= nrows) return;
    // the row particle: sorted index s (fluid rows in the kernels' target order)
    const int s = (!wall_rows && a.lists) ? a.ord[r] : r;
    double xa[3];
    if (wall_rows) {
        const double4 w = Wv.pos[s];
        xa[0] = w.x, xa[1] = w.y, xa[2] = w.z;
    } else {
        const double4 f = F.rec[2 * (size_t)s];
        xa[0] = f.x, xa[1] = f.y, xa[2] = f.z;
    }
    const int row = wall_rows ? a.w_perm[s] : F.perm[s];
    const bool wall_cols = a.pair == TPSPH_PAIR_FLUID_WALL;
    const double lower = wall_rows ? -1.0 : P.az2;
    int64_t o = FILL ? offsets[row] : 0;
    const int64_t o0 = o;
    int n_hit = 0;
    auto consider = [&](int j) {
        double xb[3];
        if (wall_cols) {
            const double4 w = Wv.pos[j];
            xb[0] = w.x, xb[1] = w.y, xb[2] = w.z;
        } else {
            const double4 f = F.rec[2 * (size_t)j];
            xb[0] = f.x, xb[1] = f.y, xb[2] = f.z;
        }
        const double d2 = dist2<ND>(xa[0] - xb[0], xa[1] - xb[1], (ND == 3) ? xa[2] - xb[2] : 0.0);
        if (d2 <= P.R2 && d2 >= lower) {
            if (FILL) idx[o++] = wall_cols ? a.w_perm[j] : F.perm[j];
            n_hit++;
        }
    };
    auto walk = [&](const int *start, double hs) {
        Pencils<ND> it;
        it.init(P, xa[0], xa[1], xa[2], hs);
        int jb, je;
        while (it.next(P, start, jb, je))
            for (int j = jb; j < je; j++) consider(j);
    };
    if (!a.lists) {
        if (wall_cols)
            walk(Wv.start, P.hs_wall);
        else
            walk(F.start, P.hs_fluid);
    } else if (wall_rows) {
        const int n = a.acnt[s];
        if (n >= 0)
            for (int k = 0; k < n; k++) consider(a.anbr[(size_t)k * a.lw + s]);
        else
            walk(F.start, P.hs_fluid);
    } else if (wall_cols) {
        const int n = a.wcnt[r];
        if (n >= 0)
            for (int k = 0; k < n; k++) consider(a.wnbr[(size_t)k * a.ls + r]);
        else
            walk(Wv.start, P.hs_wall);
    } else {
        const int n = a.cnt[r];
        int b = -1;
        if (a.bricks) b = bsearch_le(a.bstart, a.nbricks, r);
        if (b >= 0 && a.bntot[b] <= P.brick_cap - 1) {
            // the brick list: four LDS indices per group, the far record pads the last group
            const int *gb = a.brow_gb + (size_t)b * P.maxr;
            const int *lds = a.brow_lds + (size_t)b * (P.maxr + 1);
            const int ng = (max(n, 0) + 3) >> 2;
            for (int g = 0; g < 4 * ng; g++) {
                const int e = a.nbr16[((size_t)(g >> 2) * a.ls + r) * 4 + (g & 3)];
                if (e == P.brick_cap - 1) continue;
                const int row_e = bsearch_le(lds, P.maxr + 1, e);
                consider(gb[row_e] + (e - lds[row_e]));
            }
        } else if (n >= 0) {
            for (int k = 0; k < n; k++) consider(a.nbr[(size_t)k * a.ls + r]);
        } else {
            walk(F.start, P.hs_fluid);
        }
    }
    if (!FILL) {
        counts[row] = n_hit;
        return;
    }
    for (int64_t k = o0 + 1; k < o; k++) {  // ascending columns
        const int key = idx[k];
        int64_t j = k - 1;
        while (j >= o0 && idx[j] > key) {
            idx[j + 1] = idx[j];
            j--;
        }
        idx[j + 1] = key;
    }
}

// cell-sorted SoA coordinates (and optionally {x, y, z, mass} records) of n points
template <int ND>
__global__ void k_positions(int n, const double *__restrict__ x, const int *__restrict__ perm,
                            double *__restrict__ xs, double *__restrict__ ys,
                            double *__restrict__ zs, double4 *__restrict__ pos,
                            const double *__restrict__ mass)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int i = perm[s];
    const double a = x[(size_t)i * ND], b = x[(size_t)i * ND + 1];
    const double c = (ND == 3) ? x[(size_t)i * ND + 2] : 0.0;
    xs[s] = a;
    ys[s] = b;
    if (zs) zs[s] = c;
    if (pos) pos[s] = make_double4(a, b, c, mass ? mass[i] : 0.0);
}

template <typename T>
__global__ void k_unpermute(int n, const int *__restrict__ perm, const T *__restrict__ sorted,
                            T *__restrict__ out)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[perm[s]] = sorted[s];
}

// tpsph_sort_particles: row j of the sorted state is row perm[j] of the caller's state (v
// with `vrows` doubles per particle, u with ND, the mass with one); po[j] = perm[j]
template <int ND>
__global__ void k_permute_state(int n, int vrows, const int *__restrict__ perm,
                                const double *__restrict__ v, const double *__restrict__ u,
                                const double *__restrict__ m, double *__restrict__ vo,
                                double *__restrict__ uo, double *__restrict__ mo,
                                int64_t *__restrict__ po)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int i = perm[j];
    for (int d = 0; d < vrows; d++) vo[(size_t)j * vrows + d] = v[(size_t)i * vrows + d];
#pragma unroll
    for (int d = 0; d < ND; d++) uo[(size_t)j * ND + d] = u[(size_t)i * ND + d];
    mo[j] = m[i];
    if (po) po[j] = i;
}

// fluid pressure (p = .w of the second double4 of each record, half = 1) or density (.w of
// the first, half = 0) of the cell-sorted records
__global__ void k_record_pressure(int n, const double4 *__restrict__ rec, double *__restrict__ p,
                                  int half = 1)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    p[s] = rec[2 * (size_t)s + half].w;
}


}  // namespace

namespace {



// Cell list of n points x (ND doubles per point, caller order) into start / perm.
template <int ND>
int build_cells(tpsph_handle *h, const int *flag, int n, const double *x, int *cell, int *rank,
                int *count, int *start, int *perm)
{
    // count is zero on entry (cleared at create, and again by k_scatter)
    const int nc = h->P.ncells;
    if (n > 0)
        hipLaunchKernelGGL((k_cells<ND>), dim3(gate_blocks(n)), dim3(kBlock), 0, h->stream, flag, h->P, n, x, cell, rank, count);
    TRY(launch_check(h, "k_cells"));
    TRY(scan(h, flag, count, start, nc + 1));
    if (n > 0) {
        hipLaunchKernelGGL((k_scatter), dim3(gate_blocks(n)), dim3(kBlock), 0, h->stream, flag, n, cell, rank, start, perm, count);
        hipLaunchKernelGGL((k_sort_cells), dim3(gate_blocks(nc)), dim3(kBlock), 0, h->stream, flag, nc, start, perm);
    }
    return launch_check(h, "k_scatter/k_sort_cells");
}

// Fluid cell list (rebuilt only when flagged) + cell-sorted gather of (v, u).
template <int ND>
int update_fluid(tpsph_handle *h, const double *v, const double *u, bool gathered,
                 bool keep_order = false)
{
    const int nf = h->nf;
    if (nf == 0) return TPSPH_OK;
    if (keep_order) {
        // the order and the lists stand: only the records of (v, u)
        hipLaunchKernelGGL((k_set_i32), dim3(1), dim3(1), 0, h->stream, h->flag, 0);
        hipLaunchKernelGGL((k_gather<ND, 0>), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream,
            h->flag, h->bwork, h->P, h->f_perm, u, v, h->f_mass, h->xs, h->ys, h->zs, h->rec, h->vol,
            h->snap, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
            h->search == TPSPH_SEARCH_LISTS ? h->counters + 3 : nullptr, nullptr, nullptr);
        return launch_check(h, "k_gather<0>");
    }
    if (!gathered) {  // else the stage kernel wrote the records of (v, u)
        TimedScope t(h, TPSPH_TIME_GATHER);
        hipLaunchKernelGGL((k_gather<ND, 0>), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream,
            h->flag, h->bwork, h->P, h->f_perm, u, v, h->f_mass, h->xs, h->ys, h->zs, h->rec, h->vol,
            h->snap, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
            h->search == TPSPH_SEARCH_LISTS ? h->counters + 3 : nullptr, nullptr, nullptr);
        TRY(launch_check(h, "k_gather<0>"));
    }
    {
        TimedScope t(h, TPSPH_TIME_CELL_LIST);
        TRY(build_cells<ND>(h, h->flag, nf, u, h->f_cell, h->f_rank, h->f_count, h->f_start,
                            h->f_perm));
        const bool lists = h->search == TPSPH_SEARCH_LISTS;
        // with lists the gather also counts the bricks (f_cell / f_rank are free again after
        // the cell list build) and zeroes the list-build counters
        hipLaunchKernelGGL((k_gather<ND, 1>), dim3(gate_blocks(nf)), dim3(kBlock), 0, h->stream,
            h->flag, nullptr, h->P, h->f_perm, u, v, h->f_mass, h->xs, h->ys, h->zs, h->rec, h->vol,
            h->snap, h->iperm, lists ? h->f_cell : nullptr, lists ? h->f_rank : nullptr,
            lists ? h->bcount : nullptr, lists ? h->counters : nullptr, 4, nullptr, h->nrebuild,
            lists ? h->xf : nullptr);
        if (lists) {
            TRY(scan(h, h->flag, h->bcount, h->bstart, h->nbricks + 1));
            hipLaunchKernelGGL((k_brick_scatter), dim3(gate_blocks(nf)), dim3(kBlock), 0, h->stream,
                h->flag, nf, h->f_cell, h->f_rank, h->bstart, h->ord, h->bcount);
            const int gu = h->bricks ? std::max(1, std::min(h->nbricks, kGateBlocks)) : 0;
            const int gw = h->nw > 0 ? nblocks(h->nw) : 0;
            if (gu + gw > 0)
                hipLaunchKernelGGL((k_brick_union_and_walls<ND>), dim3(gu + gw), dim3(kBlock), 0,
                    h->stream, h->flag, h->P, h->fluid_view(),
                    UnionArgs{h->f_start, h->bstart, h->nbricks, h->brow_gb, h->brow_lds, h->bntot,
                              h->nel, h->counters},
                    WallListsArgs{h->wpos, h->lw, h->cap_a, h->anbr, h->acnt, h->wactive,
                                  h->counters + 1, h->pw, h->rhow, h->wpv, h->wvel},
                    gu);
            hipLaunchKernelGGL((k_lists<ND>), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream,
                h->flag, h->P, h->fluid_view(), h->wall_view(),
                ListsArgs{h->ord, h->ls, h->cap_f, h->nbr, h->cnt, h->cap_w, h->wnbr, h->wcnt,
                          h->wtargets, h->counters + 2, h->counters + 3, h->wall_near,
                          h->brow_gb, h->brow_lds, h->bntot,
                          h->bricks ? reinterpret_cast<unsigned long long *>(h->nbr16) : nullptr});
            TRY(launch_check(h, "k_lists"));
        }
        // the fluid kernel of this kick lowers the flag
        TRY(launch_check(h, "k_gather<1>"));
    }
    return TPSPH_OK;
}

// Everything a kick does before the fluid pass: the sound speed (adaptive), the cell-sorted
// records of (v, u) with the fluid pressures, the rebuild of the cell list and the lists (if
// a particle left its skin), and the Adami pass of the walls.
template <int ND>
int prepare_impl(tpsph_handle *h, const double *v, const double *u, bool gathered,
                 bool keep_order = false)
{
    const int nf = h->nf, nw = h->nw;
    if (h->dyn && nf > 0) {
        // the sound speed of this kick comes from its velocities, before any pressure: the
        // records are then gathered again with it
        TimedScope t(h, TPSPH_TIME_GATHER);
        const int nb = gate_blocks(nf);
        hipLaunchKernelGGL((k_vmax<ND>), dim3(nb), dim3(kBlock), 0, h->stream, nf, h->P.vrows, v, h->vmax_part);
        hipLaunchKernelGGL((k_sound_speed), dim3(1), dim3(kBlock), 0, h->stream, h->P, nb,
                           h->vmax_part, h->dyn, h->params.alpha, h->params.delta);
        TRY(launch_check(h, "k_vmax/k_sound_speed"));
        gathered = false;
    }
    TRY(update_fluid<ND>(h, v, u, gathered, keep_order));
    if (h->P.summation && nf > 0) {
        // compute_density! before the state equation (wcsph/system.jl:291-301, 323-326)
        TimedScope t(h, TPSPH_TIME_GATHER);
        if (h->search == TPSPH_SEARCH_LISTS && h->bricks)
            hipLaunchKernelGGL((k_sum_density_brick<ND>),
                dim3(std::max(1, std::min(h->nbricks, h->n_cu))), dim3(kBrickThreads), 0,
                h->stream, h->P, h->fluid_view(), h->wall_view(), h->ord, h->bstart, h->nel,
                h->counters, h->bwork, h->brow_gb, h->brow_lds, h->bntot, h->nbr16, h->nbr,
                h->cnt, h->wnbr, h->wcnt, h->ls, h->sum_rho);
        else if (h->search == TPSPH_SEARCH_LISTS)
            hipLaunchKernelGGL((k_sum_density_lists<ND>), dim3(nblocks(4 * (int64_t)nf)), dim3(kBlock),
                               0, h->stream, h->P, h->fluid_view(), h->wall_view(), h->ord, h->nbr,
                               h->cnt, h->wnbr, h->wcnt, h->ls, h->sum_rho);
        else
            hipLaunchKernelGGL((k_sum_density<ND>), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream,
                               h->P, h->fluid_view(), h->wall_view(), h->wall_near, h->sum_rho);
        hipLaunchKernelGGL((k_sum_finish), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream, h->P,
                           nf, h->sum_rho, h->rec, h->vol);
        TRY(launch_check(h, "k_sum_density"));
    }
    if (nw > 0) {
        TimedScope t(h, TPSPH_TIME_WALL);
        if (h->anbr)
            hipLaunchKernelGGL((k_wall<ND, 4>), dim3(gate_blocks(4 * (int64_t)nw)), dim3(kBlock), 0,
                h->stream, h->P, h->fluid_view(), h->wpos, h->pw, h->rhow, h->wpv, h->anbr, h->acnt,
                h->lw, h->wactive, h->counters + 1, h->wvel);
        else
            hipLaunchKernelGGL((k_wall<ND, 1>), dim3(nblocks(nw)), dim3(kBlock), 0, h->stream, h->P,
                h->fluid_view(), h->wpos, h->pw, h->rhow, h->wpv, nullptr, nullptr, h->lw, nullptr,
                nullptr, h->wvel);
        TRY(launch_check(h, "k_wall"));
    }
    return TPSPH_OK;
}

template <int ND>
int kick_impl(tpsph_handle *h, double *dv, const double *v, const double *u, bool gathered)
{
    const int nf = h->nf, nw = h->nw;
    TRY(prepare_impl<ND>(h, v, u, gathered));
    if (nf > 0) {
        TimedScope t(h, TPSPH_TIME_FLUID);
        if (h->P.ddgrad) {
            const bool lists = h->search == TPSPH_SEARCH_LISTS;
            if (lists && h->bricks)
                hipLaunchKernelGGL((k_antuono_gradient_brick<ND>),
                    dim3(std::max(1, std::min(h->nbricks, h->n_cu))), dim3(kBrickThreads), 0,
                    h->stream, h->P, h->fluid_view(), h->ord, h->bstart, h->nel, h->counters,
                    h->bwork, h->brow_gb, h->brow_lds, h->bntot, h->nbr16, h->nbr, h->cnt, h->ls);
            else
                hipLaunchKernelGGL((k_antuono_gradient<ND>), dim3(nblocks(4 * (int64_t)nf)), dim3(kBlock), 0,
                                   h->stream, h->P, h->fluid_view(), lists ? h->ord : nullptr,
                                   lists ? h->nbr : nullptr, lists ? h->cnt : nullptr, h->ls);
            TRY(launch_check(h, "k_antuono_gradient"));
        }
        if (h->search == TPSPH_SEARCH_LISTS && h->bricks && h->mixed) {
            TRY(launch_fluid_brick_mixed(h, dv));
            hipLaunchKernelGGL((k_fluid_brick<ND, true>), dim3(std::max(1, std::min(h->nbricks, h->n_cu))),
                dim3(kBrickThreads), 0, h->stream, h->flag, h->P, h->fluid_view(), h->wall_view(),
                h->ord, h->bstart, h->nel, h->counters, h->bwork, h->brow_gb, h->brow_lds, h->bntot,
                h->nbr16, h->nbr, h->cnt, h->wnbr, h->wcnt, h->ls, dv, h->bfall);
            if (nw > 0)
                hipLaunchKernelGGL((k_fluid_walls<ND>), dim3(gate_blocks(4 * (int64_t)nf)), dim3(kBlock), 0, h->stream,
                    h->P, h->fluid_view(), h->wall_view(), h->ord, h->wtargets, h->counters + 2,
                    h->wnbr, h->wcnt, h->ls, dv);
        } else if (h->search == TPSPH_SEARCH_LISTS && h->bricks) {
            const int vm = visc_mode(h->P);
            const bool ant = h->P.ddgrad != nullptr, sum = h->P.summation != 0;
            const auto brick =
                ant   ? (vm == kViscNone    ? k_fluid_brick<ND, false, kViscNone, true>
                         : vm == kViscAlpha ? k_fluid_brick<ND, false, kViscAlpha, true>
                                            : k_fluid_brick<ND, false, kViscAlphaBeta, true>)
                : sum ? (vm == kViscNone    ? k_fluid_brick<ND, false, kViscNone, false, true>
                         : vm == kViscAlpha ? k_fluid_brick<ND, false, kViscAlpha, false, true>
                                            : k_fluid_brick<ND, false, kViscAlphaBeta, false, true>)
                      : (vm == kViscNone    ? k_fluid_brick<ND, false, kViscNone>
                         : vm == kViscAlpha ? k_fluid_brick<ND, false, kViscAlpha>
                                            : k_fluid_brick<ND, false, kViscAlphaBeta>);
            hipLaunchKernelGGL(brick, dim3(std::max(1, std::min(h->nbricks, h->n_cu))),
                dim3(kBrickThreads), 0, h->stream, h->flag, h->P, h->fluid_view(), h->wall_view(),
                h->ord, h->bstart, h->nel, h->counters, h->bwork, h->brow_gb, h->brow_lds, h->bntot,
                h->nbr16, h->nbr, h->cnt, h->wnbr, h->wcnt, h->ls, dv, nullptr);
            const auto walls = sum ? k_fluid_walls<ND, true> : k_fluid_walls<ND, false>;
            if (nw > 0)
                hipLaunchKernelGGL(walls, dim3(gate_blocks(4 * (int64_t)nf)), dim3(kBlock), 0, h->stream,
                    h->P, h->fluid_view(), h->wall_view(), h->ord, h->wtargets, h->counters + 2,
                    h->wnbr, h->wcnt, h->ls, dv);
        } else if (h->search == TPSPH_SEARCH_LISTS) {
            hipLaunchKernelGGL((k_fluid_list<ND>), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream,
                h->flag, h->P, h->fluid_view(), h->wall_view(), h->ord, h->nbr, h->cnt, h->wnbr, h->wcnt, h->ls, dv);
        } else if (h->P.summation) {
            hipLaunchKernelGGL((k_fluid<ND, true>), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream,
                h->flag, h->P, h->fluid_view(), h->wall_view(), h->wall_near, dv);
        } else {
            hipLaunchKernelGGL((k_fluid<ND>), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream, h->flag,
                h->P, h->fluid_view(), h->wall_view(), h->wall_near, dv);
        }
        TRY(launch_check(h, "k_fluid"));
    }
    h->have_kick = true;
    return TPSPH_OK;
}



}  // namespace

// The device parameters of params for nf fluid and nw wall particles on the grid of the box
// [lo, hi] (one padding cell per side): kernel and state-equation constants, cells, skin
// radii and bricks.  Failures go to the create error (tpsph_last_error(NULL)).
int tpsph::derive_params(const tpsph_params *params, int nf, int nw, const double *lo,
                  const double *hi, DevParams &P, int &nbricks)
{
    const int nd = params->ndims;
    // --- derived parameters
    const double hl = params->smoothing_length;
    const double R = 2.0 * hl;  // compact_support(WendlandC2, h)
    P.nd = nd;
    P.nf = nf;
    P.nw = nw;
    P.use_visc = params->use_viscosity ? 1 : 0;
    P.summation = params->density_calculator == TPSPH_DENSITY_SUMMATION ? 1 : 0;
    P.vrows = nd + (P.summation ? 0 : 1);
    // density diffusion acts through the continuity equation, which SummationDensity has not
    P.use_dd = (params->use_density_diffusion && !P.summation) ? 1 : 0;
    P.clip_fluid = params->clip_fluid_eos ? 1 : 0;
    P.clip_wall = params->clip_wall_pressure ? 1 : 0;
    P.h = hl;
    P.h_inv = 1.0 / hl;
    P.R = R;
    P.R2 = R * R;
    P.two_h = 2.0 * hl;
    P.az2 = std::nextafter(R * R, INFINITY) - R * R;  // sqrt(eps(R^2))^2
    const double sigma = nd == 2 ? (7.0 / (4.0 * M_PI)) * (P.h_inv * P.h_inv)
                                 : (21.0 / (16.0 * M_PI)) * (P.h_inv * P.h_inv * P.h_inv);
    P.kval_c = sigma;
    P.kgrad_c = -5.0 * sigma * (P.h_inv * P.h_inv);
    P.rho0 = params->reference_density;
    P.c = params->sound_speed;
    P.gamma = params->exponent;
    P.inv_gamma = 1.0 / params->exponent;
    P.p_bg = params->background_pressure;
    P.B = P.rho0 * (P.c * P.c) / P.gamma;
    P.alpha = params->alpha;
    P.beta = params->beta;
    P.eps_h2 = params->epsilon * (hl * hl);
    P.av_c = params->alpha * P.c * hl;
    P.dd2 = P.use_dd ? 2.0 * params->delta * hl * P.c : 0.0;
    P.ddgrad = nullptr;  // set below for DensityDiffusionAntuono
    P.eps_hh = std::nextafter(hl * hl, INFINITY) - hl * hl;
    for (int d = 0; d < 3; d++) P.g[d] = d < nd ? params->acceleration[d] : 0.0;
    P.p_offset = params->pressure_offset;
    P.dyn = nullptr;  // set below for an adaptive state equation
    P.visc_w = params->wall_viscosity;
    P.w_alpha = params->wall_alpha;
    P.w_ac = params->wall_alpha * P.c;
    P.w_beta = params->wall_beta;
    P.w_eps_h2 = params->wall_epsilon * hl * hl;
    P.w_nu = params->wall_nu;
    P.nu_kin = params->alpha * hl * P.c / (2 * nd + 4);
    P.mach = params->mach_number_target;
    P.c_min = params->min_sound_speed;
    P.c_max = params->max_sound_speed;

    const int div = params->cell_divisions > 0 ? params->cell_divisions : 2;
    const int xdiv = params->x_cell_divisions > 0 ? params->x_cell_divisions : 2 * div;
    const double skin = params->skin >= 0.0 ? params->skin : 0.04;
    double extent = R;
    for (int d = 0; d < nd; d++) extent = std::fmax(extent, std::fmax(std::fabs(lo[d]), std::fabs(hi[d])) + 4.0 * R);
    // cell-interval rounding margin: far above the error of floor((x - lo) / cs)
    const double margin = 1e-9 * R + 64.0 * DBL_EPSILON * extent;
    P.half_skin = 0.5 * skin * R;
    P.half_skin2 = P.half_skin * P.half_skin;
    P.hs_fluid = P.half_skin + margin;
    P.hs_wall = margin;
    {
        const double rlf = R + 2.0 * P.half_skin + margin, rlw = R + P.half_skin + margin;
        P.rl2_fluid = rlf * rlf;
        P.rl2_wall = rlw * rlw;
        P.hs_list_fluid = 2.0 * P.half_skin + 2.0 * margin;
        P.hs_list_wall = P.half_skin + 2.0 * margin;
    }
    // k cells span the reach of a point: R (as bounded by the FP32 square roots of the pencil
    // widths, which run at most 2e-6 R above it), its own drift and a neighbour's drift
    const double span = (R * 1.00001 + 2.0 * P.hs_fluid) * (1.0 + 1e-9);
    int64_t ncells = 1;
    for (int d = 0; d < 3; d++) {
        const double cs = span / (d == 0 ? xdiv : div);
        P.cs[d] = cs;
        P.inv_cs[d] = 1.0 / cs;
        if (d < nd) {
            P.lo[d] = lo[d] - cs;
            int64_t n = (int64_t)std::ceil((hi[d] + cs - P.lo[d]) * P.inv_cs[d]) + 1;
            if (n < 1) n = 1;
            if (n > (1 << 20))
                return fail(nullptr, TPSPH_ERR_ARGUMENT, "domain box too large for the cell grid");
            P.n[d] = (int)n;
        } else {
            P.lo[d] = 0.0;
            P.n[d] = 1;
        }
        ncells *= P.n[d];
    }
    if (ncells > (int64_t)INT32_MAX / 4)
        return fail(nullptr, TPSPH_ERR_ARGUMENT, "cell grid too large (domain / smoothing_length)");
    P.ncells = (int)ncells;
    // FP32 pre-test of the fluid list build.  A position inside the grid box is stored as
    // fl32(x - lo), off by at most 2^-24 L (L the box's longest side), so an FP32 distance d'
    // is off by at most 3.5 * 2^-24 L + 2^-24 d from the FP64 one d, and its square by a
    // further 3.1 * 2^-24 relative: with e = 4 * 2^-24 (L + rl) the bounds leave a margin of
    // at least 2^-24 rl on either side of the list radius rl
    {
        const double u = 0x1p-24;
        double L = 0.0;
        for (int d = 0; d < 3; d++) {
            P.box[d] = d < nd ? P.n[d] * P.cs[d] : 0.0;
            L = std::fmax(L, P.box[d]);
        }
        auto bounds = [&](double rl2, float &fi, float &fo) {
            const double rl = std::sqrt(rl2), e = 4.0 * u * (L + rl);
            const double in2 = rl > e ? (rl - e) * (rl - e) * (1.0 - 4.0 * u) : -1.0;
            const double out2 = (rl + e) * (rl + e) * (1.0 + 4.0 * u);
            fi = (float)in2;
            fo = (float)out2;
            if ((double)fi > in2) fi = std::nextafter(fi, -INFINITY);
            if ((double)fo < out2) fo = std::nextafter(fo, INFINITY);
        };
        bounds(P.rl2_fluid, P.list_in2, P.list_out2);
        bounds(P.rl2_wall, P.wlist_in2, P.wlist_out2);
    }
    {
        // bricks of 3 (3D) or 8 (2D) cell spans per side, about 3.1 R (8.3 R) wide, in 3D
        // shortened along x to 2.75 spans: at h = 1.2 dx a 3D brick holds up to 448 targets
        // and a union of at most about 2020 particles at rest, inside the LDS records of the
        // brick kernel (full 3-span bricks: 512 targets, and the unions of 5-10 % of the
        // bricks exceed the records and take the slower global path)
        // DensityDiffusionAntuono stages the density gradients too: in 3D its records hold
        // kBrickCapGrad particles, so its bricks are one cell shorter along y and z (unions of
        // about 5/6 the size, about 320 targets)
        const int per = nd == 3 ? 3 : 8;
        const bool grad = params->use_density_diffusion == TPSPH_DIFFUSION_ANTUONO && nd == 3;
        P.brick_cap = grad ? kBrickCapGrad : kBrickCap;
        int64_t nb = 1;
        for (int d = 0; d < 3; d++) {
            P.bc[d] = d < nd ? per * (d == 0 ? xdiv : div) : 1;
            if (nd == 3 && d == 0) P.bc[d] = std::max(1, P.bc[d] * 11 / 12);
            if (grad && d >= 1) P.bc[d] = std::max(1, P.bc[d] - 1);
            P.nb[d] = (P.n[d] + P.bc[d] - 1) / P.bc[d];
            nb *= P.nb[d];
        }
        nbricks = (int)nb;
        // union rows: cells within the list radius of the brick box (grown by the margin)
        const double rl = R + 2.0 * P.half_skin + margin;
        int rows = 1;
        for (int d = 0; d < 3; d++) {
            P.ur[d] = (d >= 1 && d < nd) ? (int)std::ceil(rl / P.cs[d]) + 1 : 0;
            if (d >= 1 && d < nd) rows *= P.bc[d] + 2 * P.ur[d];
        }
        P.maxr = rows;
        P.rl_union2 = (rl + margin) * (rl + margin);
        P.hs_union = margin;
    }
    return TPSPH_OK;
}

int tpsph::validate(const tpsph_params *p, int64_t nf, int64_t nw, const double *fm, const double *wx,
             const double *wm)
{
    auto bad = [](const std::string &m) { return fail(nullptr, TPSPH_ERR_ARGUMENT, m); };
    if (!p) return bad("params is NULL");
    if (p->ndims != 2 && p->ndims != 3) return bad("ndims must be 2 or 3");
    if (p->kernel != TPSPH_KERNEL_WENDLAND_C2) return bad("only WendlandC2Kernel is implemented");
    if (p->memory != TPSPH_MEM_DEVICE && p->memory != TPSPH_MEM_HOST)
        return bad("memory must be TPSPH_MEM_DEVICE or TPSPH_MEM_HOST");
    if (nf < 0 || nw < 0) return bad("particle counts must be non-negative");
    if (nf > (int64_t)1 << 26 || nw > (int64_t)1 << 26)
        return bad("particle counts above 2^26 per system exceed the 32-bit byte offsets of "
                   "this build");
    if (!(p->smoothing_length > 0.0)) return bad("`smoothing_length` must be positive");
    if (!(p->reference_density > DBL_EPSILON))
        return bad("`reference_density` needs to be positive and larger than eps()");
    if (!(p->sound_speed > 0.0)) return bad("`sound_speed` must be positive");
    if (p->wall_viscosity < 0 || p->wall_viscosity > 2)
        return bad("wall_viscosity must be 0 (nothing), 1 (ArtificialViscosityMonaghan) or 2 "
                   "(ViscosityAdami)");
    if (p->wall_viscosity && p->use_viscosity != 1)
        return bad("a wall viscosity needs the fluid's ArtificialViscosityMonaghan (its "
                   "kinematic viscosity enters the wall terms)");
    if (p->wall_viscosity && !(p->wall_epsilon > 0.0)) return bad("`wall_epsilon` must be positive");
    if (p->adaptive_sound_speed) {
        if (!(p->mach_number_target > 0.0)) return bad("`mach_number_target` must be positive");
        if (!(p->min_sound_speed > 0.0)) return bad("`min_sound_speed` must be positive");
        if (!(p->max_sound_speed >= p->min_sound_speed))
            return bad("`max_sound_speed` must not be below `min_sound_speed`");
    }
    if (!(p->exponent > 0.0)) return bad("`exponent` must be positive");
    if (p->cell_divisions < 0 || p->cell_divisions > 4) return bad("cell_divisions must be in 0..4");
    if (p->x_cell_divisions < 0 || p->x_cell_divisions > 8)
        return bad("x_cell_divisions must be in 0..8");
    if (p->pair_search < TPSPH_SEARCH_AUTO || p->pair_search > TPSPH_SEARCH_CELLS)
        return bad("pair_search must be one of enum tpsph_pair_search");
    if (p->use_density_diffusion < TPSPH_DIFFUSION_NONE ||
        p->use_density_diffusion > TPSPH_DIFFUSION_ANTUONO)
        return bad("use_density_diffusion must be one of enum tpsph_density_diffusion");
    if (!(p->skin < 1.0)) return bad("skin must be below 1 (a fraction of the compact support)");
    if (p->precision != TPSPH_PRECISION_FP64 && p->precision != TPSPH_PRECISION_MIXED)
        return bad("precision must be TPSPH_PRECISION_FP64 or TPSPH_PRECISION_MIXED");
    if (p->density_calculator != TPSPH_DENSITY_CONTINUITY &&
        p->density_calculator != TPSPH_DENSITY_SUMMATION)
        return bad("density_calculator must be one of enum tpsph_density_calculator");
    if (p->density_calculator == TPSPH_DENSITY_SUMMATION && p->precision != TPSPH_PRECISION_FP64)
        return bad("SummationDensity runs in FP64 (TPSPH_PRECISION_FP64)");
    if (p->precision == TPSPH_PRECISION_MIXED && p->use_viscosity && p->beta != 0.0)
        return bad("precision TPSPH_PRECISION_MIXED evaluates ArtificialViscosityMonaghan with "
                   "beta = 0 only");
    if (nf > 0 && !fm) return bad("fluid_mass is NULL");
    if (nw > 0 && (!wx || !wm)) return bad("wall_coordinates / wall_mass is NULL");
    return TPSPH_OK;
}

int tpsph::kick_dispatch(tpsph_handle *h, double *dv, const double *v, const double *u,
                         bool gathered)
{
    return h->nd == 2 ? kick_impl<2>(h, dv, v, u, gathered) : kick_impl<3>(h, dv, v, u, gathered);
}

int tpsph::prepare_dispatch(tpsph_handle *h, const double *v, const double *u, bool keep_order)
{
    return h->nd == 2 ? prepare_impl<2>(h, v, u, false, keep_order)
                      : prepare_impl<3>(h, v, u, false, keep_order);
}

const double tpsph::kCkA[5] = {0.0, -567301805773.0 / 1357537059087.0,
                               -2404267990393.0 / 2016746695238.0,
                               -3550918686646.0 / 2091501179385.0,
                               -1275806237668.0 / 842570457699.0};
const double tpsph::kCkB[5] = {1432997174477.0 / 9575080441755.0,
                               5161836677717.0 / 13612068292357.0,
                               1720146321549.0 / 2090206949498.0,
                               3134564353537.0 / 4481467310338.0,
                               2277821191437.0 / 14882151754819.0};

int tpsph::run_ck2n54(tpsph_handle *h, double *v, double *u, double dt, int64_t nsteps,
                      double *store_v, double *store_u)
{
    if (!h->sol_dv) {
        TRY(dalloc(h, &h->sol_dv, v_len(h)));
        TRY(dalloc(h, &h->sol_tv, v_len(h)));
        TRY(dalloc(h, &h->sol_tu, u_len(h)));
    }
    if (h->P.summation) {
        // SummationDensity: v is the velocity, so the drift slope of u is v itself; the stage
        // updates u (with the v of the stage state) before v, then the next kick gathers and
        // sums the densities of the new state
        const int64_t n = (int64_t)h->nf * h->nd;
        const unsigned grid = (unsigned)((n + kBlock - 1) / kBlock);
        if (store_v) return fail(h, TPSPH_ERR_ARGUMENT, "reverse mode needs ContinuityDensity");
        for (int64_t s = 0; s < nsteps; s++) {
            for (int i = 0; i < 5; i++) {
                TRY(kick_dispatch(h, h->sol_dv, v, u, false));
                TimedScope ts(h, TPSPH_TIME_STAGE);
                hipLaunchKernelGGL((k_rk_stage), dim3(grid), dim3(kBlock), 0, h->stream, n, u,
                                   h->sol_tu, v, kCkA[i], kCkB[i], dt, (int)(i == 0));
                hipLaunchKernelGGL((k_rk_stage), dim3(grid), dim3(kBlock), 0, h->stream, n, v,
                                   h->sol_tv, h->sol_dv, kCkA[i], kCkB[i], dt, (int)(i == 0));
                TRY(launch_check(h, "k_rk_stage"));
            }
        }
        h->have_kick = false;
        return TPSPH_OK;
    }
    // every stage kernel also writes the cell-sorted records of the new state, so the kicks
    // after the first skip the gather
    bool gathered = false;
    for (int64_t s = 0; s < nsteps; s++) {
        for (int i = 0; i < 5; i++) {
            if (store_v) {  // the state this stage's kick sees
                const size_t k = (size_t)(5 * s + i);
                HIP_TRY(h, hipMemcpyAsync(store_v + k * v_len(h), v, v_len(h) * sizeof(double),
                                          hipMemcpyDeviceToDevice, h->stream));
                HIP_TRY(h, hipMemcpyAsync(store_u + k * u_len(h), u, u_len(h) * sizeof(double),
                                          hipMemcpyDeviceToDevice, h->stream));
                // the stored states are what the caller keeps: the last kick and stage would
                // only advance (v, u) past them
                if (s == nsteps - 1 && i == 4) break;
            }
            TRY(kick_dispatch(h, h->sol_dv, v, u, gathered));
            TimedScope ts(h, TPSPH_TIME_STAGE);
            if (h->nd == 2)
                hipLaunchKernelGGL((k_ck_stage_gather<2>), dim3(nblocks(h->nf)), dim3(kBlock), 0, h->stream,
                    h->flag, h->bwork, h->P, h->nf, v, u, h->sol_tv, h->sol_tu, h->sol_dv, kCkA[i], kCkB[i], dt,
                    i == 0, h->iperm, h->f_mass, h->xs, h->ys, h->zs, h->rec, h->vol, h->snap,
                    h->search == TPSPH_SEARCH_LISTS ? h->counters + 3 : nullptr);
            else
                hipLaunchKernelGGL((k_ck_stage_gather<3>), dim3(nblocks(h->nf)), dim3(kBlock), 0, h->stream,
                    h->flag, h->bwork, h->P, h->nf, v, u, h->sol_tv, h->sol_tu, h->sol_dv, kCkA[i], kCkB[i], dt,
                    i == 0, h->iperm, h->f_mass, h->xs, h->ys, h->zs, h->rec, h->vol, h->snap,
                    h->search == TPSPH_SEARCH_LISTS ? h->counters + 3 : nullptr);
            TRY(launch_check(h, "k_ck_stage_gather"));
            gathered = true;
        }
    }
    // the state has advanced past the last kick: its pressures belong to no current state
    h->have_kick = false;
    return TPSPH_OK;
}

namespace {


// CSR of nrows rows in the caller's order from a count launch and a fill launch of the same
// kernel: launch(fill, counts, offsets, idx).  offsets / idx may be host or device memory.
template <typename Launch>
int emit_csr(tpsph_handle *h, int nrows, Launch launch, int64_t *offsets, int32_t *idx,
             int64_t capacity)
{
    int *counts = nullptr;
    int64_t *doff = nullptr;
    int *didx = nullptr;
    TRY(dalloc(h, &counts, (size_t)nrows));
    std::vector<int> hc((size_t)nrows);
    std::vector<int64_t> ho((size_t)nrows + 1, 0);
    int rc = TPSPH_OK;
    if (nrows > 0) {
        launch(false, counts, nullptr, nullptr);
        rc = launch_check(h, "pair count");
        if (rc == TPSPH_OK) {
            hipError_t e = hipMemcpyAsync(hc.data(), counts, (size_t)nrows * sizeof(int), hipMemcpyDeviceToHost, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) rc = fail(h, TPSPH_ERR_HIP, std::string("neighbour counts: ") + hipGetErrorString(e));
        }
    }
    for (int r = 0; r < nrows; r++) ho[(size_t)r + 1] = ho[(size_t)r] + hc[(size_t)r];
    const int64_t total = ho[(size_t)nrows];
    if (rc == TPSPH_OK) {
        hipError_t e = hipMemcpy(offsets, ho.data(), ho.size() * sizeof(int64_t), hipMemcpyDefault);
        if (e != hipSuccess) rc = fail(h, TPSPH_ERR_HIP, std::string("offsets: ") + hipGetErrorString(e));
    }
    if (rc == TPSPH_OK && idx != nullptr && capacity < total)
        rc = fail(h, TPSPH_ERR_CAPACITY, "idx capacity is smaller than the number of pairs");
    if (rc == TPSPH_OK && idx != nullptr && total > 0) {
        rc = dalloc(h, &doff, ho.size());
        if (rc == TPSPH_OK) rc = dalloc(h, &didx, (size_t)total);
        if (rc == TPSPH_OK) {
            hipError_t e = hipMemcpy(doff, ho.data(), ho.size() * sizeof(int64_t), hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = fail(h, TPSPH_ERR_HIP, "offsets upload failed");
        }
        if (rc == TPSPH_OK) {
            launch(true, nullptr, doff, didx);
            rc = launch_check(h, "pair fill");
        }
        if (rc == TPSPH_OK) {
            hipError_t e = hipMemcpyAsync(idx, didx, (size_t)total * sizeof(int), hipMemcpyDefault, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) rc = fail(h, TPSPH_ERR_HIP, std::string("neighbour lists: ") + hipGetErrorString(e));
        }
    }
    (void)hipFree(counts);
    (void)hipFree(doff);
    (void)hipFree(didx);
    return rc;
}

// ------------------------------------------------------------------------------------------
// Kick without allocation (tpsph_kick_noalloc): every intermediate array of one kick lives in
// a file-scope device array, `scratch`, laid out anew on each call.  Host code reaches the
// array through the address hipGetSymbolAddress returns on each device (the symbol itself is
// a host-side stand-in and never a device address).  Counters and histograms are cleared
// by kernels, so the call records no memset (it can be captured in a HIP graph).
// ------------------------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

__global__ void k_fill_i32(int *p, int value, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_fill_u8(unsigned char *p, unsigned char value, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_fill_f64(double *p, double value, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = value;
}

// device address of `scratch` on the current device, looked up once per device
int scratch_base(unsigned char **base)
{
    static std::mutex mu;
    static unsigned char *addr[kMaxDevices] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(nullptr, TPSPH_ERR_HIP, "tpsph_kick_noalloc: no current HIP device");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(nullptr, TPSPH_ERR_HIP,
                        std::string("tpsph_kick_noalloc: hipGetSymbolAddress(scratch): ") +
                            hipGetErrorString(e));
        addr[dev] = static_cast<unsigned char *>(p);
    }
    *base = addr[dev];
    return TPSPH_OK;
}

// Bump allocator over the scratch array (256-byte aligned regions).
struct ScratchLayout {
    unsigned char *base;
    size_t used = 0;
    template <typename T>
    T *take(size_t count)
    {
        T *p = reinterpret_cast<T *>(base ? base + used : nullptr);
        used += (std::max<size_t>(count, 1) * sizeof(T) + 255) / 256 * 256;
        return p;
    }
};

// The kick of kick_impl with the cell walk (TPSPH_SEARCH_CELLS), fluid and wall cell lists
// built from scratch, on a handle whose arrays all point into the scratch array.  With
// base == nullptr only the layout is measured (h->... stay null, nothing is launched).
template <int ND>
int noalloc_kick(tpsph_handle *h, unsigned char *base, size_t *bytes, const double *fluid_mass,
                 const double *wall_x, const double *wall_mass, double *dv, const double *v,
                 const double *u)
{
    const int nf = h->nf, nw = h->nw;
    const int nc = h->P.ncells;
    const int tiles = (nc + 1 + kScanTile - 1) / kScanTile;
    ScratchLayout L{base};
    h->flag = L.take<int>(1);
    h->scan_tiles = L.take<int>((size_t)tiles + 1);
    h->f_cell = L.take<int>(nf);
    h->f_rank = L.take<int>(nf);
    h->f_perm = L.take<int>(nf);
    h->iperm = L.take<int>(nf);
    h->f_count = L.take<int>((size_t)nc + 1);
    h->f_start = L.take<int>((size_t)nc + 1);
    h->xs = L.take<double>((size_t)nf + kPad);
    h->ys = L.take<double>((size_t)nf + kPad);
    h->zs = L.take<double>((size_t)nf + kPad);
    h->snap = L.take<double>((size_t)ND * nf);
    h->rec = L.take<double4>(2 * (size_t)nf);
    h->vol = L.take<double2>(nf);
    int *w_cell = L.take<int>(nw), *w_rank = L.take<int>(nw), *w_count = L.take<int>((size_t)nc + 1);
    h->w_start = L.take<int>((size_t)nc + 1);
    h->w_perm = L.take<int>(nw);
    h->wx = L.take<double>((size_t)nw + kPad);
    h->wy = L.take<double>((size_t)nw + kPad);
    h->wz = L.take<double>((size_t)nw + kPad);
    h->wpos = L.take<double4>(nw);
    h->wpv = L.take<double2>(nw);
    h->pw = L.take<double>(nw);
    h->rhow = L.take<double>(nw);
    h->wall_near = L.take<unsigned char>(nc);
    *bytes = L.used;
    if (!base) return TPSPH_OK;
    hipStream_t st = h->stream;
    const double nan = __builtin_nan("");
    // what must start at zero (or NaN): histograms, the wall-near flags, the SoA padding
    hipLaunchKernelGGL((k_fill_i32), dim3(gate_blocks(nc + 1)), dim3(kBlock), 0, st, h->f_count, 0, nc + 1);
    hipLaunchKernelGGL((k_fill_i32), dim3(gate_blocks(nc + 1)), dim3(kBlock), 0, st, w_count, 0, nc + 1);
    hipLaunchKernelGGL((k_fill_u8), dim3(gate_blocks(nc)), dim3(kBlock), 0, st, h->wall_near, (unsigned char)0, nc);
    for (double *a : {h->xs, h->ys, h->zs})
        hipLaunchKernelGGL((k_fill_f64), dim3(gate_blocks(nf + kPad)), dim3(kBlock), 0, st, a, nan, nf + kPad);
    for (double *a : {h->wx, h->wy, h->wz})
        hipLaunchKernelGGL((k_fill_f64), dim3(gate_blocks(nw + kPad)), dim3(kBlock), 0, st, a, nan, nw + kPad);
    hipLaunchKernelGGL((k_set_i32), dim3(1), dim3(1), 0, st, h->flag, 1);
    TRY(launch_check(h, "tpsph_kick_noalloc clearing"));
    // static walls: cell list, sorted positions and masses, cells near walls
    if (nw > 0) {
        TRY(build_cells<ND>(h, nullptr, nw, wall_x, w_cell, w_rank, w_count, h->w_start, h->w_perm));
        hipLaunchKernelGGL((k_positions<ND>), dim3(nblocks(nw)), dim3(kBlock), 0, st, nw, wall_x,
                           h->w_perm, h->wx, h->wy, ND == 3 ? h->wz : nullptr, h->wpos, wall_mass);
        hipLaunchKernelGGL((k_wall_near<ND>), dim3(nblocks(nw)), dim3(kBlock), 0, st, h->P, nw, h->wpos, h->wall_near);
        TRY(launch_check(h, "tpsph_kick_noalloc walls"));
    }
    h->f_mass = const_cast<double *>(fluid_mass);
    return kick_impl<ND>(h, dv, v, u, false);
}

}  // namespace

// ==========================================================================================
// C ABI
// ==========================================================================================
extern "C" {

int tpsph_abi_version(void) { return TPSPH_ABI_VERSION; }

const char *tpsph_last_error(const tpsph_handle *h)
{
    return h ? h->last_error.c_str() : g_create_error.c_str();
}

int tpsph_create(const tpsph_params *params, int64_t n_fluid, int64_t n_wall,
                 const double *fluid_mass, const double *wall_coordinates,
                 const double *wall_mass, void *stream, tpsph_handle **out)
{
    if (!out) return fail(nullptr, TPSPH_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    TRY(validate(params, n_fluid, n_wall, fluid_mass, wall_coordinates, wall_mass));
    tpsph_handle *h = new (std::nothrow) tpsph_handle();
    if (!h) return fail(nullptr, TPSPH_ERR_HIP, "out of host memory");
    h->params = *params;
    h->nd = params->ndims;
    h->nf = (int)n_fluid;
    h->nw = (int)n_wall;
    h->stream = (hipStream_t)stream;
    const int nd = h->nd;
    const bool host = is_host(h);
    const hipMemcpyKind in_kind = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;

    std::vector<double> wxh((size_t)nd * h->nw);
    if (h->nw > 0) {
        hipError_t e = hipMemcpy(wxh.data(), wall_coordinates, wxh.size() * sizeof(double),
                                 host ? hipMemcpyHostToHost : hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            std::string m = std::string("copy of wall_coordinates: ") + hipGetErrorString(e);
            delete h;
            return fail(nullptr, TPSPH_ERR_HIP, m);
        }
    }

    // --- grid: the domain box (or the wall box when unset), one padding cell per side
    double lo[3], hi[3];
    bool have_box = true;
    for (int d = 0; d < nd; d++)
        if (!(params->domain_min[d] < params->domain_max[d])) have_box = false;
    for (int d = 0; d < nd; d++) {
        lo[d] = have_box ? params->domain_min[d] : INFINITY;
        hi[d] = have_box ? params->domain_max[d] : -INFINITY;
    }
    for (int i = 0; i < h->nw; i++)
        for (int d = 0; d < nd; d++) {
            lo[d] = std::fmin(lo[d], wxh[(size_t)i * nd + d]);
            hi[d] = std::fmax(hi[d], wxh[(size_t)i * nd + d]);
        }
    for (int d = 0; d < nd; d++)
        if (!(lo[d] <= hi[d])) {
            lo[d] = 0.0;
            hi[d] = 0.0;
        }
    {
        const int rc = derive_params(params, h->nf, h->nw, lo, hi, h->P, h->nbricks);
        if (rc != TPSPH_OK) {
            delete h;
            return rc;
        }
    }
    DevParams &P = h->P;
    const int64_t ncells = P.ncells;
    h->search = params->pair_search;
    // SummationDensity on the Verlet lists needs the brick kernel (P.maxr within one scan
    // block); the automatic choice takes the cell walk otherwise
    const bool sum_lists_ok = !P.summation || P.maxr <= kBrickMaxRows;
    if (h->search == TPSPH_SEARCH_AUTO)
        h->search = (P.half_skin > 0.0 && sum_lists_ok) ? TPSPH_SEARCH_LISTS : TPSPH_SEARCH_CELLS;
    if (h->search == TPSPH_SEARCH_LISTS && !sum_lists_ok) {
        delete h;
        return fail(nullptr, TPSPH_ERR_ARGUMENT,
                    "SummationDensity with TPSPH_SEARCH_LISTS needs the brick kernel: the "
                    "default cell divisions, or TPSPH_SEARCH_CELLS");
    }
    if (h->search == TPSPH_SEARCH_LISTS) {
        // capacities: 1.5 x the expected count in the list ball (particle volume m / rho0)
        // plus 16; walls fill at most 7/8 (3D) or 3/4 (2D) of the ball around a corner
        std::vector<double> mh((size_t)std::max(h->nf, h->nw));
        auto mean = [&](const double *src, int n) {
            if (n == 0) return 0.0;
            if (hipMemcpy(mh.data(), src, (size_t)n * sizeof(double),
                          host ? hipMemcpyHostToHost : hipMemcpyDeviceToHost) != hipSuccess)
                return -1.0;
            double acc = 0.0;
            for (int i = 0; i < n; i++) acc += mh[i];
            return acc / n;
        };
        const double mf = mean(fluid_mass, h->nf), mw = mean(wall_mass, h->nw);
        if (mf < 0.0 || mw < 0.0) {
            delete h;
            return fail(nullptr, TPSPH_ERR_HIP, "copy of the masses failed");
        }
        auto ball = [&](double r) { return nd == 2 ? M_PI * r * r : 4.0 / 3.0 * M_PI * r * r * r; };
        auto cap = [&](double expected) {
            const double c = std::ceil(1.5 * expected) + 16.0;
            return (int)std::min(4096.0, std::ceil(c / 8.0) * 8.0);
        };
        const double rlf = std::sqrt(P.rl2_fluid), rlw = std::sqrt(P.rl2_wall);
        h->cap_f = cap(mf > 0.0 ? ball(rlf) / (mf / P.rho0) : 0.0);
        h->cap_w = h->nw > 0 ? cap(mw > 0.0 ? (nd == 2 ? 0.75 : 0.875) * ball(rlw) / (mw / P.rho0) : 0.0) : 0;
        h->ls = (h->nf + 63) / 64 * 64;
        // a wall particle of a flat wall sees fluid in at most half its ball
        h->cap_a = h->nw > 0 ? cap(mf > 0.0 ? 0.6 * ball(rlw) / (mf / P.rho0) : 0.0) : 0;
        h->lw = (h->nw + 63) / 64 * 64;
    }
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            h->n_cu = cus;
    }

    auto cleanup_fail = [&](int rc) {
        std::string m = h->last_error;
        tpsph_destroy(h);
        return fail(nullptr, rc, m);
    };
    int rc;
#define CTRY(expr) \
    if ((rc = (expr)) != TPSPH_OK) return cleanup_fail(rc)
    const int nf = h->nf, nw = h->nw;
    const int tiles = (int)((ncells + 1 + kScanTile - 1) / kScanTile);
    h->mixed = params->precision == TPSPH_PRECISION_MIXED;
    if (h->mixed && !(h->search == TPSPH_SEARCH_LISTS && P.maxr <= kBrickMaxRows &&
                      params->use_density_diffusion != TPSPH_DIFFUSION_ANTUONO)) {
        delete h;
        return fail(nullptr, TPSPH_ERR_ARGUMENT,
                    "precision TPSPH_PRECISION_MIXED runs on the brick kernel: it needs the "
                    "Verlet lists (skin > 0), the default brick sizes and no "
                    "DensityDiffusionAntuono");
    }
    CTRY(dalloc(h, &h->flag, 1));
    CTRY(dalloc(h, &h->nrebuild, 1));
    if (hipMemset(h->nrebuild, 0, sizeof(int)) != hipSuccess) {
        h->last_error = "clearing the rebuild count failed";
        return cleanup_fail(TPSPH_ERR_HIP);
    }
    if (params->wall_viscosity && nw > 0) CTRY(dalloc(h, &h->wvel, 3 * (size_t)nw));
    if (params->use_density_diffusion == TPSPH_DIFFUSION_ANTUONO && !P.summation && nf > 0) {
        CTRY(dalloc(h, &h->ddgrad, (size_t)nd * nf));
        P.ddgrad = h->ddgrad;
    }
    if (P.summation) CTRY(dalloc(h, &h->sum_rho, (size_t)std::max(nf, 1)));
    if (params->adaptive_sound_speed) {
        CTRY(dalloc(h, &h->dyn, 4));
        CTRY(dalloc(h, &h->vmax_part, kGateBlocks));
        // until the first kick: the constants of params.sound_speed
        const double init[4] = {P.c, P.B, P.av_c, P.dd2};
        if (hipMemcpy(h->dyn, init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess) {
            h->last_error = "copy of the sound speed failed";
            return cleanup_fail(TPSPH_ERR_HIP);
        }
        P.dyn = h->dyn;
    }
    CTRY(dalloc(h, &h->scan_tiles, (size_t)tiles + 1));
    CTRY(dalloc(h, &h->f_cell, nf));
    CTRY(dalloc(h, &h->f_rank, nf));
    CTRY(dalloc(h, &h->f_perm, nf));
    CTRY(dalloc(h, &h->f_count, (size_t)ncells + 1));
    CTRY(dalloc(h, &h->f_start, (size_t)ncells + 1));
    CTRY(dalloc(h, &h->f_mass, nf));
    CTRY(dalloc(h, &h->xs, (size_t)nf + kPad));
    CTRY(dalloc(h, &h->ys, (size_t)nf + kPad));
    CTRY(dalloc(h, &h->zs, (size_t)nf + kPad));
    if (h->search == TPSPH_SEARCH_LISTS) CTRY(dalloc(h, &h->xf, (size_t)nf + kPad));
    if (h->search == TPSPH_SEARCH_LISTS && nw > 0) CTRY(dalloc(h, &h->wxf, (size_t)nw + kPad));
    CTRY(dalloc(h, &h->snap, (size_t)nd * nf));
    CTRY(dalloc(h, &h->iperm, nf));
    CTRY(dalloc(h, &h->rec, 2 * (size_t)nf));
    CTRY(dalloc(h, &h->vol, nf));
    CTRY(dalloc(h, &h->w_start, (size_t)ncells + 1));
    CTRY(dalloc(h, &h->w_perm, nw));
    CTRY(dalloc(h, &h->wx, (size_t)nw + kPad));
    CTRY(dalloc(h, &h->wy, (size_t)nw + kPad));
    CTRY(dalloc(h, &h->wz, (size_t)nw + kPad));
    CTRY(dalloc(h, &h->wall_near, (size_t)ncells));
    if (h->search == TPSPH_SEARCH_LISTS) {
        CTRY(dalloc(h, &h->nbr, (size_t)h->cap_f * h->ls));
        CTRY(dalloc(h, &h->cnt, nf));
        CTRY(dalloc(h, &h->ord, nf));
        CTRY(dalloc(h, &h->counters, 4));
        // the brick kernel (with DensityDiffusionAntuono its records also carry the density
        // gradients) whenever every brick's union rows fit one scan block
        h->bricks = P.maxr <= kBrickMaxRows;
        if (h->bricks) {
            CTRY(dalloc(h, &h->brow_gb, (size_t)h->nbricks * P.maxr));
            CTRY(dalloc(h, &h->brow_lds, (size_t)h->nbricks * (P.maxr + 1)));
            CTRY(dalloc(h, &h->bntot, h->nbricks));
            CTRY(dalloc(h, &h->nel, h->nbricks));
            // [0] the brick kernels' work counter, [1] the bricks the mixed-precision pass
            // leaves to the FP64 pass (listed in bfall), [2] the work counter of the brick
            // gradient pass (DensityDiffusionAntuono); all cleared by the kick's gather or
            // the stage kernel
            CTRY(dalloc(h, &h->bwork, 3));
            if (h->mixed) CTRY(dalloc(h, &h->bfall, (size_t)h->nbricks));
            CTRY(dalloc(h, &h->nbr16, (size_t)h->cap_f * h->ls));
        }
        CTRY(dalloc(h, &h->bcount, (size_t)h->nbricks + 1));
        CTRY(dalloc(h, &h->bstart, (size_t)h->nbricks + 1));
        if (hipMemset(h->bcount, 0, ((size_t)h->nbricks + 1) * sizeof(int)) != hipSuccess) {
            h->last_error = "clearing the brick counts failed";
            return cleanup_fail(TPSPH_ERR_HIP);
        }
        if (nw > 0) {
            CTRY(dalloc(h, &h->wnbr, (size_t)h->cap_w * h->ls));
            CTRY(dalloc(h, &h->wcnt, nf));
            CTRY(dalloc(h, &h->anbr, (size_t)h->cap_a * h->lw));
            CTRY(dalloc(h, &h->acnt, nw));
            CTRY(dalloc(h, &h->wactive, nw));
            if (h->bricks) CTRY(dalloc(h, &h->wtargets, nf));
        }
    }
    // NaN (all bits set) in the coordinate arrays, including their padding; never zero so a
    // padded entry can not pass the radius test
    for (double *a : {h->xs, h->ys, h->zs})
        CTRY(hip_fill(h, a, (size_t)nf + kPad));
    if ((h->xf && hipMemset(h->xf, 0xff, ((size_t)nf + kPad) * sizeof(float4)) != hipSuccess) ||
        (h->wxf && hipMemset(h->wxf, 0xff, ((size_t)nw + kPad) * sizeof(float4)) != hipSuccess)) {
        h->last_error = "clearing the FP32 positions failed";
        return cleanup_fail(TPSPH_ERR_HIP);
    }
    for (double *a : {h->wx, h->wy, h->wz})
        CTRY(hip_fill(h, a, (size_t)nw + kPad));
    if (hipMemset(h->f_count, 0, ((size_t)ncells + 1) * sizeof(int)) != hipSuccess) {
        h->last_error = "clearing the cell counts failed";
        return cleanup_fail(TPSPH_ERR_HIP);
    }
    if (hipMemset(h->wall_near, 0, (size_t)ncells) != hipSuccess) {
        h->last_error = "clearing wall_near failed";
        return cleanup_fail(TPSPH_ERR_HIP);
    }
    CTRY(dalloc(h, &h->wpos, nw));
    CTRY(dalloc(h, &h->wpv, nw));
    CTRY(dalloc(h, &h->pw, nw));
    CTRY(dalloc(h, &h->rhow, nw));
    hipLaunchKernelGGL((k_set_i32), dim3(1), dim3(1), 0, h->stream, h->flag, 1);  // the first kick builds the fluid cells
    CTRY(launch_check(h, "k_set_i32"));
    if (nf > 0) {
        hipError_t e = hipMemcpy(h->f_mass, fluid_mass, (size_t)nf * sizeof(double), in_kind);
        if (e != hipSuccess) {
            h->last_error = std::string("copy of fluid_mass: ") + hipGetErrorString(e);
            return cleanup_fail(TPSPH_ERR_HIP);
        }
    }
    // static wall cell list
    if (nw > 0) {
        double *wx_d = nullptr, *wm_d = nullptr;
        int *wcell = nullptr, *wrank = nullptr, *wcount = nullptr;
        rc = dalloc(h, &wx_d, (size_t)nd * nw);
        if (rc == TPSPH_OK) rc = dalloc(h, &wm_d, nw);
        if (rc == TPSPH_OK) rc = dalloc(h, &wcell, nw);
        if (rc == TPSPH_OK) rc = dalloc(h, &wrank, nw);
        if (rc == TPSPH_OK) rc = dalloc(h, &wcount, (size_t)ncells + 1);
        if (rc == TPSPH_OK && hipMemset(wcount, 0, ((size_t)ncells + 1) * sizeof(int)) != hipSuccess)
            rc = fail(h, TPSPH_ERR_HIP, "clearing the wall cell counts failed");
        if (rc == TPSPH_OK) {
            hipError_t e1 = hipMemcpy(wx_d, wxh.data(), wxh.size() * sizeof(double), hipMemcpyHostToDevice);
            hipError_t e2 = hipMemcpy(wm_d, wall_mass, (size_t)nw * sizeof(double), in_kind);
            if (e1 != hipSuccess || e2 != hipSuccess) rc = fail(h, TPSPH_ERR_HIP, "copy of wall data failed");
        }
        if (rc == TPSPH_OK)
            rc = nd == 2 ? build_cells<2>(h, nullptr, nw, wx_d, wcell, wrank, wcount, h->w_start, h->w_perm)
                         : build_cells<3>(h, nullptr, nw, wx_d, wcell, wrank, wcount, h->w_start, h->w_perm);
        if (rc == TPSPH_OK) {
            if (nd == 2)
                hipLaunchKernelGGL((k_positions<2>), dim3(nblocks(nw)), dim3(kBlock), 0, h->stream, nw, wx_d, h->w_perm, h->wx, h->wy, nullptr, h->wpos, wm_d);
            else
                hipLaunchKernelGGL((k_positions<3>), dim3(nblocks(nw)), dim3(kBlock), 0, h->stream, nw, wx_d, h->w_perm, h->wx, h->wy, h->wz, h->wpos, wm_d);
            if (nd == 2)
                hipLaunchKernelGGL((k_wall_near<2>), dim3(nblocks(nw)), dim3(kBlock), 0, h->stream, h->P, nw, h->wpos, h->wall_near);
            else
                hipLaunchKernelGGL((k_wall_near<3>), dim3(nblocks(nw)), dim3(kBlock), 0, h->stream, h->P, nw, h->wpos, h->wall_near);
            if (h->wxf && nd == 2)
                hipLaunchKernelGGL((k_wall_f32<2>), dim3(nblocks(nw)), dim3(kBlock), 0, h->stream, h->P, nw, h->wpos, h->wxf);
            else if (h->wxf)
                hipLaunchKernelGGL((k_wall_f32<3>), dim3(nblocks(nw)), dim3(kBlock), 0, h->stream, h->P, nw, h->wpos, h->wxf);
            rc = launch_check(h, "wall setup");
        }
        hipError_t es = hipStreamSynchronize(h->stream);
        (void)hipFree(wx_d);
        (void)hipFree(wm_d);
        (void)hipFree(wcell);
        (void)hipFree(wrank);
        (void)hipFree(wcount);
        if (rc == TPSPH_OK && es != hipSuccess)
            rc = fail(h, TPSPH_ERR_HIP, std::string("wall setup: ") + hipGetErrorString(es));
        if (rc != TPSPH_OK) return cleanup_fail(rc);
    }
#undef CTRY
    *out = h;
    return TPSPH_OK;
}

int tpsph_kick(tpsph_handle *h, double *dv_ode, const double *v_ode, const double *u_ode,
               double t)
{
    (void)t;
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (h->nf > 0 && (!dv_ode || !v_ode || !u_ode))
        return fail(h, TPSPH_ERR_ARGUMENT, "dv_ode, v_ode and u_ode must not be NULL");
    if (!is_host(h)) return kick_dispatch(h, dv_ode, v_ode, u_ode);
    TRY(ensure_staging(h));
    HIP_TRY(h, hipMemcpyAsync(h->st_v, v_ode, v_len(h) * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipMemcpyAsync(h->st_u, u_ode, u_len(h) * sizeof(double), hipMemcpyHostToDevice, h->stream));
    TRY(kick_dispatch(h, h->st_dv, h->st_v, h->st_u));
    HIP_TRY(h, hipMemcpyAsync(dv_ode, h->st_dv, v_len(h) * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return TPSPH_OK;
}

int tpsph_drift(tpsph_handle *h, double *du_ode, const double *v_ode, const double *u_ode,
                double t)
{
    (void)t;
    (void)u_ode;
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (h->nf == 0) return TPSPH_OK;
    if (!du_ode || !v_ode) return fail(h, TPSPH_ERR_ARGUMENT, "du_ode and v_ode must not be NULL");
    if (!is_host(h)) {
        if (h->nd == 2)
            hipLaunchKernelGGL((k_drift<2>), dim3(nblocks(h->nf)), dim3(kBlock), 0, h->stream, h->nf, h->P.vrows, v_ode, du_ode);
        else
            hipLaunchKernelGGL((k_drift<3>), dim3(nblocks(h->nf)), dim3(kBlock), 0, h->stream, h->nf, h->P.vrows, v_ode, du_ode);
        return launch_check(h, "k_drift");
    }
    for (int i = 0; i < h->nf; i++)
        for (int d = 0; d < h->nd; d++)
            du_ode[(size_t)i * h->nd + d] = v_ode[(size_t)i * h->P.vrows + d];
    return TPSPH_OK;
}

int tpsph_get_pressures(tpsph_handle *h, double *fluid_pressure, double *wall_pressure,
                        double *wall_density)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (!h->have_kick) return fail(h, TPSPH_ERR_STATE, "no kick has been evaluated yet");
    const size_t n = (size_t)std::max(h->nf, h->nw);
    double *tmp = nullptr, *pf_sorted = nullptr;
    TRY(dalloc(h, &tmp, n));
    int rc = dalloc(h, &pf_sorted, (size_t)h->nf);
    auto emit = [&](double *dst, const double *sorted, const int *perm, int cnt) -> int {
        if (!dst || cnt == 0) return TPSPH_OK;
        hipLaunchKernelGGL((k_unpermute<double>), dim3(nblocks(cnt)), dim3(kBlock), 0, h->stream, cnt, perm, sorted, tmp);
        TRY(launch_check(h, "k_unpermute"));
        HIP_TRY(h, hipMemcpyAsync(dst, tmp, (size_t)cnt * sizeof(double), hipMemcpyDefault, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        return TPSPH_OK;
    };
    if (rc == TPSPH_OK && fluid_pressure && h->nf > 0) {
        hipLaunchKernelGGL((k_record_pressure), dim3(nblocks(h->nf)), dim3(kBlock), 0, h->stream, h->nf, h->rec, pf_sorted);
        rc = launch_check(h, "k_record_pressure");
        if (rc == TPSPH_OK) rc = emit(fluid_pressure, pf_sorted, h->f_perm, h->nf);
    }
    if (rc == TPSPH_OK) rc = emit(wall_pressure, h->pw, h->w_perm, h->nw);
    if (rc == TPSPH_OK) rc = emit(wall_density, h->rhow, h->w_perm, h->nw);
    (void)hipFree(tmp);
    (void)hipFree(pf_sorted);
    return rc;
}

int tpsph_get_density(tpsph_handle *h, double *fluid_density)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (!fluid_density) return fail(h, TPSPH_ERR_ARGUMENT, "fluid_density is NULL");
    if (!h->have_kick) return fail(h, TPSPH_ERR_STATE, "no kick has been evaluated yet");
    if (h->nf == 0) return TPSPH_OK;
    double *sorted = nullptr, *tmp = nullptr;
    TRY(dalloc(h, &sorted, (size_t)h->nf));
    int rc = dalloc(h, &tmp, (size_t)h->nf);
    if (rc == TPSPH_OK) {
        hipLaunchKernelGGL((k_record_pressure), dim3(nblocks(h->nf)), dim3(kBlock), 0, h->stream,
                           h->nf, h->rec, sorted, 0);
        hipLaunchKernelGGL((k_unpermute<double>), dim3(nblocks(h->nf)), dim3(kBlock), 0, h->stream,
                           h->nf, h->f_perm, sorted, tmp);
        rc = launch_check(h, "k_record_pressure/k_unpermute");
    }
    if (rc == TPSPH_OK &&
        (hipMemcpyAsync(fluid_density, tmp, (size_t)h->nf * sizeof(double), hipMemcpyDefault,
                        h->stream) != hipSuccess ||
         hipStreamSynchronize(h->stream) != hipSuccess))
        rc = fail(h, TPSPH_ERR_HIP, "reading the densities failed");
    (void)hipFree(sorted);
    (void)hipFree(tmp);
    return rc;
}

int tpsph_get_sound_speed(tpsph_handle *h, double *sound_speed)
{
    if (!h || !sound_speed) return fail(h, TPSPH_ERR_ARGUMENT, "handle or sound_speed is NULL");
    if (!h->dyn) {
        *sound_speed = h->P.c;
        return TPSPH_OK;
    }
    if (hipMemcpyAsync(sound_speed, h->dyn, sizeof(double), hipMemcpyDeviceToHost, h->stream) !=
            hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return fail(h, TPSPH_ERR_HIP, "reading the sound speed failed");
    return TPSPH_OK;
}

int tpsph_neighbors_csr(tpsph_handle *h, const double *u_ode, int32_t pair, int64_t *offsets,
                        int32_t *idx, int64_t capacity)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (pair < 0 || pair > 2) return fail(h, TPSPH_ERR_ARGUMENT, "unknown pair kind");
    if (!offsets) return fail(h, TPSPH_ERR_ARGUMENT, "offsets must not be NULL");
    if (h->nf > 0 && !u_ode) return fail(h, TPSPH_ERR_ARGUMENT, "u_ode must not be NULL");
    const int nd = h->nd, nf = h->nf, nw = h->nw;
    const double *u = u_ode;
    if (is_host(h)) {
        TRY(ensure_staging(h));
        HIP_TRY(h, hipMemcpyAsync(h->st_u, u_ode, u_len(h) * sizeof(double), hipMemcpyHostToDevice, h->stream));
        u = h->st_u;
    }
    // fluid cell list of the current positions; the flag stays raised, so the next kick
    // rebuilds and gathers its own state again
    hipLaunchKernelGGL((k_set_i32), dim3(1), dim3(1), 0, h->stream, h->flag, 1);
    TRY(nd == 2 ? build_cells<2>(h, h->flag, nf, u, h->f_cell, h->f_rank, h->f_count, h->f_start, h->f_perm)
                : build_cells<3>(h, h->flag, nf, u, h->f_cell, h->f_rank, h->f_count, h->f_start, h->f_perm));
    if (nf > 0) {
        if (nd == 2)
            hipLaunchKernelGGL((k_positions<2>), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream, nf, u, h->f_perm, h->xs, h->ys, nullptr, nullptr, nullptr);
        else
            hipLaunchKernelGGL((k_positions<3>), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream, nf, u, h->f_perm, h->xs, h->ys, h->zs, nullptr, nullptr);
        TRY(launch_check(h, "k_positions"));
    }
    h->have_kick = false;
    int nrows;
    const double *rx, *ry, *rz, *cx, *cy, *cz;
    const int *row_perm, *col_start, *col_perm;
    double hs;
    if (pair == TPSPH_PAIR_FLUID_FLUID) {
        nrows = nf; rx = h->xs; ry = h->ys; rz = h->zs; row_perm = h->f_perm;
        col_start = h->f_start; cx = h->xs; cy = h->ys; cz = h->zs; col_perm = h->f_perm;
        hs = h->P.hs_fluid;
    } else if (pair == TPSPH_PAIR_FLUID_WALL) {
        nrows = nf; rx = h->xs; ry = h->ys; rz = h->zs; row_perm = h->f_perm;
        col_start = h->w_start; cx = h->wx; cy = h->wy; cz = h->wz; col_perm = h->w_perm;
        hs = h->P.hs_wall;
    } else {
        nrows = nw; rx = h->wx; ry = h->wy; rz = h->wz; row_perm = h->w_perm;
        col_start = h->f_start; cx = h->xs; cy = h->ys; cz = h->zs; col_perm = h->f_perm;
        hs = h->P.hs_fluid;
    }
    auto launch = [&](bool fill, int *counts, const int64_t *doff, int *didx) {
        const dim3 g(nblocks(nrows)), b(kBlock);
        if (nd == 2 && !fill)
            hipLaunchKernelGGL((k_nbr<2, false>), g, b, 0, h->stream, h->P, nrows, rx, ry, rz, row_perm, col_start, cx, cy, cz, col_perm, hs, counts, nullptr, nullptr);
        else if (nd == 2)
            hipLaunchKernelGGL((k_nbr<2, true>), g, b, 0, h->stream, h->P, nrows, rx, ry, rz, row_perm, col_start, cx, cy, cz, col_perm, hs, nullptr, doff, didx);
        else if (!fill)
            hipLaunchKernelGGL((k_nbr<3, false>), g, b, 0, h->stream, h->P, nrows, rx, ry, rz, row_perm, col_start, cx, cy, cz, col_perm, hs, counts, nullptr, nullptr);
        else
            hipLaunchKernelGGL((k_nbr<3, true>), g, b, 0, h->stream, h->P, nrows, rx, ry, rz, row_perm, col_start, cx, cy, cz, col_perm, hs, nullptr, doff, didx);
    };
    return emit_csr(h, nrows, launch, offsets, idx, capacity);
}

int tpsph_kick_pairs_csr(tpsph_handle *h, int32_t pair, int64_t *offsets, int32_t *idx,
                         int64_t capacity)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (pair < 0 || pair > 2) return fail(h, TPSPH_ERR_ARGUMENT, "unknown pair kind");
    if (!offsets) return fail(h, TPSPH_ERR_ARGUMENT, "offsets must not be NULL");
    if (!h->have_kick) return fail(h, TPSPH_ERR_STATE, "no kick has been evaluated yet");
    const bool lists = h->search == TPSPH_SEARCH_LISTS;
    const EvalPairsArgs a{pair, lists ? 1 : 0, h->bricks ? 1 : 0, h->ord, h->bstart, h->bntot,
                          h->brow_gb, h->brow_lds, h->nbricks, h->nbr16, h->nbr, h->cnt,
                          h->wnbr, h->wcnt, h->ls, h->anbr, h->acnt, h->lw, h->w_perm};
    const int nrows = pair == TPSPH_PAIR_WALL_FLUID ? h->nw : h->nf;
    if ((pair != TPSPH_PAIR_FLUID_FLUID && h->nw == 0) || h->nf == 0) {
        // no pairs: all offsets zero
        std::vector<int64_t> zero((size_t)nrows + 1, 0);
        HIP_TRY(h, hipMemcpy(offsets, zero.data(), zero.size() * sizeof(int64_t), hipMemcpyDefault));
        return TPSPH_OK;
    }
    auto launch = [&](bool fill, int *counts, const int64_t *doff, int *didx) {
        const dim3 g(nblocks(nrows)), b(kBlock);
        const FluidView F = h->fluid_view();
        const WallView W = h->wall_view();
        if (h->nd == 2 && !fill)
            hipLaunchKernelGGL((k_eval_pairs<2, false>), g, b, 0, h->stream, h->P, F, W, a, counts, nullptr, nullptr);
        else if (h->nd == 2)
            hipLaunchKernelGGL((k_eval_pairs<2, true>), g, b, 0, h->stream, h->P, F, W, a, nullptr, doff, didx);
        else if (!fill)
            hipLaunchKernelGGL((k_eval_pairs<3, false>), g, b, 0, h->stream, h->P, F, W, a, counts, nullptr, nullptr);
        else
            hipLaunchKernelGGL((k_eval_pairs<3, true>), g, b, 0, h->stream, h->P, F, W, a, nullptr, doff, didx);
    };
    return emit_csr(h, nrows, launch, offsets, idx, capacity);
}

int tpsph_axpby(tpsph_handle *h, int64_t n, double *y, double a, const double *x, double b)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (n < 0) return fail(h, TPSPH_ERR_ARGUMENT, "n must be non-negative");
    if (n == 0) return TPSPH_OK;
    if (!y || (b != 0.0 && !x)) return fail(h, TPSPH_ERR_ARGUMENT, "y (and x when b != 0) must not be NULL");
    if (is_host(h)) {
        for (int64_t i = 0; i < n; i++) {
            if (b == 0.0)
                y[i] = a == 0.0 ? 0.0 : a * y[i];
            else if (a == 0.0)
                y[i] = b * x[i];
            else
                y[i] = std::fma(b, x[i], a * y[i]);
        }
        return TPSPH_OK;
    }
    TimedScope ts(h, TPSPH_TIME_STAGE);
    hipLaunchKernelGGL((k_axpby), dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, h->stream, n, y, a, x, b);
    return launch_check(h, "k_axpby");
}

int tpsph_rk_stage(tpsph_handle *h, double *y, double *tmp, const double *k, double a,
                   double b, double dt, int64_t n, int32_t first_stage)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (n < 0) return fail(h, TPSPH_ERR_ARGUMENT, "n must be non-negative");
    if (n == 0) return TPSPH_OK;
    if (!y || !tmp || !k) return fail(h, TPSPH_ERR_ARGUMENT, "y, tmp and k must not be NULL");
    if (is_host(h)) {
        for (int64_t i = 0; i < n; i++) {
            double t = first_stage ? dt * k[i] : std::fma(dt, k[i], a * tmp[i]);
            tmp[i] = t;
            y[i] = std::fma(b, t, y[i]);
        }
        return TPSPH_OK;
    }
    TimedScope ts(h, TPSPH_TIME_STAGE);
    hipLaunchKernelGGL((k_rk_stage), dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, h->stream, n, y, tmp, k, a, b, dt, first_stage);
    return launch_check(h, "k_rk_stage");
}

int tpsph_solve_ck2n54(tpsph_handle *h, double *v_ode, double *u_ode, double t0, double dt,
                       int64_t nsteps)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (nsteps < 0) return fail(h, TPSPH_ERR_ARGUMENT, "nsteps must be non-negative");
    if (h->nf == 0 || nsteps == 0) return TPSPH_OK;
    if (!v_ode || !u_ode) return fail(h, TPSPH_ERR_ARGUMENT, "v_ode and u_ode must not be NULL");
    (void)t0;
    double *v = v_ode, *u = u_ode;
    if (is_host(h)) {
        TRY(ensure_staging(h));
        HIP_TRY(h, hipMemcpyAsync(h->st_v, v_ode, v_len(h) * sizeof(double), hipMemcpyHostToDevice, h->stream));
        HIP_TRY(h, hipMemcpyAsync(h->st_u, u_ode, u_len(h) * sizeof(double), hipMemcpyHostToDevice, h->stream));
        v = h->st_v;
        u = h->st_u;
    }
    TRY(run_ck2n54(h, v, u, dt, nsteps, nullptr, nullptr));
    if (is_host(h)) {
        HIP_TRY(h, hipMemcpyAsync(v_ode, v, v_len(h) * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(h, hipMemcpyAsync(u_ode, u, u_len(h) * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
    }
    return TPSPH_OK;
}

int tpsph_kick_noalloc(const tpsph_params *params, int64_t n_fluid, int64_t n_wall,
                       const double *fluid_mass, const double *wall_coordinates,
                       const double *wall_mass, double *dv_ode, const double *v_ode,
                       const double *u_ode, double t, void *stream)
{
    (void)t;
    TRY(validate(params, n_fluid, n_wall, fluid_mass, wall_coordinates, wall_mass));
    auto bad = [](const std::string &m) { return fail(nullptr, TPSPH_ERR_ARGUMENT, "tpsph_kick_noalloc: " + m); };
    if (params->memory != TPSPH_MEM_DEVICE) return bad("needs device pointers (TPSPH_MEM_DEVICE)");
    if (params->adaptive_sound_speed || params->use_density_diffusion == TPSPH_DIFFUSION_ANTUONO ||
        params->wall_viscosity || params->precision != TPSPH_PRECISION_FP64 ||
        params->density_calculator != TPSPH_DENSITY_CONTINUITY)
        return bad("supports ContinuityDensity, StateEquationCole, Molteni-Colagrossi diffusion, "
                   "walls without viscosity and FP64");
    if (n_fluid > 0 && (!dv_ode || !v_ode || !u_ode))
        return bad("dv_ode, v_ode and u_ode must not be NULL");
    double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < params->ndims; d++) {
        if (!(params->domain_min[d] < params->domain_max[d]))
            return bad("needs the grid box (domain_min < domain_max on every axis)");
        lo[d] = params->domain_min[d];
        hi[d] = params->domain_max[d];
    }
    // a handle on the stack whose arrays point into the scratch array; it owns nothing
    tpsph_handle h;
    h.params = *params;
    h.nd = params->ndims;
    h.nf = (int)n_fluid;
    h.nw = (int)n_wall;
    h.stream = (hipStream_t)stream;
    h.search = TPSPH_SEARCH_CELLS;
    TRY(derive_params(params, h.nf, h.nw, lo, hi, h.P, h.nbricks));
    size_t bytes = 0;
    TRY(h.nd == 2 ? noalloc_kick<2>(&h, nullptr, &bytes, fluid_mass, wall_coordinates, wall_mass, dv_ode, v_ode, u_ode)
                  : noalloc_kick<3>(&h, nullptr, &bytes, fluid_mass, wall_coordinates, wall_mass, dv_ode, v_ode, u_ode));
    if (bytes > kScratchBytes)
        return bad("needs " + std::to_string(bytes) + " bytes of scratch for " +
                   std::to_string(n_fluid) + " fluid, " + std::to_string(n_wall) +
                   " wall particles and " + std::to_string(h.P.ncells) +
                   " cells; the limit is " + std::to_string(kScratchBytes) + " bytes");
    unsigned char *base = nullptr;
    TRY(scratch_base(&base));
    const int rc = h.nd == 2 ? noalloc_kick<2>(&h, base, &bytes, fluid_mass, wall_coordinates, wall_mass, dv_ode, v_ode, u_ode)
                             : noalloc_kick<3>(&h, base, &bytes, fluid_mass, wall_coordinates, wall_mass, dv_ode, v_ode, u_ode);
    if (rc != TPSPH_OK) return fail(nullptr, rc, h.last_error);
    return TPSPH_OK;
}

int tpsph_integrals(tpsph_handle *h, const double *v_ode, const double *u_ode, double *out)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (!out) return fail(h, TPSPH_ERR_ARGUMENT, "out must not be NULL");
    if (h->nf > 0 && (!v_ode || !u_ode)) return fail(h, TPSPH_ERR_ARGUMENT, "v_ode and u_ode must not be NULL");
    if (h->P.summation)
        return fail(h, TPSPH_ERR_ARGUMENT, "tpsph_integrals reads the density rows of v_ode "
                                           "(ContinuityDensity)");
    const double *v = v_ode, *u = u_ode;
    if (is_host(h) && h->nf > 0) {
        TRY(ensure_staging(h));
        HIP_TRY(h, hipMemcpyAsync(h->st_v, v_ode, v_len(h) * sizeof(double), hipMemcpyHostToDevice, h->stream));
        HIP_TRY(h, hipMemcpyAsync(h->st_u, u_ode, u_len(h) * sizeof(double), hipMemcpyHostToDevice, h->stream));
        v = h->st_v;
        u = h->st_u;
    }
    if (!h->integ) TRY(dalloc(h, &h->integ, (size_t)(kIntegralBlocks + 1) * kIntegrals));
    const int nb = std::max(1, std::min(nblocks(h->nf), kIntegralBlocks));
    if (h->nd == 2)
        hipLaunchKernelGGL((k_integrals<2>), dim3(nb), dim3(kBlock), 0, h->stream, h->nf, v, u, h->f_mass, h->P, h->integ);
    else
        hipLaunchKernelGGL((k_integrals<3>), dim3(nb), dim3(kBlock), 0, h->stream, h->nf, v, u, h->f_mass, h->P, h->integ);
    double *res = h->integ + (size_t)kIntegralBlocks * kIntegrals;
    hipLaunchKernelGGL((k_integrals_final), dim3(1), dim3(kBlock), 0, h->stream, nb, h->integ, res);
    TRY(launch_check(h, "k_integrals"));
    HIP_TRY(h, hipMemcpyAsync(out, res, kIntegrals * sizeof(double), hipMemcpyDefault, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return TPSPH_OK;
}

int tpsph_rebuild_count(tpsph_handle *h, int64_t *count)
{
    if (!h || !count) return fail(h, TPSPH_ERR_ARGUMENT, "handle or count is NULL");
    int n = 0;
    if (h->nrebuild) {
        HIP_TRY(h, hipMemcpyAsync(&n, h->nrebuild, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
    }
    *count = n;
    return TPSPH_OK;
}

int tpsph_sort_particles(tpsph_handle *h, double *v_ode, double *u_ode, int64_t *perm)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    const int nf = h->nf, nd = h->nd;
    if (nf == 0) return TPSPH_OK;
    if (!v_ode || !u_ode) return fail(h, TPSPH_ERR_ARGUMENT, "v_ode and u_ode must not be NULL");
    TRY(ensure_staging(h));
    const size_t vl = v_len(h), ul = u_len(h);
    const bool host = is_host(h);
    // device copies of the state before the sort (st_v, st_u) and of the masses (st_dv);
    // host mode writes the sorted state to temporaries and copies it back
    double *vo = v_ode, *uo = u_ode, *tmp = nullptr;
    int64_t *po = perm;
    const hipMemcpyKind in_kind = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    HIP_TRY(h, hipMemcpyAsync(h->st_v, v_ode, vl * sizeof(double), in_kind, h->stream));
    HIP_TRY(h, hipMemcpyAsync(h->st_u, u_ode, ul * sizeof(double), in_kind, h->stream));
    HIP_TRY(h, hipMemcpyAsync(h->st_dv, h->f_mass, (size_t)nf * sizeof(double),
                              hipMemcpyDeviceToDevice, h->stream));
    if (host) {
        TRY(dalloc(h, &tmp, vl + ul + (size_t)nf));
        vo = tmp;
        uo = tmp + vl;
        po = perm ? reinterpret_cast<int64_t *>(tmp + vl + ul) : nullptr;
    }
    // the cell-sorted order of the current positions (the flag stays raised until the
    // preparation below has rebuilt everything from the sorted state)
    hipLaunchKernelGGL((k_set_i32), dim3(1), dim3(1), 0, h->stream, h->flag, 1);
    int rc = nd == 2 ? build_cells<2>(h, h->flag, nf, h->st_u, h->f_cell, h->f_rank, h->f_count,
                                      h->f_start, h->f_perm)
                     : build_cells<3>(h, h->flag, nf, h->st_u, h->f_cell, h->f_rank, h->f_count,
                                      h->f_start, h->f_perm);
    if (rc == TPSPH_OK) {
        if (nd == 2)
            hipLaunchKernelGGL((k_permute_state<2>), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream, nf,
                               h->P.vrows, h->f_perm, h->st_v, h->st_u, h->st_dv, vo, uo, h->f_mass, po);
        else
            hipLaunchKernelGGL((k_permute_state<3>), dim3(nblocks(nf)), dim3(kBlock), 0, h->stream, nf,
                               h->P.vrows, h->f_perm, h->st_v, h->st_u, h->st_dv, vo, uo, h->f_mass, po);
        rc = launch_check(h, "k_permute_state");
    }
    // rebuild from the sorted state: the cell-sorted order is the one just computed (within
    // a cell the order is by caller index, which now increases along it), so the internal
    // permutation is the identity; pressures and wall values of the sorted state are ready
    // for tpsph_get_pressures, and the next kick need not rebuild
    if (rc == TPSPH_OK) rc = prepare_dispatch(h, vo, uo);
    if (rc == TPSPH_OK) {
        hipLaunchKernelGGL((k_set_i32), dim3(1), dim3(1), 0, h->stream, h->flag, 0);
        rc = launch_check(h, "k_set_i32");
    }
    if (rc == TPSPH_OK && host) {
        if (hipMemcpyAsync(v_ode, vo, vl * sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipMemcpyAsync(u_ode, uo, ul * sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            (perm && hipMemcpyAsync(perm, po, (size_t)nf * sizeof(int64_t), hipMemcpyDeviceToHost,
                                    h->stream) != hipSuccess) ||
            hipStreamSynchronize(h->stream) != hipSuccess)
            rc = fail(h, TPSPH_ERR_HIP, "copy of the sorted state to the host failed");
    }
    if (tmp) (void)hipFree(tmp);
    // a recorded trajectory (reverse mode) belongs to the old order
    h->adj_ready = false;
    h->have_kick = rc == TPSPH_OK;
    return rc;
}

int tpsph_synchronize(tpsph_handle *h)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return TPSPH_OK;
}

int tpsph_enable_timing(tpsph_handle *h, int32_t enable)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    h->timing = enable != 0;
    h->records.clear();
    h->events_used = 0;
    return TPSPH_OK;
}

int tpsph_kernel_time(tpsph_handle *h, int32_t which, int64_t *launches, double *total_ms)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (which < 0 || which >= TPSPH_TIME_COUNT) return fail(h, TPSPH_ERR_ARGUMENT, "unknown kernel id");
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    int64_t n = 0;
    double sum = 0.0;
    for (const auto &r : h->records) {
        if (r.kind != which) continue;
        float ms = 0.0f;
        HIP_TRY(h, hipEventElapsedTime(&ms, r.a, r.b));
        sum += ms;
        n++;
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = sum;
    return TPSPH_OK;
}

void tpsph_destroy(tpsph_handle *h)
{
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    void *ptrs[] = {h->flag, h->f_cell, h->f_rank, h->f_count, h->f_start, h->f_perm,
                    h->scan_tiles, h->f_mass, h->xs, h->ys, h->zs, h->xf, h->wxf, h->snap, h->rec, h->vol,
                    h->w_start, h->w_perm, h->wx, h->wy, h->wz, h->wpos, h->wpv, h->wall_near,
                    h->pw, h->nbr, h->cnt, h->wnbr, h->wcnt, h->anbr, h->acnt, h->ord, h->bcount, h->bstart,
                    h->brow_gb, h->brow_lds, h->bntot, h->nel, h->counters, h->wactive, h->bwork, h->nbr16,
                    h->wtargets, h->iperm, h->dyn, h->vmax_part, h->wvel, h->ddgrad,
                    h->rhow, h->st_v, h->st_u, h->st_dv, h->st_du, h->sol_dv, h->sol_tv,
                    h->sol_tu, h->integ, h->nrebuild, h->adj_kb, h->adj_coef, h->adj_th,
                    h->adj_gw, h->adj_red, h->adj_vb, h->adj_ub, h->adj_tvb, h->adj_tub,
                    h->adj_ck, h->adj_st, h->sum_rho, h->bfall};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : h->event_pool) (void)hipEventDestroy(e);
    delete h;
}

}  // extern "C"
