// tpsph.hip -- MI355X (gfx950) WCSPH right-hand side behind the C ABI of include/tpsph.h.
//
// Pipeline of one kick! (semidiscretization.jl:589-612), all on the handle's stream:
//   1. fluid cell list: linear cell id + per-cell histogram (atomics)      [k_fluid_cells]
//      exclusive scan of the histogram                                     [k_scan_*]
//      scatter + per-cell index sort (a stable counting sort)              [k_scatter, k_sort_cells]
//   2. gather the fluid state into cell-sorted SoA, fluid state equation   [k_gather]
//   3. Adami pressure extrapolation + inverse state equation per wall      [k_wall]
//   4. fused fluid-fluid + fluid-wall interaction + gravity, written back
//      to the caller's particle order                                      [k_fluid]
// The static wall cell list is built once at create.
//
// Cells have edge R (1 + 2^-30), R = compact support 2h, so every pair with d^2 <= R^2 lies
// in adjacent cells even after rounding of the cell coordinates.  Cell coordinates are
// clamped into the grid, which keeps the search exact for particles outside the box.
//
// The radius test evaluates d^2 = (dx*dx + dy*dy) + dz*dz with FP contraction off, the
// order the oracle uses, so neighbour sets agree bit for bit.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "tpsph.h"

namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

thread_local std::string g_create_error;

struct DevParams {
    int nd;
    int nf, nw;
    int use_visc, use_dd, clip_fluid, clip_wall;
    double h, h_inv, R2, two_h, almostzero;
    double kval_c;   // kernel normalisation sigma * h^-d
    double kgrad_c;  // -5 * sigma * h^-d * h^-2
    double rho0, gamma, inv_gamma, p_bg, B, c;
    double alpha, beta, eps_h2;  // epsilon * h^2
    double dd_coef;              // delta * h * c
    double g[3];
    double p_offset;
    double lo[3];
    double inv_cell;
    int n[3];
};

// ------------------------------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------------------------------

// Julia's max(x, 0) for Float64: NaN propagates, -0.0 becomes +0.0.
__device__ __forceinline__ double julia_max0(double x) { return (x != x) ? x : (x > 0.0 ? x : 0.0); }

// Squared distance in the shared evaluation order, no contraction.
template <int ND>
__device__ __forceinline__ double dist2(const double *pd)
{
#pragma clang fp contract(off)
    double s = pd[0] * pd[0] + pd[1] * pd[1];
    if constexpr (ND == 3) s = s + pd[2] * pd[2];
    return s;
}

__device__ __forceinline__ int cell_1d(double x, double lo, double inv, int n)
{
    double f = floor((x - lo) * inv);
    f = fmin(fmax(f, 0.0), (double)(n - 1));  // fmax/fmin also map NaN into the grid
    return (int)f;
}

template <int ND>
__device__ __forceinline__ void cell_of(const DevParams &P, const double *x, int *c)
{
    c[0] = cell_1d(x[0], P.lo[0], P.inv_cell, P.n[0]);
    c[1] = cell_1d(x[1], P.lo[1], P.inv_cell, P.n[1]);
    c[2] = (ND == 3) ? cell_1d(x[2], P.lo[2], P.inv_cell, P.n[2]) : 0;
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

// Iterate over the candidate ranges of the 3^ND cells around cell c: one contiguous range
// of the cell-sorted arrays per (y, z) row of 3 cells.
template <int ND, typename F>
__device__ __forceinline__ void foreach_row(const DevParams &P, const int *c, const int *start,
                                            F &&f)
{
    const int x0 = max(c[0] - 1, 0), x1 = min(c[0] + 1, P.n[0] - 1);
    const int zlo = (ND == 3) ? max(c[2] - 1, 0) : 0;
    const int zhi = (ND == 3) ? min(c[2] + 1, P.n[2] - 1) : 0;
    const int ylo = max(c[1] - 1, 0), yhi = min(c[1] + 1, P.n[1] - 1);
    for (int z = zlo; z <= zhi; z++)
        for (int y = ylo; y <= yhi; y++) {
            const int row = P.n[0] * (y + P.n[1] * z);
            f(start[row + x0], start[row + x1 + 1]);
        }
}

// ------------------------------------------------------------------------------------------
// cell list (counting sort)
// ------------------------------------------------------------------------------------------
__global__ void k_fill_i32(int *p, int value, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

template <int ND>
__global__ void k_cells(DevParams P, int n, const double *__restrict__ x, int stride,
                        int *__restrict__ cell, int *__restrict__ rank, int *__restrict__ count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double xi[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < ND; d++) xi[d] = x[(size_t)i * stride + d];
    int c[3];
    cell_of<ND>(P, xi, c);
    int lin = c[0] + P.n[0] * (c[1] + P.n[1] * c[2]);
    cell[i] = lin;
    rank[i] = atomicAdd(&count[lin], 1);
}

__device__ __forceinline__ int block_exclusive_scan(int value, int *s, int &total)
{
    s[threadIdx.x] = value;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        int t = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    total = s[kBlock - 1];
    int excl = s[threadIdx.x] - value;
    __syncthreads();
    return excl;
}

// Exclusive scan of one tile of kScanTile elements per block; tile totals to tile_sums.
__global__ void __launch_bounds__(kBlock) k_scan_tiles(const int *in, int *out, int *tile_sums, int n)
{
    __shared__ int s[kBlock];
    int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        sum += v[k];
    }
    int total;
    int run = block_exclusive_scan(sum, s, total);
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// In-place exclusive scan of n values by a single block.
__global__ void __launch_bounds__(kBlock) k_scan_single(int *data, int n)
{
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

__global__ void k_scan_add(int *out, const int *tile_sums, int n)
{
    int i = blockIdx.x * kScanTile + threadIdx.x;
    int add = tile_sums[blockIdx.x];
    for (int k = 0; k < kScanItems; k++, i += kBlock)
        if (i < n) out[i] += add;
}

__global__ void k_scatter(int n, const int *__restrict__ cell, const int *__restrict__ rank,
                          const int *__restrict__ start, int *__restrict__ perm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    perm[start[cell[i]] + rank[i]] = i;
}

// Order the members of every cell by particle index: the result equals a stable counting
// sort, independent of the order in which the histogram atomics were served.
__global__ void k_sort_cells(int ncells, const int *__restrict__ start, int *__restrict__ perm)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
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

// ------------------------------------------------------------------------------------------
// gather into cell-sorted SoA + fluid state equation (wcsph/system.jl:426-438)
// ------------------------------------------------------------------------------------------
template <int ND>
__global__ void k_gather(DevParams P, const int *__restrict__ perm, const double *__restrict__ u,
                         const double *__restrict__ v, const double *__restrict__ m,
                         double *__restrict__ xs, double *__restrict__ vs,
                         double *__restrict__ rhos, double *__restrict__ ps,
                         double *__restrict__ ms)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.nf) return;
    int i = perm[s];
    const size_t nf = (size_t)P.nf;
#pragma unroll
    for (int d = 0; d < ND; d++) {
        xs[d * nf + s] = u[(size_t)i * ND + d];
        vs[d * nf + s] = v[(size_t)i * (ND + 1) + d];
    }
    double rho = v[(size_t)i * (ND + 1) + ND];
    rhos[s] = rho;
    ps[s] = state_equation(P, rho);
    ms[s] = m[i];
}

// ------------------------------------------------------------------------------------------
// Adami pressure extrapolation (dummy_particles.jl:489-569), one thread per wall particle in
// wall-cell order, over the fluid neighbours within 2h.
// ------------------------------------------------------------------------------------------
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_wall(DevParams P, const double *__restrict__ wxs, const int *__restrict__ fstart,
           const double *__restrict__ xs, const double *__restrict__ rhos,
           const double *__restrict__ ps, double *__restrict__ pw, double *__restrict__ rhow)
{
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= P.nw) return;
    const size_t nw = (size_t)P.nw, nf = (size_t)P.nf;
    double xw[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < ND; d++) xw[d] = wxs[d * nw + w];
    int c[3];
    cell_of<ND>(P, xw, c);
    double pressure = 0.0, volume = 0.0;
    foreach_row<ND>(P, c, fstart, [&](int jb, int je) {
        for (int j = jb; j < je; j++) {
            double pd[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int d = 0; d < ND; d++) pd[d] = xw[d] - xs[d * nf + j];
            double d2 = dist2<ND>(pd);
            if (d2 > P.R2) continue;
            double r = sqrt(d2);
            double W = wendland_c2(P, r);
            double rho_f = rhos[j];
            double hydro = P.g[0] * (rho_f * pd[0]) + P.g[1] * (rho_f * pd[1]);
            if constexpr (ND == 3) hydro += P.g[2] * (rho_f * pd[2]);
            double sum_pressures = (P.p_offset + ps[j]) + hydro;
            pressure += sum_pressures * W;
            volume += W;
        }
    });
    if (volume > DBL_EPSILON) pressure /= volume;
    if (P.clip_wall) pressure = julia_max0(pressure);
    pw[w] = pressure;
    rhow[w] = P.rho0 * pow((pressure - P.p_bg) / P.B + 1.0, P.inv_gamma);
}

// ------------------------------------------------------------------------------------------
// Fused WCSPH interaction (wcsph/rhs.jl:5-127) for fluid-fluid and fluid-wall neighbours,
// plus gravity (semidiscretization.jl:720-728).  One thread per fluid particle in
// cell-sorted order; results go to the caller's order through perm.
//
// Per fluid-fluid pair with grad W = kdr * r_ab, kdr = -5 sigma h^-2 (1 - q/2)^3:
//   dv   += -m_b (p_a + p_b)/(rho_a rho_b) grad W                    (pressure, continuity form)
//   dv   += (m_b alpha c mu + m_b beta mu^2)/rho_mean grad W, if v_ab.r_ab < 0   (Monaghan AV)
//   drho += rho_a/rho_b m_b (v_ab . grad W)                           (continuity)
//   drho += delta h c (m_b/rho_b) psi . grad W, psi = 2(rho_a-rho_b) r_ab/|r_ab|^2 (Molteni)
// psi . grad W = 2 (rho_a - rho_b) kdr (|r_ab|^2/|r_ab|^2), so the diffusion term needs no
// division.  Wall neighbours: v_b = 0, p_b / rho_b from the Adami pass, no viscosity (the
// wall model's viscosity is `nothing`) and no density diffusion.
// ------------------------------------------------------------------------------------------
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_fluid(DevParams P, const int *__restrict__ perm, const int *__restrict__ fstart,
            const double *__restrict__ xs, const double *__restrict__ vs,
            const double *__restrict__ rhos, const double *__restrict__ ps,
            const double *__restrict__ ms, const int *__restrict__ wstart,
            const double *__restrict__ wxs, const double *__restrict__ wms,
            const double *__restrict__ pw, const double *__restrict__ rhow,
            double *__restrict__ dv)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.nf) return;
    const size_t nf = (size_t)P.nf, nw = (size_t)P.nw;
    double xa[3] = {0.0, 0.0, 0.0}, va[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < ND; d++) {
        xa[d] = xs[d * nf + s];
        va[d] = vs[d * nf + s];
    }
    const double rho_a = rhos[s], p_a = ps[s];
    int c[3];
    cell_of<ND>(P, xa, c);

    double acc[3] = {0.0, 0.0, 0.0}, drho = 0.0;
    foreach_row<ND>(P, c, fstart, [&](int jb, int je) {
        for (int j = jb; j < je; j++) {
            double pd[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int d = 0; d < ND; d++) pd[d] = xa[d] - xs[d * nf + j];
            double d2 = dist2<ND>(pd);
            if (d2 > P.R2) continue;
            double r = sqrt(d2);
            if (r < P.almostzero) continue;
            double q = r * P.h_inv;
            double w = 1.0 - q * 0.5;
            double kdr = P.kgrad_c * (w * w * w);
            double m_b = ms[j], rho_b = rhos[j], p_b = ps[j];
            double vd[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int d = 0; d < ND; d++) vd[d] = va[d] - vs[d * nf + j];
            double vr = vd[0] * pd[0] + vd[1] * pd[1];
            if constexpr (ND == 3) vr += vd[2] * pd[2];
            double fac = -m_b * ((p_a + p_b) / (rho_a * rho_b));
            if (P.use_visc && vr < 0.0) {
                double mu = (P.h * vr) / (d2 + P.eps_h2);
                double rho_mean = (rho_a + rho_b) * 0.5;
                fac += (m_b * P.alpha * P.c * mu + m_b * P.beta * (mu * mu)) / rho_mean;
            }
            fac *= kdr;
#pragma unroll
            for (int d = 0; d < ND; d++) acc[d] += fac * pd[d];
            drho += (rho_a / rho_b) * m_b * (kdr * vr);
            if (P.use_dd) drho += P.dd_coef * (m_b / rho_b) * (2.0 * (rho_a - rho_b) * kdr);
        }
    });

    double accw[3] = {0.0, 0.0, 0.0}, drhow = 0.0;
    if (P.nw > 0) {
        foreach_row<ND>(P, c, wstart, [&](int jb, int je) {
            for (int j = jb; j < je; j++) {
                double pd[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int d = 0; d < ND; d++) pd[d] = xa[d] - wxs[d * nw + j];
                double d2 = dist2<ND>(pd);
                if (d2 > P.R2) continue;
                double r = sqrt(d2);
                if (r < P.almostzero) continue;
                double q = r * P.h_inv;
                double w = 1.0 - q * 0.5;
                double kdr = P.kgrad_c * (w * w * w);
                double m_b = wms[j], rho_b = rhow[j], p_b = pw[j];
                double fac = -m_b * ((p_a + p_b) / (rho_a * rho_b)) * kdr;
#pragma unroll
                for (int d = 0; d < ND; d++) accw[d] += fac * pd[d];
                double vr = va[0] * pd[0] + va[1] * pd[1];
                if constexpr (ND == 3) vr += va[2] * pd[2];
                drhow += (rho_a / rho_b) * m_b * (kdr * vr);
            }
        });
    }

    const int i = perm[s];
#pragma unroll
    for (int d = 0; d < ND; d++) dv[(size_t)i * (ND + 1) + d] = (acc[d] + accw[d]) + P.g[d];
    dv[(size_t)i * (ND + 1) + ND] = drho + drhow;
}

// drift!: du = v (semidiscretization.jl:561-571; delta_v = 0 without shifting)
template <int ND>
__global__ void k_drift(int nf, const double *__restrict__ v, double *__restrict__ du)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
#pragma unroll
    for (int d = 0; d < ND; d++) du[(size_t)i * ND + d] = v[(size_t)i * (ND + 1) + d];
}

// One 2N-storage stage of the partition (v, u) with k = (dv, v): the drift slope is read
// before v is updated, as the integrator evaluates kick! and drift! at the same stage state.
template <int ND>
__global__ void k_ck_stage(int nf, double *__restrict__ v, double *__restrict__ u,
                           double *__restrict__ tv, double *__restrict__ tu,
                           const double *__restrict__ dv, double a, double b, double dt,
                           int first)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
#pragma unroll
    for (int d = 0; d < ND; d++) {
        const size_t ku = (size_t)i * ND + d, kv = (size_t)i * (ND + 1) + d;
        double slope = v[kv];
        double t = first ? dt * slope : fma(dt, slope, a * tu[ku]);
        tu[ku] = t;
        u[ku] = fma(b, t, u[ku]);
    }
#pragma unroll
    for (int d = 0; d <= ND; d++) {
        const size_t kv = (size_t)i * (ND + 1) + d;
        double t = first ? dt * dv[kv] : fma(dt, dv[kv], a * tv[kv]);
        tv[kv] = t;
        v[kv] = fma(b, t, v[kv]);
    }
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

// ------------------------------------------------------------------------------------------
// neighbour CSR export (parity tests)
// ------------------------------------------------------------------------------------------
// rows: points xr (SoA, cell-sorted, row_perm maps to caller order), columns: cell-sorted
// points xc with col_start cell offsets and col_perm to caller order.
template <int ND>
__global__ void k_nbr_count(DevParams P, int nrows, const double *__restrict__ xr,
                            const int *__restrict__ row_perm, int ncols,
                            const double *__restrict__ xc, const int *__restrict__ col_start,
                            int *__restrict__ counts)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nrows) return;
    double xa[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < ND; d++) xa[d] = xr[(size_t)d * nrows + s];
    int c[3];
    cell_of<ND>(P, xa, c);
    int cnt = 0;
    foreach_row<ND>(P, c, col_start, [&](int jb, int je) {
        for (int j = jb; j < je; j++) {
            double pd[3] = {0.0, 0.0, 0.0};
            for (int d = 0; d < ND; d++) pd[d] = xa[d] - xc[(size_t)d * ncols + j];
            if (dist2<ND>(pd) <= P.R2) cnt++;
        }
    });
    counts[row_perm[s]] = cnt;
}

template <int ND>
__global__ void k_nbr_fill(DevParams P, int nrows, const double *__restrict__ xr,
                           const int *__restrict__ row_perm, int ncols,
                           const double *__restrict__ xc, const int *__restrict__ col_start,
                           const int *__restrict__ col_perm, const int64_t *__restrict__ offsets,
                           int *__restrict__ idx)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nrows) return;
    double xa[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < ND; d++) xa[d] = xr[(size_t)d * nrows + s];
    int c[3];
    cell_of<ND>(P, xa, c);
    const int row = row_perm[s];
    int64_t o = offsets[row];
    const int64_t b = o;
    foreach_row<ND>(P, c, col_start, [&](int jb, int je) {
        for (int j = jb; j < je; j++) {
            double pd[3] = {0.0, 0.0, 0.0};
            for (int d = 0; d < ND; d++) pd[d] = xa[d] - xc[(size_t)d * ncols + j];
            if (dist2<ND>(pd) <= P.R2) idx[o++] = col_perm[j];
        }
    });
    for (int64_t k = b + 1; k < o; k++) {  // ascending column order
        int key = idx[k];
        int64_t j = k - 1;
        while (j >= b && idx[j] > key) {
            idx[j + 1] = idx[j];
            j--;
        }
        idx[j + 1] = key;
    }
}

template <int ND>
__global__ void k_aos_to_soa(int n, const double *__restrict__ x, double *__restrict__ xs,
                             const int *__restrict__ perm)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int i = perm[s];
    for (int d = 0; d < ND; d++) xs[(size_t)d * n + s] = x[(size_t)i * ND + d];
}

template <typename T>
__global__ void k_unpermute(int n, const int *__restrict__ perm, const T *__restrict__ sorted,
                            T *__restrict__ out)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[perm[s]] = sorted[s];
}

inline int nblocks(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace

// ==========================================================================================
// handle
// ==========================================================================================
struct tpsph_handle {
    tpsph_params params{};
    DevParams P{};
    int nd = 0;
    int nf = 0, nw = 0;
    int ncells = 0;
    hipStream_t stream = nullptr;
    std::string last_error;

    // fluid, per kick
    int *f_cell = nullptr, *f_rank = nullptr, *f_count = nullptr, *f_start = nullptr,
        *f_perm = nullptr;
    int *scan_tiles = nullptr;
    double *f_mass = nullptr;  // caller order
    double *xs = nullptr, *vs = nullptr, *rhos = nullptr, *ps = nullptr, *ms = nullptr;
    // wall, static (cell-sorted) + per kick
    int *w_start = nullptr, *w_perm = nullptr;
    double *wxs = nullptr, *wms = nullptr, *pw = nullptr, *rhow = nullptr;
    // host-mode staging and solver scratch
    double *st_v = nullptr, *st_u = nullptr, *st_dv = nullptr, *st_du = nullptr;
    double *sol_dv = nullptr, *sol_tv = nullptr, *sol_tu = nullptr;
    bool have_kick = false;

    // timing
    bool timing = false;
    struct Rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> event_pool;
    size_t events_used = 0;
    std::vector<Rec> records;
};

namespace {

int fail(tpsph_handle *h, int code, const std::string &msg)
{
    if (h)
        h->last_error = msg;
    else
        g_create_error = msg;
    return code;
}

#define HIP_TRY(h, expr)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((h), TPSPH_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

#define TRY(expr)                  \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != TPSPH_OK) return rc_; \
    } while (0)

template <typename T>
int dalloc(tpsph_handle *h, T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    HIP_TRY(h, hipMalloc((void **)p, count * sizeof(T)));
    return TPSPH_OK;
}

int launch_check(tpsph_handle *h, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(h, TPSPH_ERR_HIP, std::string("launch of ") + what + ": " + hipGetErrorString(e));
    return TPSPH_OK;
}

struct TimedScope {
    tpsph_handle *h;
    int kind;
    hipEvent_t b = nullptr;
    TimedScope(tpsph_handle *h_, int kind_) : h(h_), kind(kind_)
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

// exclusive scan of n ints (in) into out, on the handle's stream
int scan(tpsph_handle *h, const int *in, int *out, int n)
{
    int tiles = (n + kScanTile - 1) / kScanTile;
    k_scan_tiles<<<tiles, kBlock, 0, h->stream>>>(in, out, h->scan_tiles, n);
    k_scan_single<<<1, kBlock, 0, h->stream>>>(h->scan_tiles, tiles);
    k_scan_add<<<tiles, kBlock, 0, h->stream>>>(out, h->scan_tiles, n);
    return launch_check(h, "scan");
}

// Build the cell list of n points x (stride doubles per point) into start/perm.
template <int ND>
int build_cells(tpsph_handle *h, int n, const double *x, int stride, int *cell, int *rank,
                int *count, int *start, int *perm)
{
    const int nc = h->ncells;
    k_fill_i32<<<nblocks(nc + 1), kBlock, 0, h->stream>>>(count, 0, nc + 1);
    if (n > 0) k_cells<ND><<<nblocks(n), kBlock, 0, h->stream>>>(h->P, n, x, stride, cell, rank, count);
    TRY(launch_check(h, "k_cells"));
    TRY(scan(h, count, start, nc + 1));
    if (n > 0) {
        k_scatter<<<nblocks(n), kBlock, 0, h->stream>>>(n, cell, rank, start, perm);
        k_sort_cells<<<nblocks(nc), kBlock, 0, h->stream>>>(nc, start, perm);
    }
    return launch_check(h, "k_scatter/k_sort_cells");
}

template <int ND>
int kick_impl(tpsph_handle *h, double *dv, const double *v, const double *u)
{
    const int nf = h->nf, nw = h->nw;
    {
        TimedScope t(h, TPSPH_TIME_CELL_LIST);
        TRY(build_cells<ND>(h, nf, u, ND, h->f_cell, h->f_rank, h->f_count, h->f_start,
                            h->f_perm));
    }
    if (nf > 0) {
        TimedScope t(h, TPSPH_TIME_GATHER);
        k_gather<ND><<<nblocks(nf), kBlock, 0, h->stream>>>(h->P, h->f_perm, u, v, h->f_mass,
                                                             h->xs, h->vs, h->rhos, h->ps, h->ms);
        TRY(launch_check(h, "k_gather"));
    }
    if (nw > 0) {
        TimedScope t(h, TPSPH_TIME_WALL);
        k_wall<ND><<<nblocks(nw), kBlock, 0, h->stream>>>(h->P, h->wxs, h->f_start, h->xs,
                                                           h->rhos, h->ps, h->pw, h->rhow);
        TRY(launch_check(h, "k_wall"));
    }
    if (nf > 0) {
        TimedScope t(h, TPSPH_TIME_FLUID);
        k_fluid<ND><<<nblocks(nf), kBlock, 0, h->stream>>>(
            h->P, h->f_perm, h->f_start, h->xs, h->vs, h->rhos, h->ps, h->ms, h->w_start,
            h->wxs, h->wms, h->pw, h->rhow, dv);
        TRY(launch_check(h, "k_fluid"));
    }
    h->have_kick = true;
    return TPSPH_OK;
}

int kick_dispatch(tpsph_handle *h, double *dv, const double *v, const double *u)
{
    return h->nd == 2 ? kick_impl<2>(h, dv, v, u) : kick_impl<3>(h, dv, v, u);
}

bool is_host(const tpsph_handle *h) { return h->params.memory == TPSPH_MEM_HOST; }

size_t v_len(const tpsph_handle *h) { return (size_t)h->nf * (h->nd + 1); }
size_t u_len(const tpsph_handle *h) { return (size_t)h->nf * h->nd; }

int ensure_staging(tpsph_handle *h)
{
    if (h->st_v) return TPSPH_OK;
    TRY(dalloc(h, &h->st_v, v_len(h)));
    TRY(dalloc(h, &h->st_u, u_len(h)));
    TRY(dalloc(h, &h->st_dv, v_len(h)));
    TRY(dalloc(h, &h->st_du, u_len(h)));
    return TPSPH_OK;
}

int validate(const tpsph_params *p, int64_t nf, int64_t nw, const double *fm, const double *wx,
             const double *wm)
{
    auto bad = [](const std::string &m) { return fail(nullptr, TPSPH_ERR_ARGUMENT, m); };
    if (!p) return bad("params is NULL");
    if (p->ndims != 2 && p->ndims != 3) return bad("ndims must be 2 or 3");
    if (p->kernel != TPSPH_KERNEL_WENDLAND_C2) return bad("only WendlandC2Kernel is implemented");
    if (p->memory != TPSPH_MEM_DEVICE && p->memory != TPSPH_MEM_HOST)
        return bad("memory must be TPSPH_MEM_DEVICE or TPSPH_MEM_HOST");
    if (nf < 0 || nw < 0) return bad("particle counts must be non-negative");
    if (nf > (int64_t)INT32_MAX / 4 || nw > (int64_t)INT32_MAX / 4)
        return bad("particle counts exceed the 32-bit index range of this build");
    if (!(p->smoothing_length > 0.0)) return bad("`smoothing_length` must be positive");
    if (!(p->reference_density > DBL_EPSILON))
        return bad("`reference_density` needs to be positive and larger than eps()");
    if (!(p->sound_speed > 0.0)) return bad("`sound_speed` must be positive");
    if (!(p->exponent > 0.0)) return bad("`exponent` must be positive");
    if (nf > 0 && !fm) return bad("fluid_mass is NULL");
    if (nw > 0 && (!wx || !wm)) return bad("wall_coordinates / wall_mass is NULL");
    return TPSPH_OK;
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

    // wall coordinates to the host for the bounding box (the grid is sized at create)
    std::vector<double> wx((size_t)nd * h->nw);
    if (h->nw > 0) {
        hipError_t e = hipMemcpy(wx.data(), wall_coordinates, wx.size() * sizeof(double),
                                 host ? hipMemcpyHostToHost : hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            std::string m = std::string("copy of wall_coordinates: ") + hipGetErrorString(e);
            delete h;
            return fail(nullptr, TPSPH_ERR_HIP, m);
        }
    }

    // --- derived parameters
    const double hl = params->smoothing_length;
    const double R = 2.0 * hl;  // compact_support(WendlandC2, h)
    DevParams &P = h->P;
    P.nd = nd;
    P.nf = h->nf;
    P.nw = h->nw;
    P.use_visc = params->use_viscosity ? 1 : 0;
    P.use_dd = params->use_density_diffusion ? 1 : 0;
    P.clip_fluid = params->clip_fluid_eos ? 1 : 0;
    P.clip_wall = params->clip_wall_pressure ? 1 : 0;
    P.h = hl;
    P.h_inv = 1.0 / hl;
    P.R2 = R * R;
    P.two_h = 2.0 * hl;
    P.almostzero = std::sqrt(std::nextafter(R * R, INFINITY) - R * R);
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
    P.dd_coef = params->delta * ((hl + hl) / 2.0) * P.c;
    for (int d = 0; d < 3; d++) P.g[d] = d < nd ? params->acceleration[d] : 0.0;
    P.p_offset = params->pressure_offset;

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
            lo[d] = std::fmin(lo[d], wx[(size_t)i * nd + d]);
            hi[d] = std::fmax(hi[d], wx[(size_t)i * nd + d]);
        }
    for (int d = 0; d < nd; d++)
        if (!(lo[d] <= hi[d])) {
            lo[d] = 0.0;
            hi[d] = 0.0;
        }
    const double cell = R * (1.0 + std::ldexp(1.0, -30));
    P.inv_cell = 1.0 / cell;
    int64_t ncells = 1;
    for (int d = 0; d < 3; d++) {
        if (d < nd) {
            P.lo[d] = lo[d] - cell;
            int64_t n = (int64_t)std::ceil((hi[d] + cell - P.lo[d]) * P.inv_cell) + 1;
            if (n < 1) n = 1;
            if (n > (1 << 20)) {
                delete h;
                return fail(nullptr, TPSPH_ERR_ARGUMENT, "domain box too large for the cell grid");
            }
            P.n[d] = (int)n;
        } else {
            P.lo[d] = 0.0;
            P.n[d] = 1;
        }
        ncells *= P.n[d];
    }
    if (ncells > (int64_t)INT32_MAX / 2) {
        delete h;
        return fail(nullptr, TPSPH_ERR_ARGUMENT, "cell grid too large (domain / smoothing_length)");
    }
    h->ncells = (int)ncells;

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
    CTRY(dalloc(h, &h->scan_tiles, (size_t)tiles + 1));
    CTRY(dalloc(h, &h->f_cell, nf));
    CTRY(dalloc(h, &h->f_rank, nf));
    CTRY(dalloc(h, &h->f_perm, nf));
    CTRY(dalloc(h, &h->f_count, (size_t)ncells + 1));
    CTRY(dalloc(h, &h->f_start, (size_t)ncells + 1));
    CTRY(dalloc(h, &h->f_mass, nf));
    CTRY(dalloc(h, &h->xs, (size_t)nd * nf));
    CTRY(dalloc(h, &h->vs, (size_t)nd * nf));
    CTRY(dalloc(h, &h->rhos, nf));
    CTRY(dalloc(h, &h->ps, nf));
    CTRY(dalloc(h, &h->ms, nf));
    CTRY(dalloc(h, &h->w_start, (size_t)ncells + 1));
    CTRY(dalloc(h, &h->w_perm, nw));
    CTRY(dalloc(h, &h->wxs, (size_t)nd * nw));
    CTRY(dalloc(h, &h->wms, nw));
    CTRY(dalloc(h, &h->pw, nw));
    CTRY(dalloc(h, &h->rhow, nw));
    if (nf > 0) {
        hipError_t e = hipMemcpy(h->f_mass, fluid_mass, (size_t)nf * sizeof(double), in_kind);
        if (e != hipSuccess) {
            h->last_error = std::string("copy of fluid_mass: ") + hipGetErrorString(e);
            return cleanup_fail(TPSPH_ERR_HIP);
        }
    }
    // static wall cell list: sort walls by cell once, keep SoA coordinates and masses
    if (nw > 0) {
        double *wx_d = nullptr, *wm_d = nullptr;
        int *wcell = nullptr, *wrank = nullptr, *wcount = nullptr;
        CTRY(dalloc(h, &wx_d, (size_t)nd * nw));
        CTRY(dalloc(h, &wm_d, nw));
        CTRY(dalloc(h, &wcell, nw));
        CTRY(dalloc(h, &wrank, nw));
        CTRY(dalloc(h, &wcount, (size_t)ncells + 1));
        hipError_t e1 = hipMemcpy(wx_d, wx.data(), wx.size() * sizeof(double), hipMemcpyHostToDevice);
        hipError_t e2 = hipMemcpy(wm_d, wall_mass, (size_t)nw * sizeof(double), in_kind);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            h->last_error = "copy of wall data failed";
            (void)hipFree(wx_d); (void)hipFree(wm_d); (void)hipFree(wcell); (void)hipFree(wrank); (void)hipFree(wcount);
            return cleanup_fail(TPSPH_ERR_HIP);
        }
        if (nd == 2)
            rc = build_cells<2>(h, nw, wx_d, 2, wcell, wrank, wcount, h->w_start, h->w_perm);
        else
            rc = build_cells<3>(h, nw, wx_d, 3, wcell, wrank, wcount, h->w_start, h->w_perm);
        if (rc == TPSPH_OK) {
            if (nd == 2)
                k_aos_to_soa<2><<<nblocks(nw), kBlock, 0, h->stream>>>(nw, wx_d, h->wxs, h->w_perm);
            else
                k_aos_to_soa<3><<<nblocks(nw), kBlock, 0, h->stream>>>(nw, wx_d, h->wxs, h->w_perm);
            // wms[s] = wm[w_perm[s]]: gather through a 1-D "AoS" of stride 1
            k_aos_to_soa<1><<<nblocks(nw), kBlock, 0, h->stream>>>(nw, wm_d, h->wms, h->w_perm);
            rc = launch_check(h, "wall setup");
        }
        hipError_t es = hipStreamSynchronize(h->stream);
        (void)hipFree(wx_d); (void)hipFree(wm_d); (void)hipFree(wcell); (void)hipFree(wrank); (void)hipFree(wcount);
        if (rc != TPSPH_OK) return cleanup_fail(rc);
        if (es != hipSuccess) {
            h->last_error = std::string("wall setup: ") + hipGetErrorString(es);
            return cleanup_fail(TPSPH_ERR_HIP);
        }
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
            k_drift<2><<<nblocks(h->nf), kBlock, 0, h->stream>>>(h->nf, v_ode, du_ode);
        else
            k_drift<3><<<nblocks(h->nf), kBlock, 0, h->stream>>>(h->nf, v_ode, du_ode);
        return launch_check(h, "k_drift");
    }
    for (int i = 0; i < h->nf; i++)
        for (int d = 0; d < h->nd; d++)
            du_ode[(size_t)i * h->nd + d] = v_ode[(size_t)i * (h->nd + 1) + d];
    return TPSPH_OK;
}

int tpsph_get_pressures(tpsph_handle *h, double *fluid_pressure, double *wall_pressure,
                        double *wall_density)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    if (!h->have_kick) return fail(h, TPSPH_ERR_STATE, "no kick has been evaluated yet");
    double *tmp = nullptr;
    const size_t n = (size_t)std::max(h->nf, h->nw);
    TRY(dalloc(h, &tmp, n));
    int rc = TPSPH_OK;
    auto emit = [&](double *dst, const double *sorted, const int *perm, int cnt) -> int {
        if (!dst || cnt == 0) return TPSPH_OK;
        k_unpermute<double><<<nblocks(cnt), kBlock, 0, h->stream>>>(cnt, perm, sorted, tmp);
        TRY(launch_check(h, "k_unpermute"));
        HIP_TRY(h, hipMemcpyAsync(dst, tmp, (size_t)cnt * sizeof(double), hipMemcpyDefault, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        return TPSPH_OK;
    };
    rc = emit(fluid_pressure, h->ps, h->f_perm, h->nf);
    if (rc == TPSPH_OK) rc = emit(wall_pressure, h->pw, h->w_perm, h->nw);
    if (rc == TPSPH_OK) rc = emit(wall_density, h->rhow, h->w_perm, h->nw);
    (void)hipFree(tmp);
    return rc;
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
    // fluid cell list + sorted fluid coordinates for the current positions
    TRY(nd == 2 ? build_cells<2>(h, nf, u, 2, h->f_cell, h->f_rank, h->f_count, h->f_start, h->f_perm)
                : build_cells<3>(h, nf, u, 3, h->f_cell, h->f_rank, h->f_count, h->f_start, h->f_perm));
    if (nf > 0) {
        if (nd == 2) k_aos_to_soa<2><<<nblocks(nf), kBlock, 0, h->stream>>>(nf, u, h->xs, h->f_perm);
        else k_aos_to_soa<3><<<nblocks(nf), kBlock, 0, h->stream>>>(nf, u, h->xs, h->f_perm);
        TRY(launch_check(h, "k_aos_to_soa"));
    }
    h->have_kick = false;  // xs no longer matches a kick
    int nrows, ncols;
    const double *xr, *xc;
    const int *row_perm, *col_start, *col_perm;
    if (pair == TPSPH_PAIR_FLUID_FLUID) {
        nrows = nf; ncols = nf; xr = h->xs; xc = h->xs;
        row_perm = h->f_perm; col_start = h->f_start; col_perm = h->f_perm;
    } else if (pair == TPSPH_PAIR_FLUID_WALL) {
        nrows = nf; ncols = nw; xr = h->xs; xc = h->wxs;
        row_perm = h->f_perm; col_start = h->w_start; col_perm = h->w_perm;
    } else {
        nrows = nw; ncols = nf; xr = h->wxs; xc = h->xs;
        row_perm = h->w_perm; col_start = h->f_start; col_perm = h->f_perm;
    }
    int *counts = nullptr;
    int64_t *doff = nullptr;
    int *didx = nullptr;
    TRY(dalloc(h, &counts, (size_t)nrows));
    std::vector<int> hc((size_t)nrows);
    std::vector<int64_t> ho((size_t)nrows + 1, 0);
    int rc = TPSPH_OK;
    if (nrows > 0) {
        if (nd == 2)
            k_nbr_count<2><<<nblocks(nrows), kBlock, 0, h->stream>>>(h->P, nrows, xr, row_perm, ncols, xc, col_start, counts);
        else
            k_nbr_count<3><<<nblocks(nrows), kBlock, 0, h->stream>>>(h->P, nrows, xr, row_perm, ncols, xc, col_start, counts);
        rc = launch_check(h, "k_nbr_count");
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
            if (nd == 2)
                k_nbr_fill<2><<<nblocks(nrows), kBlock, 0, h->stream>>>(h->P, nrows, xr, row_perm, ncols, xc, col_start, col_perm, doff, didx);
            else
                k_nbr_fill<3><<<nblocks(nrows), kBlock, 0, h->stream>>>(h->P, nrows, xr, row_perm, ncols, xc, col_start, col_perm, doff, didx);
            rc = launch_check(h, "k_nbr_fill");
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
    k_rk_stage<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, h->stream>>>(n, y, tmp, k, a, b, dt, first_stage);
    return launch_check(h, "k_rk_stage");
}

int tpsph_solve_ck2n54(tpsph_handle *h, double *v_ode, double *u_ode, double t0, double dt,
                       int64_t nsteps)
{
    static const double A[5] = {0.0, -567301805773.0 / 1357537059087.0,
                                -2404267990393.0 / 2016746695238.0,
                                -3550918686646.0 / 2091501179385.0,
                                -1275806237668.0 / 842570457699.0};
    static const double B[5] = {1432997174477.0 / 9575080441755.0,
                                5161836677717.0 / 13612068292357.0,
                                1720146321549.0 / 2090206949498.0,
                                3134564353537.0 / 4481467310338.0,
                                2277821191437.0 / 14882151754819.0};
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
    if (!h->sol_dv) {
        TRY(dalloc(h, &h->sol_dv, v_len(h)));
        TRY(dalloc(h, &h->sol_tv, v_len(h)));
        TRY(dalloc(h, &h->sol_tu, u_len(h)));
    }
    for (int64_t s = 0; s < nsteps; s++) {
        for (int i = 0; i < 5; i++) {
            TRY(kick_dispatch(h, h->sol_dv, v, u));
            TimedScope ts(h, TPSPH_TIME_STAGE);
            if (h->nd == 2)
                k_ck_stage<2><<<nblocks(h->nf), kBlock, 0, h->stream>>>(h->nf, v, u, h->sol_tv, h->sol_tu, h->sol_dv, A[i], B[i], dt, i == 0);
            else
                k_ck_stage<3><<<nblocks(h->nf), kBlock, 0, h->stream>>>(h->nf, v, u, h->sol_tv, h->sol_tu, h->sol_dv, A[i], B[i], dt, i == 0);
            TRY(launch_check(h, "k_ck_stage"));
        }
    }
    if (is_host(h)) {
        HIP_TRY(h, hipMemcpyAsync(v_ode, v, v_len(h) * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(h, hipMemcpyAsync(u_ode, u, u_len(h) * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
    }
    return TPSPH_OK;
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
    void *ptrs[] = {h->f_cell, h->f_rank, h->f_count, h->f_start, h->f_perm, h->scan_tiles,
                    h->f_mass, h->xs, h->vs, h->rhos, h->ps, h->ms, h->w_start, h->w_perm,
                    h->wxs, h->wms, h->pw, h->rhow, h->st_v, h->st_u, h->st_dv, h->st_du,
                    h->sol_dv, h->sol_tv, h->sol_tu};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : h->event_pool) (void)hipEventDestroy(e);
    delete h;
}

}  // extern "C"
