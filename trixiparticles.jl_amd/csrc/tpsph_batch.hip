// tpsph_batch.hip -- an ensemble of independent WCSPH cases in one set of kernels per stage.
//
// The cases share the smoothing kernel and length, the state equation's rho0, exponent and
// background pressure, the viscosity's beta and epsilon, gravity and the cell size; each
// has its own particles, walls, grid box, sound speed, Monaghan alpha and diffusion delta.
// All cases live in packed arrays: fluid particles case after case (the caller's v_ode /
// u_ode of the cases concatenated), wall particles likewise, and the cells of all case grids
// one after the other (case c owns the global cells cell_off[c] .. cell_off[c] + ncells_c).
//
// One kick is one pass of each kernel over the whole ensemble:
//   k_bcells + scan + k_scatter + k_sort_cells   a counting sort over the global cell index;
//                                                cases own contiguous cell ranges, so the
//                                                sorted order keeps every case's particles
//                                                together and, inside a case, is exactly the
//                                                order the case sorted alone would have
//   k_bgather     cell-sorted records with the case's state equation
//   k_bwall       Adami pass of every wall particle over its case's fluid cells
//   k_bfluid      fused fluid-fluid + fluid-wall interaction + gravity, the cell walk of the
//                 single-case cell search (k_fluid) on the case's grid and constants
// and one CK2N54 stage is k_bstage over all cases' (v, u).  The cell order is rebuilt at
// every kick, as the reference does; nothing depends on a skin, so a case's sums follow the
// same order whether it runs in the ensemble or alone, and the results agree bit for bit
// with tpsph_kick / tpsph_solve_ck2n54 of the case on a handle with pair_search
// TPSPH_SEARCH_CELLS, skin 0 and the same box.
//
// tpsph_batch_solve_ck2n54 captures one step (five kicks and five stages) in a HIP graph and
// replays it.  Counters and histograms are cleared by kernels (k_scatter clears the cell
// counts it consumed), so the graph holds no memset node.
#include "tpsph_internal.h"
#include "tpsph_batch.h"

using namespace tpsph;

namespace {

struct CaseDev {
    double lo[3], cs[3], inv_cs[3];
    double hs_fluid, hs_wall;  // cell growth: it depends on the box (rounding margin)
    int n[3];
    int cell_off;            // first global cell of the case
    int ncells;
    int f_off, nf;           // fluid particles (caller and sorted order)
    int w_off, nw;           // wall particles
    double c, B, av_c, dd2;  // sound speed and the constants that depend on it
};

// The constants of case cs: the shared ones with the case's grid and sound speed.
__device__ __forceinline__ void apply_case(DevParams &Q, const CaseDev &cs)
{
#pragma unroll
    for (int d = 0; d < 3; d++) {
        Q.lo[d] = cs.lo[d];
        Q.cs[d] = cs.cs[d];
        Q.inv_cs[d] = cs.inv_cs[d];
        Q.n[d] = cs.n[d];
    }
    Q.hs_fluid = cs.hs_fluid;
    Q.hs_wall = cs.hs_wall;
    Q.c = cs.c;
    Q.B = cs.B;
    Q.av_c = cs.av_c;
    Q.dd2 = cs.dd2;
}

template <int ND>
__device__ __forceinline__ int local_cell(const DevParams &Q, const double *x)
{
    const int c0 = cell_1d(x[0], Q.lo[0], Q.inv_cs[0], Q.n[0]);
    const int c1 = cell_1d(x[1], Q.lo[1], Q.inv_cs[1], Q.n[1]);
    const int c2 = (ND == 3) ? cell_1d(x[2], Q.lo[2], Q.inv_cs[2], Q.n[2]) : 0;
    return c0 + Q.n[0] * (c1 + Q.n[1] * c2);
}

// global cell of every point (n points, ND doubles each, case pcase[i]) and its rank
template <int ND>
__global__ void k_bcells(DevParams P, const CaseDev *__restrict__ cases,
                         const int *__restrict__ pcase, int n, const double *__restrict__ x,
                         int *__restrict__ cell, int *__restrict__ rank, int *__restrict__ count)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const CaseDev &cs = cases[pcase[i]];
        DevParams Q = P;
        apply_case(Q, cs);
        const int g = cs.cell_off + local_cell<ND>(Q, x + (size_t)i * ND);
        cell[i] = g;
        rank[i] = atomicAdd(&count[g], 1);
    }
}

// cell-sorted records of the fluid state (the fields of k_gather) with the case's state
// equation; fcase is indexed by the caller's index, whose case ranges the sorted order keeps
template <int ND>
__global__ void k_bgather(DevParams P, const CaseDev *__restrict__ cases,
                          const int *__restrict__ fcase, int nf, const int *__restrict__ perm,
                          const double *__restrict__ u, const double *__restrict__ v,
                          const double *__restrict__ m, double *__restrict__ xs,
                          double *__restrict__ ys, double *__restrict__ zs,
                          double4 *__restrict__ rec, double2 *__restrict__ vol)
{
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nf; s += gridDim.x * blockDim.x) {
        const int i = perm[s];
        DevParams Q = P;
        apply_case(Q, cases[fcase[i]]);
        double x[3] = {0.0, 0.0, 0.0}, vv[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < ND; d++) {
            x[d] = u[(size_t)i * ND + d];
            vv[d] = v[(size_t)i * (ND + 1) + d];
        }
        const double rho = v[(size_t)i * (ND + 1) + ND];
        xs[s] = x[0];
        ys[s] = x[1];
        if constexpr (ND == 3) zs[s] = x[2];
        rec[2 * (size_t)s] = make_double4(x[0], x[1], x[2], rho);
        rec[2 * (size_t)s + 1] = make_double4(vv[0], vv[1], vv[2], state_equation(Q, rho));
        vol[s] = make_double2(m[i] / rho, m[i]);
    }
}

// Adami pass (k_wall's cell walk) of every wall particle over its case's fluid cells
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_bwall(DevParams P, const CaseDev *__restrict__ cases, const int *__restrict__ wcase,
            FluidView F, const double4 *__restrict__ wpos, double *__restrict__ pw,
            double *__restrict__ rhow, double2 *__restrict__ wpv)
{
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < P.nw; w += gridDim.x * blockDim.x) {
        const CaseDev &cs = cases[wcase[w]];
        DevParams Q = P;
        apply_case(Q, cs);
        const double4 xw = wpos[w];
        double pressure = 0.0, volume = 0.0, vs[3] = {0.0, 0.0, 0.0};
        Pencils<ND> it;
        it.init(Q, xw.x, xw.y, xw.z, Q.hs_fluid);
        const int *start = F.start + cs.cell_off;
        int jb, je;
        while (it.next(Q, start, jb, je))
            for (int j = jb; j < je; j++) adami_term<ND>(Q, F, xw, j, pressure, volume, vs);
        adami_finish(Q, w, xw.w, pressure, volume, pw, rhow, wpv, vs, nullptr);
    }
}

// Fused fluid-fluid + fluid-wall interaction and gravity (k_fluid) per sorted fluid particle
// on its case's grid; BLOCK threads per block, CAP list entries per lane in LDS.
template <int ND, int BLOCK, int CAP>
__global__ void __launch_bounds__(BLOCK)
    k_bfluid(DevParams P, const CaseDev *__restrict__ cases, const int *__restrict__ fcase,
             FluidView F, WallView Wv, const unsigned char *__restrict__ wall_near,
             double *__restrict__ dv)
{
    __shared__ int lists[CAP * BLOCK];
    int *lst = lists + threadIdx.x;
    const int s = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    const bool active = s < P.nf;
    const int sc = active ? s : P.nf - 1;  // an in-bounds particle for the idle lanes
    const CaseDev &cs = cases[fcase[sc]];
    DevParams Q = P;
    apply_case(Q, cs);
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
    double acc[3] = {0.0, 0.0, 0.0}, drho = 0.0;
    neighbour_sweep<ND, false, CAP, BLOCK>(Q, F.start + cs.cell_off, F.x, F.y, F.z, Q.hs_fluid,
                                           active, F, Wv, lst, xa, va, rho_a, p_a, inv_rho_a,
                                           acc, drho);
    double accw[3] = {0.0, 0.0, 0.0}, drhow = 0.0;
    if (P.nw > 0) {
        bool near = false;
        if (active) near = wall_near[cs.cell_off + local_cell<ND>(Q, xa)] != 0;
        if (__any(near))
            neighbour_sweep<ND, true, CAP, BLOCK>(Q, Wv.start + cs.cell_off, Wv.x, Wv.y, Wv.z,
                                                  Q.hs_wall, near, F, Wv, lst, xa, va, rho_a,
                                                  p_a, inv_rho_a, accw, drhow);
    }
    if (!active) return;
    const int i = F.perm[s];
#pragma unroll
    for (int d = 0; d < ND; d++) dv[(size_t)i * (ND + 1) + d] = (acc[d] + accw[d]) + P.g[d];
    dv[(size_t)i * (ND + 1) + ND] = drho + drhow;
}

// One 2N-storage stage of (v, u) with slopes (dv, v) over every case: a block owns kBlock
// particles and updates their u (with the v of the stage state) before their v.
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_bstage(int nf, double *__restrict__ v, double *__restrict__ u, double *__restrict__ tv,
             double *__restrict__ tu, const double *__restrict__ dv, double a, double b,
             double dt, int first)
{
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
}

// static wall data after the wall sort: sorted positions + masses, case of every sorted wall
template <int ND>
__global__ void k_bwall_setup(int nw, const double *__restrict__ x, const double *__restrict__ m,
                              const int *__restrict__ perm, const int *__restrict__ wcase_in,
                              double *__restrict__ wx, double *__restrict__ wy,
                              double *__restrict__ wz, double4 *__restrict__ wpos,
                              int *__restrict__ wcase)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nw) return;
    const int i = perm[s];
    const double a = x[(size_t)i * ND], b = x[(size_t)i * ND + 1];
    const double c = (ND == 3) ? x[(size_t)i * ND + 2] : 0.0;
    wx[s] = a;
    wy[s] = b;
    if constexpr (ND == 3) wz[s] = c;
    wpos[s] = make_double4(a, b, c, m[i]);
    wcase[s] = wcase_in[i];
}

// every global cell with a wall particle of its case within reach (k_wall_near per case)
template <int ND>
__global__ void k_bwall_near(DevParams P, const CaseDev *__restrict__ cases,
                             const int *__restrict__ wcase, int nw,
                             const double4 *__restrict__ wpos, unsigned char *__restrict__ near)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const CaseDev &cs = cases[wcase[w]];
    DevParams Q = P;
    apply_case(Q, cs);
    const double4 x = wpos[w];
    const double reach = Q.R + Q.hs_wall;
    const int x0 = cell_1d(x.x - reach, Q.lo[0], Q.inv_cs[0], Q.n[0]);
    const int x1 = cell_1d(x.x + reach, Q.lo[0], Q.inv_cs[0], Q.n[0]);
    const int y0 = cell_1d(x.y - reach, Q.lo[1], Q.inv_cs[1], Q.n[1]);
    const int y1 = cell_1d(x.y + reach, Q.lo[1], Q.inv_cs[1], Q.n[1]);
    const int z0 = (ND == 3) ? cell_1d(x.z - reach, Q.lo[2], Q.inv_cs[2], Q.n[2]) : 0;
    const int z1 = (ND == 3) ? cell_1d(x.z + reach, Q.lo[2], Q.inv_cs[2], Q.n[2]) : 0;
    for (int cz = z0; cz <= z1; cz++)
        for (int cy = y0; cy <= y1; cy++)
            for (int cx = x0; cx <= x1; cx++)
                near[cs.cell_off + cx + Q.n[0] * (cy + Q.n[1] * cz)] = 1;
}

}  // namespace

// ==========================================================================================
// the ensemble
// ==========================================================================================
struct tpsph_batch {
    tpsph_handle h;  // shared parameters, stream and the packed arrays (fluid_view / wall_view)
    int ncases = 0;
    int64_t ncells_total = 0;
    std::vector<CaseDev> cases_h;
    CaseDev *cases = nullptr;
    int *fcase = nullptr;   // case of every fluid particle (caller = sorted ranges)
    int *wcase = nullptr;   // case of every sorted wall particle
    // tuned launch shapes (tpsph_batch_set_launch)
    int fluid_block = kBlock, wall_block = kBlock;
    // the captured step
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    const double *graph_v = nullptr, *graph_u = nullptr;
    double graph_dt = 0.0;
    bool own_stream = false;
};

namespace {

int bfail(tpsph_batch *b, int code, const std::string &msg)
{
    return fail(b ? &b->h : nullptr, code, msg);
}

template <int ND>
int batch_kick(tpsph_batch *b, double *dv, const double *v, const double *u)
{
    tpsph_handle *h = &b->h;
    const int nf = h->nf, nw = h->nw;
    const int nc = (int)b->ncells_total;
    hipStream_t st = h->stream;
    if (nf == 0) return TPSPH_OK;
    hipLaunchKernelGGL((k_bcells<ND>), dim3(gate_blocks(nf)), dim3(kBlock), 0, st, h->P, b->cases,
                       b->fcase, nf, u, h->f_cell, h->f_rank, h->f_count);
    TRY(launch_check(h, "k_bcells"));
    TRY(scan(h, nullptr, h->f_count, h->f_start, nc + 1));
    hipLaunchKernelGGL((k_scatter), dim3(gate_blocks(nf)), dim3(kBlock), 0, st, nullptr, nf, h->f_cell,
                       h->f_rank, h->f_start, h->f_perm, h->f_count);
    hipLaunchKernelGGL((k_sort_cells), dim3(gate_blocks(nc)), dim3(kBlock), 0, st, nullptr, nc,
                       h->f_start, h->f_perm);
    hipLaunchKernelGGL((k_bgather<ND>), dim3(gate_blocks(nf)), dim3(kBlock), 0, st, h->P, b->cases,
                       b->fcase, nf, h->f_perm, u, v, h->f_mass, h->xs, h->ys, h->zs, h->rec, h->vol);
    TRY(launch_check(h, "batch cell list"));
    if (nw > 0) {
        hipLaunchKernelGGL((k_bwall<ND>), dim3(gate_blocks(nw)), dim3(b->wall_block), 0, st, h->P,
                           b->cases, b->wcase, h->fluid_view(), h->wpos, h->pw, h->rhow, h->wpv);
        TRY(launch_check(h, "k_bwall"));
    }
    const int fb = b->fluid_block;
    const dim3 grid((nf + fb - 1) / fb);
    const FluidView F = h->fluid_view();
    const WallView W = h->wall_view();
    if (fb == 64)
        hipLaunchKernelGGL((k_bfluid<ND, 64, kCap>), grid, dim3(64), 0, st, h->P, b->cases, b->fcase, F, W, h->wall_near, dv);
    else if (fb == 128)
        hipLaunchKernelGGL((k_bfluid<ND, 128, kCap>), grid, dim3(128), 0, st, h->P, b->cases, b->fcase, F, W, h->wall_near, dv);
    else
        hipLaunchKernelGGL((k_bfluid<ND, 256, kCap>), grid, dim3(256), 0, st, h->P, b->cases, b->fcase, F, W, h->wall_near, dv);
    return launch_check(h, "k_bfluid");
}

int batch_kick_dispatch(tpsph_batch *b, double *dv, const double *v, const double *u)
{
    return b->h.nd == 2 ? batch_kick<2>(b, dv, v, u) : batch_kick<3>(b, dv, v, u);
}

// one CK2N54 step: five kicks, each followed by its stage update
int batch_step(tpsph_batch *b, double *v, double *u, double dt)
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
    tpsph_handle *h = &b->h;
    for (int i = 0; i < 5; i++) {
        TRY(batch_kick_dispatch(b, h->sol_dv, v, u));
        if (h->nd == 2)
            hipLaunchKernelGGL((k_bstage<2>), dim3(nblocks(h->nf)), dim3(kBlock), 0, h->stream, h->nf, v, u,
                               h->sol_tv, h->sol_tu, h->sol_dv, A[i], B[i], dt, i == 0);
        else
            hipLaunchKernelGGL((k_bstage<3>), dim3(nblocks(h->nf)), dim3(kBlock), 0, h->stream, h->nf, v, u,
                               h->sol_tv, h->sol_tu, h->sol_dv, A[i], B[i], dt, i == 0);
        TRY(launch_check(h, "k_bstage"));
    }
    return TPSPH_OK;
}

void drop_graph(tpsph_batch *b)
{
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
}

}  // namespace

extern "C" {

int tpsph_batch_create(const tpsph_params *params, int32_t n_cases, const tpsph_case *cases,
                       const double *fluid_mass, const double *wall_coordinates,
                       const double *wall_mass, void *stream, tpsph_batch **out)
{
    if (!out) return fail(nullptr, TPSPH_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!params) return fail(nullptr, TPSPH_ERR_ARGUMENT, "params is NULL");
    if (n_cases < 1 || !cases) return fail(nullptr, TPSPH_ERR_ARGUMENT, "at least one case is needed");
    if (params->memory != TPSPH_MEM_DEVICE)
        return fail(nullptr, TPSPH_ERR_ARGUMENT, "the ensemble takes device pointers (TPSPH_MEM_DEVICE)");
    if (params->adaptive_sound_speed || params->use_density_diffusion == TPSPH_DIFFUSION_ANTUONO ||
        params->wall_viscosity || params->precision != TPSPH_PRECISION_FP64 ||
        params->density_calculator != TPSPH_DENSITY_CONTINUITY)
        return fail(nullptr, TPSPH_ERR_ARGUMENT,
                    "the ensemble runs ContinuityDensity, StateEquationCole, Monaghan viscosity, "
                    "Molteni-Colagrossi or no diffusion, walls without viscosity, FP64");
    int64_t nf = 0, nw = 0;
    for (int c = 0; c < n_cases; c++) {
        if (cases[c].n_fluid < 0 || cases[c].n_wall < 0)
            return fail(nullptr, TPSPH_ERR_ARGUMENT, "particle counts must be non-negative");
        if (!(cases[c].sound_speed > 0.0))
            return fail(nullptr, TPSPH_ERR_ARGUMENT, "`sound_speed` of every case must be positive");
        for (int d = 0; d < params->ndims; d++)
            if (!(cases[c].domain_min[d] < cases[c].domain_max[d]))
                return fail(nullptr, TPSPH_ERR_ARGUMENT, "every case needs its grid box (domain_min < domain_max)");
        nf += cases[c].n_fluid;
        nw += cases[c].n_wall;
    }
    {
        tpsph_params pv = *params;
        pv.sound_speed = cases[0].sound_speed;
        TRY(validate(&pv, nf, nw, fluid_mass, wall_coordinates, wall_mass));
    }
    tpsph_batch *b = new (std::nothrow) tpsph_batch();
    if (!b) return fail(nullptr, TPSPH_ERR_HIP, "out of host memory");
    tpsph_handle *h = &b->h;
    h->params = *params;
    h->params.skin = 0.0;
    h->params.pair_search = TPSPH_SEARCH_CELLS;
    h->nd = params->ndims;
    h->nf = (int)nf;
    h->nw = (int)nw;
    h->search = TPSPH_SEARCH_CELLS;
    b->ncases = n_cases;
    auto cleanup = [&](int rc) {
        const std::string m = h->last_error.empty() ? g_create_error : h->last_error;
        tpsph_batch_destroy(b);
        return fail(nullptr, rc, m);
    };
    if (stream) {
        h->stream = (hipStream_t)stream;
    } else {
        // graph capture needs a stream of its own (default priority)
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
            delete b;
            return fail(nullptr, TPSPH_ERR_HIP, "stream creation failed");
        }
        b->own_stream = true;
    }
    // per case: the grid and constants a single handle with its box would have
    b->cases_h.resize(n_cases);
    int64_t cell_off = 0, f_off = 0, w_off = 0;
    for (int c = 0; c < n_cases; c++) {
        tpsph_params pc = h->params;
        pc.sound_speed = cases[c].sound_speed;
        pc.alpha = cases[c].alpha;
        pc.delta = cases[c].delta;
        DevParams Pc{};
        int nbricks = 0;
        double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < params->ndims; d++) {
            lo[d] = cases[c].domain_min[d];
            hi[d] = cases[c].domain_max[d];
        }
        const int rc = derive_params(&pc, (int)cases[c].n_fluid, (int)cases[c].n_wall, lo, hi, Pc, nbricks);
        if (rc != TPSPH_OK) return cleanup(rc);
        if (c == 0) h->P = Pc;  // the shared constants (cell size, kernel, radii)
        CaseDev &cd = b->cases_h[c];
        for (int d = 0; d < 3; d++) {
            cd.lo[d] = Pc.lo[d];
            cd.cs[d] = Pc.cs[d];
            cd.inv_cs[d] = Pc.inv_cs[d];
            cd.n[d] = Pc.n[d];
        }
        cd.hs_fluid = Pc.hs_fluid;
        cd.hs_wall = Pc.hs_wall;
        cd.cell_off = (int)cell_off;
        cd.ncells = Pc.ncells;
        cd.f_off = (int)f_off;
        cd.nf = (int)cases[c].n_fluid;
        cd.w_off = (int)w_off;
        cd.nw = (int)cases[c].n_wall;
        cd.c = Pc.c;
        cd.B = Pc.B;
        cd.av_c = Pc.av_c;
        cd.dd2 = Pc.dd2;
        cell_off += Pc.ncells;
        f_off += cd.nf;
        w_off += cd.nw;
        if (cell_off > (int64_t)INT32_MAX / 4) {
            h->last_error = "the cells of all cases exceed the 32-bit cell index";
            return cleanup(TPSPH_ERR_ARGUMENT);
        }
    }
    b->ncells_total = cell_off;
    h->P.nf = (int)nf;
    h->P.nw = (int)nw;
    h->P.ncells = (int)cell_off;
    const int nd = h->nd;
    const int nc = (int)cell_off;
    int rc;
#define BTRY(expr) \
    if ((rc = (expr)) != TPSPH_OK) return cleanup(rc)
    const int tiles = (nc + 1 + kScanTile - 1) / kScanTile;
    BTRY(dalloc(h, &b->cases, n_cases));
    BTRY(dalloc(h, &b->fcase, nf));
    BTRY(dalloc(h, &b->wcase, nw));
    BTRY(dalloc(h, &h->scan_tiles, (size_t)tiles + 1));
    BTRY(dalloc(h, &h->f_cell, nf));
    BTRY(dalloc(h, &h->f_rank, nf));
    BTRY(dalloc(h, &h->f_perm, nf));
    BTRY(dalloc(h, &h->f_count, (size_t)nc + 1));
    BTRY(dalloc(h, &h->f_start, (size_t)nc + 1));
    BTRY(dalloc(h, &h->f_mass, nf));
    BTRY(dalloc(h, &h->xs, (size_t)nf + kPad));
    BTRY(dalloc(h, &h->ys, (size_t)nf + kPad));
    BTRY(dalloc(h, &h->zs, (size_t)nf + kPad));
    BTRY(dalloc(h, &h->rec, 2 * (size_t)nf));
    BTRY(dalloc(h, &h->vol, nf));
    BTRY(dalloc(h, &h->w_start, (size_t)nc + 1));
    BTRY(dalloc(h, &h->w_perm, nw));
    BTRY(dalloc(h, &h->wx, (size_t)nw + kPad));
    BTRY(dalloc(h, &h->wy, (size_t)nw + kPad));
    BTRY(dalloc(h, &h->wz, (size_t)nw + kPad));
    BTRY(dalloc(h, &h->wpos, nw));
    BTRY(dalloc(h, &h->wpv, nw));
    BTRY(dalloc(h, &h->pw, nw));
    BTRY(dalloc(h, &h->rhow, nw));
    BTRY(dalloc(h, &h->wall_near, (size_t)nc));
    BTRY(dalloc(h, &h->sol_dv, (size_t)nf * (nd + 1)));
    BTRY(dalloc(h, &h->sol_tv, (size_t)nf * (nd + 1)));
    BTRY(dalloc(h, &h->sol_tu, (size_t)nf * nd));
    {
        std::vector<int> fc((size_t)nf), wc((size_t)nw);
        for (int c = 0; c < n_cases; c++) {
            const CaseDev &cd = b->cases_h[c];
            for (int i = 0; i < cd.nf; i++) fc[(size_t)cd.f_off + i] = c;
            for (int i = 0; i < cd.nw; i++) wc[(size_t)cd.w_off + i] = c;
        }
        hipError_t e = hipMemcpy(b->cases, b->cases_h.data(), n_cases * sizeof(CaseDev), hipMemcpyHostToDevice);
        if (e == hipSuccess && nf > 0) e = hipMemcpy(b->fcase, fc.data(), fc.size() * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess && nf > 0) e = hipMemcpy(h->f_mass, fluid_mass, (size_t)nf * sizeof(double), hipMemcpyDeviceToDevice);
        if (e == hipSuccess) e = hipMemset(h->f_count, 0, ((size_t)nc + 1) * sizeof(int));
        if (e == hipSuccess) e = hipMemset(h->wall_near, 0, (size_t)nc);
        if (e != hipSuccess) {
            h->last_error = std::string("ensemble set-up: ") + hipGetErrorString(e);
            return cleanup(TPSPH_ERR_HIP);
        }
        for (double *a : {h->xs, h->ys, h->zs}) BTRY(hip_fill(h, a, (size_t)nf + kPad));
        for (double *a : {h->wx, h->wy, h->wz}) BTRY(hip_fill(h, a, (size_t)nw + kPad));
        // static walls: the same counting sort over the global cells, once
        if (nw > 0) {
            int *wcell = nullptr, *wrank = nullptr, *wcount = nullptr, *wc_d = nullptr;
            rc = dalloc(h, &wcell, nw);
            if (rc == TPSPH_OK) rc = dalloc(h, &wrank, nw);
            if (rc == TPSPH_OK) rc = dalloc(h, &wcount, (size_t)nc + 1);
            if (rc == TPSPH_OK) rc = dalloc(h, &wc_d, nw);
            if (rc == TPSPH_OK && (hipMemset(wcount, 0, ((size_t)nc + 1) * sizeof(int)) != hipSuccess ||
                                   hipMemcpy(wc_d, wc.data(), wc.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess))
                rc = fail(h, TPSPH_ERR_HIP, "wall set-up copies failed");
            if (rc == TPSPH_OK) {
                hipStream_t st = h->stream;
                if (nd == 2)
                    hipLaunchKernelGGL((k_bcells<2>), dim3(gate_blocks(nw)), dim3(kBlock), 0, st, h->P, b->cases, wc_d, (int)nw, wall_coordinates, wcell, wrank, wcount);
                else
                    hipLaunchKernelGGL((k_bcells<3>), dim3(gate_blocks(nw)), dim3(kBlock), 0, st, h->P, b->cases, wc_d, (int)nw, wall_coordinates, wcell, wrank, wcount);
                rc = scan(h, nullptr, wcount, h->w_start, nc + 1);
                if (rc == TPSPH_OK) {
                    hipLaunchKernelGGL((k_scatter), dim3(gate_blocks(nw)), dim3(kBlock), 0, st, nullptr, (int)nw, wcell, wrank, h->w_start, h->w_perm, wcount);
                    hipLaunchKernelGGL((k_sort_cells), dim3(gate_blocks(nc)), dim3(kBlock), 0, st, nullptr, nc, h->w_start, h->w_perm);
                    if (nd == 2) {
                        hipLaunchKernelGGL((k_bwall_setup<2>), dim3(nblocks(nw)), dim3(kBlock), 0, st, (int)nw, wall_coordinates, wall_mass, h->w_perm, wc_d, h->wx, h->wy, h->wz, h->wpos, b->wcase);
                        hipLaunchKernelGGL((k_bwall_near<2>), dim3(nblocks(nw)), dim3(kBlock), 0, st, h->P, b->cases, b->wcase, (int)nw, h->wpos, h->wall_near);
                    } else {
                        hipLaunchKernelGGL((k_bwall_setup<3>), dim3(nblocks(nw)), dim3(kBlock), 0, st, (int)nw, wall_coordinates, wall_mass, h->w_perm, wc_d, h->wx, h->wy, h->wz, h->wpos, b->wcase);
                        hipLaunchKernelGGL((k_bwall_near<3>), dim3(nblocks(nw)), dim3(kBlock), 0, st, h->P, b->cases, b->wcase, (int)nw, h->wpos, h->wall_near);
                    }
                    rc = launch_check(h, "ensemble wall set-up");
                }
            }
            const hipError_t es = hipStreamSynchronize(h->stream);
            (void)hipFree(wcell);
            (void)hipFree(wrank);
            (void)hipFree(wcount);
            (void)hipFree(wc_d);
            if (rc == TPSPH_OK && es != hipSuccess)
                rc = fail(h, TPSPH_ERR_HIP, std::string("ensemble wall set-up: ") + hipGetErrorString(es));
            if (rc != TPSPH_OK) return cleanup(rc);
        }
    }
#undef BTRY
    *out = b;
    return TPSPH_OK;
}

int tpsph_batch_kick(tpsph_batch *b, double *dv_ode, const double *v_ode, const double *u_ode,
                     double t)
{
    (void)t;
    if (!b) return fail(nullptr, TPSPH_ERR_ARGUMENT, "ensemble is NULL");
    if (b->h.nf > 0 && (!dv_ode || !v_ode || !u_ode))
        return bfail(b, TPSPH_ERR_ARGUMENT, "dv_ode, v_ode and u_ode must not be NULL");
    return batch_kick_dispatch(b, dv_ode, v_ode, u_ode);
}

int tpsph_batch_solve_ck2n54(tpsph_batch *b, double *v_ode, double *u_ode, double t0, double dt,
                             int64_t nsteps, int32_t use_graph)
{
    (void)t0;
    if (!b) return fail(nullptr, TPSPH_ERR_ARGUMENT, "ensemble is NULL");
    if (nsteps < 0) return bfail(b, TPSPH_ERR_ARGUMENT, "nsteps must be non-negative");
    tpsph_handle *h = &b->h;
    if (h->nf == 0 || nsteps == 0) return TPSPH_OK;
    if (!v_ode || !u_ode) return bfail(b, TPSPH_ERR_ARGUMENT, "v_ode and u_ode must not be NULL");
    // the caller's arrays were written on other streams: start from a settled device
    HIP_TRY(h, hipDeviceSynchronize());
    if (!use_graph) {
        for (int64_t s = 0; s < nsteps; s++) TRY(batch_step(b, v_ode, u_ode, dt));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        return TPSPH_OK;
    }
    if (!b->graph_exec || b->graph_v != v_ode || b->graph_u != u_ode || b->graph_dt != dt) {
        drop_graph(b);
        HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        const int rc = batch_step(b, v_ode, u_ode, dt);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(h->stream, &g);
        if (rc != TPSPH_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return bfail(b, TPSPH_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
        b->graph = g;
        HIP_TRY(h, hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
        b->graph_v = v_ode;
        b->graph_u = u_ode;
        b->graph_dt = dt;
    }
    for (int64_t s = 0; s < nsteps; s++) HIP_TRY(h, hipGraphLaunch(b->graph_exec, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return TPSPH_OK;
}

int tpsph_batch_set_launch(tpsph_batch *b, int32_t which, int32_t value)
{
    if (!b) return fail(nullptr, TPSPH_ERR_ARGUMENT, "ensemble is NULL");
    if (which == TPSPH_BATCH_FLUID_BLOCK) {
        if (value != 64 && value != 128 && value != 256)
            return bfail(b, TPSPH_ERR_ARGUMENT, "the fluid block is 64, 128 or 256 threads");
        b->fluid_block = value;
    } else if (which == TPSPH_BATCH_WALL_BLOCK) {
        if (value != 64 && value != 128 && value != 256)
            return bfail(b, TPSPH_ERR_ARGUMENT, "the wall block is 64, 128 or 256 threads");
        b->wall_block = value;
    } else {
        return bfail(b, TPSPH_ERR_ARGUMENT, "unknown launch parameter");
    }
    drop_graph(b);  // the captured step holds the old shapes
    return TPSPH_OK;
}

int tpsph_batch_synchronize(tpsph_batch *b)
{
    if (!b) return fail(nullptr, TPSPH_ERR_ARGUMENT, "ensemble is NULL");
    HIP_TRY(&b->h, hipStreamSynchronize(b->h.stream));
    return TPSPH_OK;
}

const char *tpsph_batch_last_error(const tpsph_batch *b)
{
    return b ? b->h.last_error.c_str() : g_create_error.c_str();
}

void tpsph_batch_destroy(tpsph_batch *b)
{
    if (!b) return;
    tpsph_handle *h = &b->h;
    (void)hipStreamSynchronize(h->stream);
    drop_graph(b);
    void *ptrs[] = {b->cases, b->fcase, b->wcase, h->scan_tiles, h->f_cell, h->f_rank, h->f_perm,
                    h->f_count, h->f_start, h->f_mass, h->xs, h->ys, h->zs, h->rec, h->vol,
                    h->w_start, h->w_perm, h->wx, h->wy, h->wz, h->wpos, h->wpv, h->pw,
                    h->rhow, h->wall_near, h->sol_dv, h->sol_tv, h->sol_tu};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->own_stream) (void)hipStreamDestroy(h->stream);
    delete b;
}

}  // extern "C"
