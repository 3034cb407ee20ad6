// tpsph_adjoint.hip -- reverse mode of the right-hand side (the pullback of kick!) and of a
// CarpenterKennedy2N54 trajectory with step checkpoints.
//
// The reference has no automatic differentiation through its right-hand side; the pullback
// here is written by hand from the terms of the forward pass (wcsph/rhs.jl:5-127,
// fluid.jl:160-197, density_diffusion.jl:213-240, pressure_acceleration.jl:34-41,
// viscosity.jl:89-132, dummy_particles.jl:489-569 and 586-672, state_equations.jl:100-139).
// Differentiated: the fluid state (v, rho, x) and the scalars alpha (Monaghan), delta
// (Molteni-Colagrossi) and g (gravity = g * unit direction of the acceleration).  Options:
// WendlandC2, StateEquationCole with a fixed sound speed, ArtificialViscosityMonaghan,
// Molteni-Colagrossi or no density diffusion, Adami walls without viscosity, FP64.
//
// Every pass is a gather, so no floating-point atomics and the results are deterministic:
//
//   phi_ab = Gbar_a . dv_ab + rbar_a drho_ab
//
// is the adjoint-weighted pair term of particle a's rates (Gbar_a, rbar_a: the adjoint of a's
// dv and drho).  Fluid particle c collects d phi / d state_c over its own terms (phi_ca) and
// its partners' terms (phi_ac): the neighbour relation is symmetric, so the partners are its
// own neighbours.  The wall pressures and densities of the Adami pass depend on the fluid
// near each wall particle; their adjoints are gathered per wall particle first (k_adj_wall),
// which turns them into two coefficients per wall particle that the fluid pass (k_adj_fluid)
// then reads.  Non-smooth points take the branch the forward pass took: the viscosity switch
// v_ab.r_ab < 0, the clipping of the wall pressure, the volume bound of the Adami pass.
//
// Checkpoints: tpsph_adjoint_forward keeps the state at the start of every segment of
// `checkpoint_every` steps; tpsph_adjoint_reverse runs the segments backwards, recomputes each
// from its checkpoint while keeping the state of every stage, and sweeps the stages in
// reverse.  The stage map of the 2N-storage scheme,
//     tmp_i = A_i tmp_{i-1} + dt k(y_{i-1}),   y_i = y_{i-1} + B_i tmp_i,
// reverses to
//     tbar += B_i ybar,   kbar = dt tbar,   ybar += J(y_{i-1})^T kbar,   tbar *= A_i
// with k = (kick!, drift!) and J^T kbar the pullback of kick! plus the drift's vbar += kbar_u.
#include "tpsph_internal.h"

#include <algorithm>

using namespace tpsph;

namespace {

constexpr int kRedBlocks = 256;  // blocks of the deterministic parameter reduction
constexpr int kAdjLanes = 4;     // lanes per particle of the adjoint passes

// Constants of the parameter derivatives: d dd2 / d delta and the unit gravity direction.
struct AdjConsts {
    double dd2_per_delta;
    double ghat[3];
};

// The geometry of one pair at squared distance d2 (shared by the pair's two terms):
// distance d and 1/d, the kernel factor kdr (grad W = kdr r_ab), (d kdr / d distance) / d
// (d kdr / d r_ab = dkdr r_ab) and 1 / (d2 + eps h^2) of the viscosity.  sqrt and the
// reciprocals are the refined hardware approximations of the forward pass (within a few ulp).
struct PairGeom {
    double d, inv_d, kdr, dkdr, inv_q;
};

__device__ __forceinline__ PairGeom pair_geom(const DevParams &P, double d2)
{
    PairGeom g;
    g.d = sqrt_pos(d2);
    g.inv_d = rcp_pos(g.d);
    const double w = 1.0 - 0.5 * P.h_inv * g.d;
    g.kdr = P.kgrad_c * (w * w * w);
    g.dkdr = P.kgrad_c * (-1.5 * P.h_inv) * (w * w) * g.inv_d;
    g.inv_q = rcp_pos(d2 + P.eps_h2);
    return g;
}

// d phi_ab / d(inputs) of one pair term in the equation of particle a with partner b:
//   r_ab = x_a - x_b (r), the pair geometry pg, v_ab (vab, zero for a wall partner),
//   vr = v_ab . r_ab, the densities with their reciprocals, the pressures, V_b = m_b / rho_b
//   and m_b of the pair, 1 / (rho_a + rho_b), and the adjoint weights G (of dv_a) and gr (of
//   drho_a).
// Outputs: d phi / d r_ab, d phi / d vr (d phi / d v_a = vr r_ab, d phi / d v_b = -vr r_ab),
// d phi / d rho_a, d phi / d rho_b (total: also through V_b), d phi / d p_a = d phi / d p_b,
// and the derivatives along alpha and delta.  FLUID: fluid partner (viscosity and density
// diffusion); a wall partner has neither (the wall model's viscosity is `nothing`).
struct PairAdj {
    double r[3];
    double vr;
    double rho_a, rho_b, p;
    double alpha, delta;
};

template <int ND, bool FLUID>
__device__ __forceinline__ PairAdj pair_adj(const DevParams &P, const AdjConsts &K,
                                            const double *r, const PairGeom &pg,
                                            const double *vab, double vr, double rho_a,
                                            double inv_rho_a, double rho_b, double inv_rho_b,
                                            double inv_sum, double p_a, double p_b,
                                            double V_b, double m_b, const double *G, double gr)
{
    PairAdj o;
    const double kdr = pg.kdr, dkdr = pg.dkdr;
    double A = G[0] * r[0] + G[1] * r[1];
    if constexpr (ND == 3) A += G[2] * r[2];
    double fac = -V_b * (p_a + p_b) * inv_rho_a;
    double visc = 0.0, dvisc_dvr = 0.0, dvisc_dq = 0.0, dvisc_dalpha = 0.0;
    if (FLUID && P.use_visc && vr < 0.0) {
        // (m_b alpha c mu + m_b beta mu^2) / rho_mean, mu = h vr / q, q = d2 + eps h^2
        const double mu = P.h * vr * pg.inv_q;
        const double two_m = 2.0 * m_b * inv_sum;
        visc = two_m * (P.alpha * P.c * mu + P.beta * (mu * mu));
        const double dmu = two_m * (P.alpha * P.c + 2.0 * P.beta * mu);
        dvisc_dvr = dmu * P.h * pg.inv_q;
        dvisc_dq = -dmu * mu * pg.inv_q;
        dvisc_dalpha = two_m * P.c * mu;
    }
    fac += visc;
    const double dd2 = FLUID ? P.dd2 : 0.0;
    const double E = dd2 * (rho_a - rho_b) + rho_a * vr;  // drho_ab = V_b kdr E
    const double kA = kdr * A;                            // dv_ab = fac kdr r_ab
    const double dphi_dkdr = fac * A + gr * V_b * E;
    o.vr = kA * dvisc_dvr + gr * V_b * kdr * rho_a;
    const double radial = dphi_dkdr * dkdr + 2.0 * kA * dvisc_dq;
#pragma unroll
    for (int k = 0; k < 3; k++) o.r[k] = 0.0;
#pragma unroll
    for (int k = 0; k < ND; k++) o.r[k] = kdr * fac * G[k] + o.vr * vab[k] + radial * r[k];
    o.p = -kA * V_b * inv_rho_a;
    o.rho_a = kA * (V_b * (p_a + p_b) * (inv_rho_a * inv_rho_a) - visc * inv_sum) +
              gr * V_b * kdr * (dd2 + vr);
    const double dphi_dVb = -kA * (p_a + p_b) * inv_rho_a + gr * kdr * E;
    o.rho_b = -kA * visc * inv_sum - gr * V_b * kdr * dd2 - dphi_dVb * V_b * inv_rho_b;
    o.alpha = kA * dvisc_dalpha;
    o.delta = FLUID ? gr * V_b * kdr * K.dd2_per_delta * (rho_a - rho_b) : 0.0;
    return o;
}

// d p / d rho of StateEquationCole (state_equations.jl:100-129), zero where it clips
__host__ __device__ __forceinline__ double dpressure_drho(const DevParams &P, double rho)
{
    const double x = rho / P.rho0;
    if (P.clip_fluid && !(P.B * (pow(x, P.gamma) - 1.0) + P.p_bg > 0.0)) return 0.0;
    return P.B * P.gamma / P.rho0 * pow(x, P.gamma - 1.0);
}

// Every candidate of one particle: its Verlet list (list != nullptr, n >= 0), else the cell
// walk around (x, y, z) with cells grown by hs.  fn(j) applies the exact radius test.  LANES
// lanes share the particle: lane q takes list entries q, q + LANES, ...; the cell walk of an
// overflowed list runs on lane 0 alone.
template <int ND, int LANES = 1, typename Fn>
__device__ __forceinline__ void for_each_candidate(const DevParams &P, const int *__restrict__ list,
                                                   int n, int stride, int col,
                                                   const int *__restrict__ start, double hs,
                                                   double x, double y, double z, int q, Fn &&fn)
{
    if (list && n >= 0) {
        // the next entry is loaded before the current one is evaluated
        int j = q < n ? list[(size_t)q * stride + col] : 0;
        for (int k = q; k < n; k += LANES) {
            const int kn = min(k + LANES, n - 1);
            const int jn = list[(size_t)kn * stride + col];
            fn(j);
            j = jn;
        }
    } else if (q == 0) {
        Pencils<ND> it;
        it.init(P, x, y, z, hs);
        int jb, je;
        while (it.next(P, start, jb, je))
            for (int j = jb; j < je; j++) fn(j);
    }
}

// Sum of v over the LANES lanes of a group (a fixed butterfly: every lane gets the total).
template <int LANES>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int o = 1; o < LANES; o <<= 1) v += __shfl_xor(v, o);
    return v;
}

// The adjoint weights of the rates in cell-sorted order: kb[s] = {Gbar (3), rbar}.
//   STAGE = false: from a caller's dv_bar (the pullback of one kick);
//   STAGE = true: the reversed stage map, tbar += B ybar, kbar = dt tbar, the drift's
//   vbar += kbar_u, tbar *= A (register adjoints tvb / tub, state adjoints vb / ub).
template <int ND, bool STAGE>
__global__ void __launch_bounds__(kBlock)
    k_adj_weights(int nf, const int *__restrict__ iperm, const double *__restrict__ dv_bar,
                  double *__restrict__ vb, double *__restrict__ ub, double *__restrict__ tvb,
                  double *__restrict__ tub, double a, double b, double dt,
                  double4 *__restrict__ kb)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nf) return;
    double k[ND + 1];
    if constexpr (STAGE) {
#pragma unroll
        for (int c = 0; c <= ND; c++) {
            const size_t e = (size_t)i * (ND + 1) + c;
            const double t = fma(b, vb[e], tvb[e]);
            k[c] = dt * t;
            tvb[e] = a * t;
        }
#pragma unroll
        for (int d = 0; d < ND; d++) {
            const size_t e = (size_t)i * ND + d;
            const double t = fma(b, ub[e], tub[e]);
            tub[e] = a * t;
            vb[(size_t)i * (ND + 1) + d] += dt * t;  // drift!: du = v
        }
    } else {
#pragma unroll
        for (int c = 0; c <= ND; c++) k[c] = dv_bar[(size_t)i * (ND + 1) + c];
    }
    kb[iperm[i]] = make_double4(k[0], k[1], ND == 3 ? k[2] : 0.0, k[ND]);
}

// Adjoint of the Adami pass, four lanes per wall particle (wall-sorted order w): the adjoints
// of p_w and rho_w from the fluid-wall terms of the fluid near w, through the inverse state
// equation and the clipping to the adjoint of the unclipped p_w = N_w / V_w with
//   N_w = sum_f (p_offset + p_f + rho_f g.(x_w - x_f)) W_wf,   V_w = sum_f W_wf,
// stored as coef[w] = {cN, cV}: d L / d N_w = cN, d L / d V_w = cV.  gw[w] accumulates the
// partial along g: cN sum_f rho_f ghat.(x_w - x_f) W_wf.
template <int ND>
__global__ void __launch_bounds__(kBlock)
    k_adj_wall(DevParams P, AdjConsts K, FluidView F, const double4 *__restrict__ wpos,
               const double *__restrict__ pw, const double *__restrict__ rhow,
               const double2 *__restrict__ wpv, const int *__restrict__ alist,
               const int *__restrict__ acnt, int lw, const double4 *__restrict__ kb,
               double2 *__restrict__ coef, double *__restrict__ gw)
{
    // four lanes per wall particle (the loop bound is uniform per group of lanes, so the
    // partners shuffle together); lane q takes the list entries q, q + 4, ...
    const int q = threadIdx.x % kAdjLanes;
    const int w = (blockIdx.x * kBlock + (int)threadIdx.x) / kAdjLanes;
    if (w >= P.nw) return;  // whole groups
    const int n = alist ? acnt[w] : -1;
    if (n == 0) {  // no fluid within reach: p_w and rho_w are constants
        if (q == 0) coef[w] = make_double2(0.0, 0.0);
        return;
    }
    const double4 xw = wpos[w];
    const double p_w = pw[w], rho_w = rhow[w], V_w = wpv[w].y, m_w = xw.w;
    const double inv_rho_w = rcp_pos(rho_w);
    double pbar = 0.0, rbar = 0.0, vol = 0.0, gsum = 0.0;
    for_each_candidate<ND, kAdjLanes>(P, alist, n, lw, w, F.start, P.hs_fluid, xw.x, xw.y, xw.z, q, [&](int j) {
        const double4 a = F.rec[2 * (size_t)j];
        // the Adami pass's difference x_w - x_f
        const double p0 = xw.x - a.x, p1 = xw.y - a.y, p2 = (ND == 3) ? xw.z - a.z : 0.0;
        const double d2 = dist2<ND>(p0, p1, p2);
        if (d2 > P.R2) return;
        const double W = wendland_c2(P, sqrt(d2));
        vol += W;
        double gp = K.ghat[0] * p0 + K.ghat[1] * p1;
        if constexpr (ND == 3) gp += K.ghat[2] * p2;
        gsum += a.w * gp * W;
        if (d2 < P.az2) return;
        const double4 b = F.rec[2 * (size_t)j + 1];
        const double4 g = kb[j];
        const double r[3] = {-p0, -p1, -p2};  // r_ab = x_f - x_w
        const double va[3] = {b.x, b.y, b.z};
        double vr = va[0] * r[0] + va[1] * r[1];
        if constexpr (ND == 3) vr += va[2] * r[2];
        const double G[3] = {g.x, g.y, g.z};
        const PairGeom pg = pair_geom(P, d2);
        const PairAdj o = pair_adj<ND, false>(P, K, r, pg, va, vr, a.w, rcp_pos(a.w), rho_w,
                                              inv_rho_w, 0.0, b.w, p_w, V_w, m_w, G, g.w);
        pbar += o.p;
        rbar += o.rho_b;
    });
    pbar = group_sum<kAdjLanes>(pbar);
    rbar = group_sum<kAdjLanes>(rbar);
    vol = group_sum<kAdjLanes>(vol);
    gsum = group_sum<kAdjLanes>(gsum);
    if (q != 0) return;
    // inverse state equation: d rho_w / d p_w = rho_w / (gamma (p_w - p_bg + B))
    const double ptot = pbar + rbar * rho_w / (P.gamma * (p_w - P.p_bg + P.B));
    double cN = 0.0, cV = 0.0;
    if (!(P.clip_wall && !(p_w > 0.0))) {
        if (vol > DBL_EPSILON) {
            cN = ptot / vol;
            cV = -ptot * p_w / vol;
        } else {
            cN = ptot;
        }
    }
    coef[w] = make_double2(cN, cV);
    gw[w] += cN * gsum;
}

// Adjoint of the fluid terms of fluid particle c, four lanes per particle (lane q takes the
// list entries q, q + 4, ...; the partial sums are added across the four lanes in a fixed
// order): its own fluid-fluid and fluid-wall terms, its partners' fluid-fluid terms, its part
// in the Adami sums of the wall particles near it, the state equation and gravity.  Adds
// (ACC) or writes d L / d state_c into vb (v, rho) and ub (x) in the caller's order, and adds
// the partials along alpha, delta and g into th[3 i ..].
template <int ND, bool ACC>
__global__ void __launch_bounds__(kBlock)
    k_adj_fluid(int *flag, DevParams P, AdjConsts K, FluidView F, WallView Wv,
                const int *__restrict__ ord, const int *__restrict__ nbr,
                const int *__restrict__ cnt, const int *__restrict__ wnbr,
                const int *__restrict__ wcnt, int ls, const double4 *__restrict__ kb,
                const double2 *__restrict__ coef, double *__restrict__ vb,
                double *__restrict__ ub, double *__restrict__ th)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;  // the rebuild (if any) is done
    const int q = threadIdx.x % kAdjLanes;
    const int tt = (xcd_tile(blockIdx.x, gridDim.x) * kBlock + (int)threadIdx.x) / kAdjLanes;
    // the loop bound is uniform per group of lanes, so the partners shuffle together
    const bool active = tt < P.nf;
    const int t = active ? tt : P.nf - 1;
    const int s = ord ? ord[t] : t;
    const double4 a = F.rec[2 * (size_t)s], b = F.rec[2 * (size_t)s + 1];
    const double2 vc = F.vol[s];
    const double4 g = kb[s];
    const double xc[3] = {a.x, a.y, a.z}, v_c[3] = {b.x, b.y, b.z};
    const double rho_c = a.w, p_c = b.w, V_c = vc.x, m_c = vc.y, inv_rho_c = rcp_pos(rho_c);
    const double G[3] = {g.x, g.y, g.z};
    const double gr = g.w;
    double xb[3] = {0.0, 0.0, 0.0}, vbr[3] = {0.0, 0.0, 0.0};
    double rhob = 0.0, pb = 0.0, alpha = 0.0, delta = 0.0;
    const int n = nbr ? cnt[t] : -1;
    for_each_candidate<ND, kAdjLanes>(P, nbr, active ? n : 0, ls, t, F.start, P.hs_fluid, a.x, a.y,
                                      a.z, active ? q : kAdjLanes, [&](int j) {
        // every load of the candidate at once: one round trip instead of two
        const double4 na = F.rec[2 * (size_t)j];
        const double4 nb = F.rec[2 * (size_t)j + 1];
        const double2 nv = F.vol[j];
        const double4 gn = kb[j];
        double r[3] = {xc[0] - na.x, xc[1] - na.y, (ND == 3) ? xc[2] - na.z : 0.0};
        const double d2 = dist2<ND>(r[0], r[1], r[2]);
        if (!(d2 <= P.R2 && d2 >= P.az2)) return;
        double vab[3] = {v_c[0] - nb.x, v_c[1] - nb.y, (ND == 3) ? v_c[2] - nb.z : 0.0};
        double vr = vab[0] * r[0] + vab[1] * r[1];
        if constexpr (ND == 3) vr += vab[2] * r[2];
        // the geometry and the density reciprocals of the pair serve both of its terms
        const PairGeom pg = pair_geom(P, d2);
        const double inv_rho_j = rcp_pos(na.w), inv_sum = rcp_pos(rho_c + na.w);
        // own term: c's rates with partner j
        const PairAdj o = pair_adj<ND, true>(P, K, r, pg, vab, vr, rho_c, inv_rho_c, na.w,
                                             inv_rho_j, inv_sum, p_c, nb.w, nv.x, nv.y, G, gr);
#pragma unroll
        for (int k = 0; k < ND; k++) {
            xb[k] += o.r[k];
            vbr[k] += o.vr * r[k];
        }
        rhob += o.rho_a;
        pb += o.p;
        alpha += o.alpha;
        delta += o.delta;
        // partner term: j's rates with partner c (r_jc = -r, v_jc = -v_ab, same vr)
        const double rn[3] = {-r[0], -r[1], -r[2]};
        const double vn[3] = {-vab[0], -vab[1], -vab[2]};
        const double Gn[3] = {gn.x, gn.y, gn.z};
        const PairAdj qa = pair_adj<ND, true>(P, K, rn, pg, vn, vr, na.w, inv_rho_j, rho_c,
                                              inv_rho_c, inv_sum, nb.w, p_c, V_c, m_c, Gn, gn.w);
#pragma unroll
        for (int k = 0; k < ND; k++) {
            xb[k] -= qa.r[k];
            vbr[k] += qa.vr * r[k];
        }
        rhob += qa.rho_b;
        pb += qa.p;
    });
    if (P.nw > 0) {
        const int nw = wnbr ? wcnt[t] : -1;
        for_each_candidate<ND, kAdjLanes>(P, wnbr, active ? nw : 0, ls, t, Wv.start, P.hs_wall,
                                          a.x, a.y, a.z, active ? q : kAdjLanes, [&](int w) {
            const double4 xw = Wv.pos[w];
            double r[3] = {xc[0] - xw.x, xc[1] - xw.y, (ND == 3) ? xc[2] - xw.z : 0.0};
            const double d2 = dist2<ND>(r[0], r[1], r[2]);
            if (d2 > P.R2) return;
            const double2 pv = Wv.pv[w];
            const double2 cf = coef[w];
            // the Adami sums of w: N_w and V_w depend on p_c, rho_c and x_c
            const double dist = sqrt(d2);  // d2 may be 0 here
            const double W = wendland_c2(P, dist);
            const double ww = 1.0 - 0.5 * P.h_inv * dist;
            const double kdrW = dist < P.two_h ? P.kgrad_c * (ww * ww * ww) : 0.0;
            double gx = -(P.g[0] * r[0] + P.g[1] * r[1]);  // g.(x_w - x_c)
            if constexpr (ND == 3) gx -= P.g[2] * r[2];
            pb += cf.x * W;
            rhob += cf.x * gx * W;
            const double radial = cf.x * ((P.p_offset + p_c) + rho_c * gx) + cf.y;
#pragma unroll
            for (int k = 0; k < ND; k++) xb[k] += -cf.x * rho_c * P.g[k] * W + radial * kdrW * r[k];
            if (d2 < P.az2) return;
            // own fluid-wall term (v_w = 0)
            double vr = v_c[0] * r[0] + v_c[1] * r[1];
            if constexpr (ND == 3) vr += v_c[2] * r[2];
            const double inv_rho_w = pv.y / xw.w, rho_w = xw.w / pv.y;
            const PairGeom pg = pair_geom(P, d2);
            const PairAdj o = pair_adj<ND, false>(P, K, r, pg, v_c, vr, rho_c, inv_rho_c, rho_w,
                                                  inv_rho_w, 0.0, p_c, pv.x, pv.y, xw.w, G, gr);
#pragma unroll
            for (int k = 0; k < ND; k++) {
                xb[k] += o.r[k];
                vbr[k] += o.vr * r[k];
            }
            rhob += o.rho_a;
            pb += o.p;
        });
    }
#pragma unroll
    for (int k = 0; k < ND; k++) {
        xb[k] = group_sum<kAdjLanes>(xb[k]);
        vbr[k] = group_sum<kAdjLanes>(vbr[k]);
    }
    rhob = group_sum<kAdjLanes>(rhob);
    pb = group_sum<kAdjLanes>(pb);
    alpha = group_sum<kAdjLanes>(alpha);
    delta = group_sum<kAdjLanes>(delta);
    if (!active || q != 0) return;
    rhob += pb * dpressure_drho(P, rho_c);
    double gpart = G[0] * K.ghat[0] + G[1] * K.ghat[1];  // gravity: dv += g ghat
    if constexpr (ND == 3) gpart += G[2] * K.ghat[2];
    const int i = F.perm[s];
    double *vi = vb + (size_t)i * (ND + 1);
    double *ui = ub + (size_t)i * ND;
#pragma unroll
    for (int k = 0; k < ND; k++) {
        vi[k] = ACC ? vi[k] + vbr[k] : vbr[k];
        ui[k] = ACC ? ui[k] + xb[k] : xb[k];
    }
    vi[ND] = ACC ? vi[ND] + rhob : rhob;
    th[3 * (size_t)i] += alpha;
    th[3 * (size_t)i + 1] += delta;
    th[3 * (size_t)i + 2] += gpart;
}

__global__ void __launch_bounds__(kBlock) k_adj_fill(double *p, size_t n, double value)
{
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        p[i] = value;
}

// Deterministic sums of the parameter partials: block k sums a fixed strided share of the
// fluid partials th (3 per particle) and of the wall partials gw (added to g) with a fixed
// tree; k_adj_sum_final adds the block results in block order.
__global__ void __launch_bounds__(kBlock)
    k_adj_sum(int nf, const double *__restrict__ th, int nw, const double *__restrict__ gw,
              double *__restrict__ part)
{
    __shared__ double s[3][kBlock];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < nf; i += gridDim.x * kBlock) {
#pragma unroll
        for (int k = 0; k < 3; k++) acc[k] += th[3 * (size_t)i + k];
    }
    for (int w = blockIdx.x * kBlock + threadIdx.x; w < nw; w += gridDim.x * kBlock) acc[2] += gw[w];
#pragma unroll
    for (int k = 0; k < 3; k++) s[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
#pragma unroll
            for (int k = 0; k < 3; k++) s[k][threadIdx.x] += s[k][threadIdx.x + o];
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < 3) part[(size_t)blockIdx.x * 3 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void k_adj_sum_final(int nparts, const double *__restrict__ part, double *__restrict__ out)
{
    if ((int)threadIdx.x >= 3) return;
    double acc = 0.0;
    for (int b = 0; b < nparts; b++) acc += part[(size_t)b * 3 + threadIdx.x];
    out[threadIdx.x] = acc;
}

int check_supported(tpsph_handle *h)
{
    const tpsph_params &p = h->params;
    if (is_host(h))
        return fail(h, TPSPH_ERR_ARGUMENT, "reverse mode needs device pointers (TPSPH_MEM_DEVICE)");
    if (p.adaptive_sound_speed || p.use_density_diffusion == TPSPH_DIFFUSION_ANTUONO ||
        p.wall_viscosity || p.precision != TPSPH_PRECISION_FP64 ||
        p.density_calculator != TPSPH_DENSITY_CONTINUITY)
        return fail(h, TPSPH_ERR_ARGUMENT,
                    "reverse mode supports ContinuityDensity, StateEquationCole with a fixed "
                    "sound speed, Molteni-Colagrossi or no density diffusion, walls without "
                    "viscosity and FP64");
    return TPSPH_OK;
}

AdjConsts adj_consts(const tpsph_handle *h)
{
    AdjConsts K{};
    const DevParams &P = h->P;
    K.dd2_per_delta = P.use_dd ? 2.0 * P.h * P.c : 0.0;  // Molteni-Colagrossi: dd2 = 2 delta h c
    double g2 = 0.0;
    for (int d = 0; d < 3; d++) g2 += P.g[d] * P.g[d];
    const double g = std::sqrt(g2);
    for (int d = 0; d < 3; d++) K.ghat[d] = g > 0.0 ? P.g[d] / g : 0.0;
    return K;
}

int fill(tpsph_handle *h, double *p, size_t n, double value)
{
    if (n == 0) return TPSPH_OK;
    const int grid = (int)std::max<size_t>(1, std::min<size_t>((n + kBlock - 1) / kBlock, 2048));
    hipLaunchKernelGGL(k_adj_fill, dim3(grid), dim3(kBlock), 0, h->stream, p, n, value);
    return launch_check(h, "k_adj_fill");
}

int ensure_buffers(tpsph_handle *h)
{
    if (h->adj_kb) return TPSPH_OK;
    TRY(dalloc(h, &h->adj_kb, (size_t)h->nf));
    TRY(dalloc(h, &h->adj_coef, (size_t)h->nw));
    TRY(dalloc(h, &h->adj_th, 3 * (size_t)h->nf));
    TRY(dalloc(h, &h->adj_gw, (size_t)h->nw));
    TRY(dalloc(h, &h->adj_red, 3 * (size_t)kRedBlocks + 3));
    TRY(dalloc(h, &h->adj_vb, v_len(h)));
    TRY(dalloc(h, &h->adj_ub, u_len(h)));
    TRY(dalloc(h, &h->adj_tvb, v_len(h)));
    TRY(dalloc(h, &h->adj_tub, u_len(h)));
    return TPSPH_OK;
}

int grow(tpsph_handle *h, double **p, size_t *cap, size_t n)
{
    if (*cap >= n) return TPSPH_OK;
    if (*p) HIP_TRY(h, hipFree(*p));
    *p = nullptr;
    *cap = 0;
    TRY(dalloc(h, p, n));
    *cap = n;
    return TPSPH_OK;
}

// The adjoint passes at the state the handle has just prepared: kb must hold the weights.
// Adds (ACC) or writes d L / d (v, u) into vb / ub; the parameter partials go to adj_th /
// adj_gw.
template <int ND, bool ACC>
int adjoint_passes(tpsph_handle *h, double *vb, double *ub)
{
    TimedScope ts(h, TPSPH_TIME_ADJOINT);
    const AdjConsts K = adj_consts(h);
    const bool lists = h->search == TPSPH_SEARCH_LISTS;
    if (h->nw > 0) {
        hipLaunchKernelGGL((k_adj_wall<ND>), dim3(nblocks(kAdjLanes * (int64_t)h->nw)), dim3(kBlock), 0, h->stream,
                           h->P, K, h->fluid_view(), h->wpos, h->pw, h->rhow, h->wpv,
                           lists ? h->anbr : nullptr, lists ? h->acnt : nullptr, h->lw,
                           h->adj_kb, h->adj_coef, h->adj_gw);
        TRY(launch_check(h, "k_adj_wall"));
    }
    hipLaunchKernelGGL((k_adj_fluid<ND, ACC>), dim3(nblocks(kAdjLanes * (int64_t)h->nf)), dim3(kBlock), 0, h->stream,
                       h->flag, h->P, K, h->fluid_view(), h->wall_view(),
                       lists ? h->ord : nullptr, lists ? h->nbr : nullptr,
                       lists ? h->cnt : nullptr, lists ? h->wnbr : nullptr,
                       lists ? h->wcnt : nullptr, h->ls, h->adj_kb, h->adj_coef, vb, ub,
                       h->adj_th);
    return launch_check(h, "k_adj_fluid");
}

int reduce_params(tpsph_handle *h, double *param_bar)
{
    const int grid = std::max(1, std::min(nblocks(std::max(h->nf, h->nw)), kRedBlocks));
    hipLaunchKernelGGL(k_adj_sum, dim3(grid), dim3(kBlock), 0, h->stream, h->nf, h->adj_th, h->nw,
                       h->adj_gw, h->adj_red);
    double *res = h->adj_red + 3 * (size_t)kRedBlocks;
    hipLaunchKernelGGL(k_adj_sum_final, dim3(1), dim3(64), 0, h->stream, grid, h->adj_red, res);
    TRY(launch_check(h, "k_adj_sum"));
    if (param_bar) {
        HIP_TRY(h, hipMemcpyAsync(param_bar, res, 3 * sizeof(double), hipMemcpyDefault, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
    }
    return TPSPH_OK;
}

template <int ND>
int kick_vjp_impl(tpsph_handle *h, const double *dv_bar, const double *v, const double *u,
                  double *v_bar, double *u_bar, double *param_bar)
{
    TRY(prepare_dispatch(h, v, u));
    TRY(fill(h, h->adj_th, 3 * (size_t)h->nf, 0.0));
    TRY(fill(h, h->adj_gw, (size_t)h->nw, 0.0));
    hipLaunchKernelGGL((k_adj_weights<ND, false>), dim3(nblocks(h->nf)), dim3(kBlock), 0,
                       h->stream, h->nf, h->iperm, dv_bar, nullptr, nullptr, nullptr, nullptr,
                       0.0, 0.0, 0.0, h->adj_kb);
    TRY(launch_check(h, "k_adj_weights"));
    TRY((adjoint_passes<ND, false>(h, v_bar, u_bar)));
    h->have_kick = false;
    return reduce_params(h, param_bar);
}

template <int ND>
int reverse_impl(tpsph_handle *h, const double *vT_bar, const double *uT_bar, double *v0_bar,
                 double *u0_bar, double *param_bar)
{
    const size_t nv = v_len(h), nu = u_len(h);
    const int64_t every = h->adj_every, nsteps = h->adj_nsteps;
    const int64_t nseg = (nsteps + every - 1) / every;
    // the stage states of one segment, then the working state of its recomputation
    TRY(grow(h, &h->adj_st, &h->adj_st_cap, (size_t)(5 * every + 1) * (nv + nu)));
    double *st_v = h->adj_st, *st_u = h->adj_st + (size_t)5 * every * nv;
    double *wv = st_u + (size_t)5 * every * nu, *wu = wv + nv;
    HIP_TRY(h, hipMemcpyAsync(h->adj_vb, vT_bar, nv * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(h, hipMemcpyAsync(h->adj_ub, uT_bar, nu * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    TRY(fill(h, h->adj_tvb, nv, 0.0));
    TRY(fill(h, h->adj_tub, nu, 0.0));
    TRY(fill(h, h->adj_th, 3 * (size_t)h->nf, 0.0));
    TRY(fill(h, h->adj_gw, (size_t)h->nw, 0.0));
    for (int64_t k = nseg - 1; k >= 0; k--) {
        const int64_t steps = std::min(every, nsteps - k * every);
        const double *ck = h->adj_ck + (size_t)k * (nv + nu);
        HIP_TRY(h, hipMemcpyAsync(wv, ck, nv * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        HIP_TRY(h, hipMemcpyAsync(wu, ck + nv, nu * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        int64_t n0 = 0, n1 = 0;
        TRY(tpsph_rebuild_count(h, &n0));
        TRY(run_ck2n54(h, wv, wu, h->adj_dt, steps, st_v, st_u));
        TRY(tpsph_rebuild_count(h, &n1));
        int moved = 1;
        HIP_TRY(h, hipMemcpyAsync(&moved, h->flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        // no rebuild in the recomputation and the last stored state (which no kick saw) within
        // its skin: every stored state is within the skin of the current lists, so the reverse
        // stages keep the order
        const bool keep_order = n1 == n0 && moved == 0;
        for (int64_t s = steps - 1; s >= 0; s--) {
            for (int i = 4; i >= 0; i--) {
                const size_t e = (size_t)(5 * s + i);
                const double *yv = st_v + e * nv, *yu = st_u + e * nu;
                TRY(prepare_dispatch(h, yv, yu, keep_order));
                hipLaunchKernelGGL((k_adj_weights<ND, true>), dim3(nblocks(h->nf)), dim3(kBlock), 0,
                                   h->stream, h->nf, h->iperm, nullptr, h->adj_vb, h->adj_ub,
                                   h->adj_tvb, h->adj_tub, kCkA[i], kCkB[i], h->adj_dt, h->adj_kb);
                TRY(launch_check(h, "k_adj_weights"));
                TRY((adjoint_passes<ND, true>(h, h->adj_vb, h->adj_ub)));
            }
        }
    }
    h->have_kick = false;
    HIP_TRY(h, hipMemcpyAsync(v0_bar, h->adj_vb, nv * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(h, hipMemcpyAsync(u0_bar, h->adj_ub, nu * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    return reduce_params(h, param_bar);
}

}  // namespace

extern "C" {

int tpsph_kick_vjp(tpsph_handle *h, const double *dv_bar, const double *v_ode,
                   const double *u_ode, double t, double *v_bar, double *u_bar,
                   double *param_bar)
{
    (void)t;
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    TRY(check_supported(h));
    if (h->nf == 0) {
        if (param_bar) {
            const double zero[3] = {0.0, 0.0, 0.0};
            HIP_TRY(h, hipMemcpy(param_bar, zero, sizeof(zero), hipMemcpyDefault));
        }
        return TPSPH_OK;
    }
    if (!dv_bar || !v_ode || !u_ode || !v_bar || !u_bar)
        return fail(h, TPSPH_ERR_ARGUMENT, "dv_bar, v_ode, u_ode, v_bar and u_bar must not be NULL");
    TRY(ensure_buffers(h));
    return h->nd == 2 ? kick_vjp_impl<2>(h, dv_bar, v_ode, u_ode, v_bar, u_bar, param_bar)
                      : kick_vjp_impl<3>(h, dv_bar, v_ode, u_ode, v_bar, u_bar, param_bar);
}

int tpsph_adjoint_forward(tpsph_handle *h, const double *v0_ode, const double *u0_ode,
                          double t0, double dt, int64_t nsteps, int64_t checkpoint_every,
                          double *v_ode, double *u_ode)
{
    (void)t0;
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    TRY(check_supported(h));
    if (nsteps < 0) return fail(h, TPSPH_ERR_ARGUMENT, "nsteps must be non-negative");
    if (checkpoint_every < 1) return fail(h, TPSPH_ERR_ARGUMENT, "checkpoint_every must be at least 1");
    if (h->nf > 0 && (!v0_ode || !u0_ode || !v_ode || !u_ode))
        return fail(h, TPSPH_ERR_ARGUMENT, "v0_ode, u0_ode, v_ode and u_ode must not be NULL");
    h->adj_ready = false;
    if (h->nf == 0) {
        h->adj_nsteps = nsteps;
        h->adj_every = checkpoint_every;
        h->adj_dt = dt;
        h->adj_ready = true;
        return TPSPH_OK;
    }
    TRY(ensure_buffers(h));
    const size_t nv = v_len(h), nu = u_len(h);
    const int64_t nseg = (nsteps + checkpoint_every - 1) / checkpoint_every;
    TRY(grow(h, &h->adj_ck, &h->adj_ck_cap, (size_t)std::max<int64_t>(nseg, 1) * (nv + nu)));
    if (v_ode != v0_ode)
        HIP_TRY(h, hipMemcpyAsync(v_ode, v0_ode, nv * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    if (u_ode != u0_ode)
        HIP_TRY(h, hipMemcpyAsync(u_ode, u0_ode, nu * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    for (int64_t k = 0; k < nseg; k++) {
        double *ck = h->adj_ck + (size_t)k * (nv + nu);
        HIP_TRY(h, hipMemcpyAsync(ck, v_ode, nv * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        HIP_TRY(h, hipMemcpyAsync(ck + nv, u_ode, nu * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        TRY(run_ck2n54(h, v_ode, u_ode, dt, std::min(checkpoint_every, nsteps - k * checkpoint_every),
                       nullptr, nullptr));
    }
    h->adj_nsteps = nsteps;
    h->adj_every = checkpoint_every;
    h->adj_dt = dt;
    h->adj_ready = true;
    return TPSPH_OK;
}

int tpsph_adjoint_reverse(tpsph_handle *h, const double *v_bar, const double *u_bar,
                          double *v0_bar, double *u0_bar, double *param_bar)
{
    if (!h) return fail(nullptr, TPSPH_ERR_ARGUMENT, "handle is NULL");
    TRY(check_supported(h));
    if (!h->adj_ready)
        return fail(h, TPSPH_ERR_STATE, "no trajectory stored: call tpsph_adjoint_forward first");
    if (h->nf == 0) {
        if (param_bar) {
            const double zero[3] = {0.0, 0.0, 0.0};
            HIP_TRY(h, hipMemcpy(param_bar, zero, sizeof(zero), hipMemcpyDefault));
        }
        return TPSPH_OK;
    }
    if (!v_bar || !u_bar || !v0_bar || !u0_bar)
        return fail(h, TPSPH_ERR_ARGUMENT, "v_bar, u_bar, v0_bar and u0_bar must not be NULL");
    return h->nd == 2 ? reverse_impl<2>(h, v_bar, u_bar, v0_bar, u0_bar, param_bar)
                      : reverse_impl<3>(h, v_bar, u_bar, v0_bar, u0_bar, param_bar);
}

}  // extern "C"
