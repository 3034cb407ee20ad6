/*
 * sph_dual.cpp -- forward-mode derivatives of the oracle's right-hand side and of a
 * CarpenterKennedy2N54 trajectory, by dual numbers.
 *
 * TEST INFRASTRUCTURE ONLY (linked into the oracle library; loaded by tests/ alone).  The
 * reference has no automatic differentiation through its right-hand side (ForwardDiff only
 * appears at src/schemes/boundary/prescribed_motion.jl:96-99 and in
 * test/general/smoothing_kernels.jl:121), so the gradients of the adjoint are checked
 * against this restatement: the kick of sph_oracle.c, with the same expressions in the same
 * order, evaluated on Dual<4> numbers that carry four tangents beside each value.  The value
 * part equals sph_oracle.c's kick and trajectory bit for bit (the neighbours are enumerated
 * by oracle_neighbors_celllist, in the oracle's cell order); the tangents are the
 * directional derivatives along four seeds.
 *
 * Differentiated inputs: Monaghan alpha, Molteni-Colagrossi delta, the gravity magnitude g
 * (gravity = g * P->gravity / |P->gravity|, so g = |P->gravity| reproduces the values) and
 * the initial state (v0, rho0, u0).  Options: WendlandC2, StateEquationCole with a fixed
 * sound speed (clipping as configured), ArtificialViscosityMonaghan, Molteni-Colagrossi
 * diffusion, Adami walls without wall viscosity.
 *
 * Non-smooth points take the branch the values take (as ForwardDiff does): the viscosity
 * switch v_ab.r_ab < 0, the clipping max(p, 0), the volume bound of the Adami pass, and the
 * square root at zero distance (whose tangent is set to zero).
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

extern "C" {
/* the oracle's parameter struct (sph_oracle.c), same layout */
typedef struct {
    int32_t ndims;
    int32_t kernel;
    int32_t clip_fluid_eos;
    int32_t clip_wall_pressure;
    int32_t use_viscosity;
    int32_t use_density_diffusion;
    double smoothing_length;
    double reference_density;
    double sound_speed;
    double exponent;
    double background_pressure;
    double alpha;
    double beta;
    double epsilon;
    double delta;
    double gravity[3];
    double pressure_offset;
    int32_t adaptive_sound_speed;
    int32_t density_calculator;
    double mach_number_target;
    double min_sound_speed;
    double max_sound_speed;
    int32_t wall_viscosity;
    int32_t reserved1;
    double wall_alpha;
    double wall_beta;
    double wall_epsilon;
    double wall_nu;
} oracle_params;

int oracle_neighbors_celllist(int nd, double R, int64_t na, const double *xa, int64_t nb,
                              const double *xb, int64_t *counts, const int64_t *offsets,
                              int64_t *idx);
}

namespace {

constexpr int NT = 4;  // tangents carried per number

struct Dual {
    double v;
    double d[NT];
    Dual() : v(0.0), d{0.0, 0.0, 0.0, 0.0} {}
    Dual(double x) : v(x), d{0.0, 0.0, 0.0, 0.0} {}  // NOLINT: constants promote
};

inline Dual operator+(const Dual &a, const Dual &b)
{
    Dual r;
    r.v = a.v + b.v;
    for (int k = 0; k < NT; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
inline Dual operator-(const Dual &a, const Dual &b)
{
    Dual r;
    r.v = a.v - b.v;
    for (int k = 0; k < NT; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
inline Dual operator-(const Dual &a)
{
    Dual r;
    r.v = -a.v;
    for (int k = 0; k < NT; k++) r.d[k] = -a.d[k];
    return r;
}
inline Dual operator*(const Dual &a, const Dual &b)
{
    Dual r;
    r.v = a.v * b.v;
    for (int k = 0; k < NT; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
inline Dual operator/(const Dual &a, const Dual &b)
{
    Dual r;
    r.v = a.v / b.v;
    for (int k = 0; k < NT; k++) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r;
}
inline Dual &operator+=(Dual &a, const Dual &b) { return a = a + b; }
inline Dual &operator-=(Dual &a, const Dual &b) { return a = a - b; }
inline Dual &operator/=(Dual &a, const Dual &b) { return a = a / b; }

inline double val(const Dual &a) { return a.v; }
inline double val(double a) { return a; }

inline Dual dsqrt(const Dual &a)
{
    Dual r;
    r.v = sqrt(a.v);
    for (int k = 0; k < NT; k++) r.d[k] = r.v > 0.0 ? a.d[k] / (2.0 * r.v) : 0.0;
    return r;
}
inline Dual dpow(const Dual &a, double e)
{
    Dual r;
    r.v = pow(a.v, e);
    const double dd = e * pow(a.v, e - 1.0);
    for (int k = 0; k < NT; k++) r.d[k] = dd * a.d[k];
    return r;
}
// fma(a, b, c) with a constant: value std::fma (the oracle's stage update), tangent a b' + c'
inline Dual dfma(double a, const Dual &b, const Dual &c)
{
    Dual r;
    r.v = fma(a, b.v, c.v);
    for (int k = 0; k < NT; k++) r.d[k] = a * b.d[k] + c.d[k];
    return r;
}

struct Consts {
    int nd;
    double h, R, almostzero;
    double ghat[3];
    const oracle_params *P;
};

double julia_eps(double x) { return nextafter(x, INFINITY) - x; }

double wc2_normalization(int nd, double h_inv)
{
    if (nd == 2) return (7.0 / (4.0 * M_PI)) * (h_inv * h_inv);
    return (21.0 / (16.0 * M_PI)) * (h_inv * h_inv * h_inv);
}

// sph_oracle.c: wc2_kernel_unsafe with the r < 2h guard of oracle_kernel
Dual kernel_w(int nd, const Dual &r, double h)
{
    if (!(val(r) < 2.0 * h)) return Dual(0.0);
    double h_inv = 1.0 / h;
    Dual q = r * Dual(h_inv);
    Dual w = Dual(1.0) - q / Dual(2.0);
    Dual w2 = w * w;
    return Dual(wc2_normalization(nd, h_inv)) * (w2 * w2) * (Dual(2.0) * q + Dual(1.0));
}

// sph_oracle.c: wc2_deriv_div_r_unsafe
Dual kernel_kdr(int nd, const Dual &r, double h)
{
    double h_inv = 1.0 / h;
    Dual q = r * Dual(h_inv);
    Dual w = Dual(1.0) - q / Dual(2.0);
    return Dual(-5.0 * wc2_normalization(nd, h_inv)) * (w * w * w) * Dual(h_inv * h_inv);
}

Dual eos(const oracle_params *P, const Dual &rho)
{
    double c = P->sound_speed;
    double B = P->reference_density * (c * c) / P->exponent;
    Dual p = Dual(B) * (dpow(rho / Dual(P->reference_density), P->exponent) - Dual(1.0)) +
             Dual(P->background_pressure);
    if (P->clip_fluid_eos) return val(p) > 0.0 ? p : Dual(0.0);
    return p;
}

Dual inverse_eos(const oracle_params *P, const Dual &p)
{
    double c = P->sound_speed;
    double B = P->reference_density * (c * c) / P->exponent;
    Dual tmp = (p - Dual(P->background_pressure)) / Dual(B) + Dual(1.0);
    return Dual(P->reference_density) * dpow(tmp, 1.0 / P->exponent);
}

Dual dot3(int nd, const Dual *a, const Dual *b)
{
    Dual s = a[0] * b[0] + a[1] * b[1];
    if (nd == 3) s = s + a[2] * b[2];
    return s;
}

struct Csr {
    std::vector<int64_t> off, idx;
};

// the oracle's neighbour enumeration (cell order) of the value positions
void neighbours(int nd, double R, int64_t na, const std::vector<double> &xa, int64_t nb,
                const std::vector<double> &xb, Csr &c)
{
    std::vector<int64_t> counts((size_t)na + 1, 0);
    c.off.assign((size_t)na + 1, 0);
    oracle_neighbors_celllist(nd, R, na, xa.data(), nb, xb.data(), counts.data(), nullptr, nullptr);
    for (int64_t a = 0; a < na; a++) c.off[a + 1] = c.off[a] + counts[a];
    c.idx.assign((size_t)std::max<int64_t>(c.off[na], 1), 0);
    oracle_neighbors_celllist(nd, R, na, xa.data(), nb, xb.data(), counts.data(), c.off.data(),
                              c.idx.data());
}

// oracle_kick on duals: v ((nd+1) x nf), u (nd x nf) -> dv; alpha, delta, g are duals
void kick(const Consts &K, const Dual &alpha, const Dual &delta, const Dual &g, int64_t nf,
          const std::vector<Dual> &v, const std::vector<Dual> &u, const double *m, int64_t nw,
          const double *xw, const double *mw, std::vector<Dual> &dv)
{
    const oracle_params *P = K.P;
    const int nd = K.nd;
    const double h = K.h, R = K.R;
    std::vector<double> xv((size_t)nf * nd), xwv(xw, xw + nw * nd);
    for (size_t k = 0; k < xv.size(); k++) xv[k] = val(u[k]);
    Dual gvec[3];
    for (int d = 0; d < 3; d++) gvec[d] = g * Dual(K.ghat[d]);
    // compute_pressure!
    std::vector<Dual> pf((size_t)nf);
    for (int64_t a = 0; a < nf; a++) pf[a] = eos(P, v[a * (nd + 1) + nd]);
    // Adami pass
    std::vector<Dual> pw((size_t)nw), rhow((size_t)nw);
    if (nw > 0) {
        Csr c;
        neighbours(nd, R, nw, xwv, nf, xv, c);
#pragma omp parallel for schedule(static)
        for (int64_t w = 0; w < nw; w++) {
            Dual pressure(0.0), vol(0.0);
            for (int64_t k = c.off[w]; k < c.off[w + 1]; k++) {
                const int64_t b = c.idx[k];
                Dual pos_diff[3];
                for (int d = 0; d < nd; d++) pos_diff[d] = Dual(xw[w * nd + d]) - u[b * nd + d];
                Dual d2 = pos_diff[0] * pos_diff[0] + pos_diff[1] * pos_diff[1];
                if (nd == 3) d2 = d2 + pos_diff[2] * pos_diff[2];
                Dual distance = dsqrt(d2);
                Dual density_neighbor = v[b * (nd + 1) + nd];
                Dual fluid_pressure = pf[b];
                Dual resulting_acceleration[3], scaled[3];
                for (int d = 0; d < nd; d++) {
                    resulting_acceleration[d] = gvec[d] - Dual(0.0);
                    scaled[d] = density_neighbor * pos_diff[d];
                }
                Dual hydrostatic_pressure = dot3(nd, resulting_acceleration, scaled);
                Dual sum_pressures = Dual(P->pressure_offset) + fluid_pressure + Dual(0.0) +
                                     hydrostatic_pressure;
                Dual kernel_weight = kernel_w(nd, distance, h);
                pressure += sum_pressures * kernel_weight;
                vol += kernel_weight;
            }
            if (val(vol) > DBL_EPSILON) pressure /= vol;
            if (P->clip_wall_pressure) pressure = val(pressure) > 0.0 ? pressure : Dual(0.0);
            pw[w] = pressure;
            rhow[w] = inverse_eos(P, pressure);
        }
    }
    dv.assign((size_t)nf * (nd + 1), Dual(0.0));
    const double sound_speed = P->sound_speed;
    // fluid-fluid
    {
        Csr c;
        neighbours(nd, R, nf, xv, nf, xv, c);
#pragma omp parallel for schedule(static)
        for (int64_t a = 0; a < nf; a++) {
            Dual p_a = pf[a];
            Dual v_a[3];
            for (int d = 0; d < nd; d++) v_a[d] = v[a * (nd + 1) + d];
            Dual rho_a = v[a * (nd + 1) + nd];
            Dual dv_particle[3], drho_particle(0.0);
            for (int64_t k = c.off[a]; k < c.off[a + 1]; k++) {
                const int64_t b = c.idx[k];
                Dual pos_diff[3];
                for (int d = 0; d < nd; d++) pos_diff[d] = u[a * nd + d] - u[b * nd + d];
                Dual d2 = pos_diff[0] * pos_diff[0] + pos_diff[1] * pos_diff[1];
                if (nd == 3) d2 = d2 + pos_diff[2] * pos_diff[2];
                Dual distance = dsqrt(d2);
                if (val(distance) < K.almostzero) continue;
                Dual kdr = kernel_kdr(nd, distance, h);
                Dual grad_kernel[3];
                for (int d = 0; d < nd; d++) grad_kernel[d] = kdr * pos_diff[d];
                Dual m_b(m[b]);
                Dual v_b[3];
                for (int d = 0; d < nd; d++) v_b[d] = v[b * (nd + 1) + d];
                Dual rho_b = v[b * (nd + 1) + nd];
                Dual p_b = pf[b];
                Dual pfac = -m_b * ((p_a + p_b) / (rho_a * rho_b));
                for (int d = 0; d < nd; d++) dv_particle[d] += (pfac * grad_kernel[d]) * Dual(1.0);
                if (P->use_viscosity) {  // ArtificialViscosityMonaghan (av_monaghan)
                    Dual v_diff[3];
                    for (int i = 0; i < nd; i++) v_diff[i] = v_a[i] - v_b[i];
                    Dual vr = dot3(nd, v_diff, pos_diff);
                    if (val(vr) < 0.0) {
                        double hh = (h + h) / 2.0;
                        Dual rho_mean = (rho_a + rho_b) / Dual(2.0);
                        Dual mu = (Dual(hh) * vr) / (distance * distance + Dual(P->epsilon * (hh * hh)));
                        Dual s = (m_b * alpha * Dual(sound_speed) * mu +
                                  m_b * Dual(P->beta) * (mu * mu)) / rho_mean;
                        for (int i = 0; i < nd; i++) dv_particle[i] += Dual(1.0) * (s * grad_kernel[i]);
                    }
                }
                Dual v_diff[3];
                for (int d = 0; d < nd; d++) v_diff[d] = v_a[d] - v_b[d];
                drho_particle += (rho_a / rho_b) * m_b * dot3(nd, v_diff, grad_kernel);
                if (P->use_density_diffusion) {
                    Dual volume_b = m_b / rho_b;
                    Dual psi_s = (Dual(2.0) * (rho_a - rho_b)) / (distance * distance);
                    Dual psi[3];
                    for (int d = 0; d < nd; d++) psi[d] = psi_s * pos_diff[d];
                    Dual term = volume_b * dot3(nd, psi, grad_kernel);
                    double h_avg = (h + h) / 2.0;
                    drho_particle += delta * Dual(h_avg) * Dual(sound_speed) * term;
                }
            }
            for (int d = 0; d < nd; d++) dv[a * (nd + 1) + d] += dv_particle[d];
            dv[a * (nd + 1) + nd] += drho_particle;
        }
    }
    // fluid-wall
    if (nw > 0) {
        Csr c;
        neighbours(nd, R, nf, xv, nw, xwv, c);
#pragma omp parallel for schedule(static)
        for (int64_t a = 0; a < nf; a++) {
            Dual p_a = pf[a];
            Dual v_a[3];
            for (int d = 0; d < nd; d++) v_a[d] = v[a * (nd + 1) + d];
            Dual rho_a = v[a * (nd + 1) + nd];
            Dual dv_particle[3], drho_particle(0.0);
            for (int64_t k = c.off[a]; k < c.off[a + 1]; k++) {
                const int64_t b = c.idx[k];
                Dual pos_diff[3];
                for (int d = 0; d < nd; d++) pos_diff[d] = u[a * nd + d] - Dual(xw[b * nd + d]);
                Dual d2 = pos_diff[0] * pos_diff[0] + pos_diff[1] * pos_diff[1];
                if (nd == 3) d2 = d2 + pos_diff[2] * pos_diff[2];
                Dual distance = dsqrt(d2);
                if (val(distance) < K.almostzero) continue;
                Dual kdr = kernel_kdr(nd, distance, h);
                Dual grad_kernel[3];
                for (int d = 0; d < nd; d++) grad_kernel[d] = kdr * pos_diff[d];
                Dual m_b(mw[b]);
                Dual rho_b = rhow[b];
                Dual p_b = pw[b];
                Dual pfac = -m_b * ((p_a + p_b) / (rho_a * rho_b));
                for (int d = 0; d < nd; d++) dv_particle[d] += (pfac * grad_kernel[d]) * Dual(1.0);
                Dual v_diff[3];
                for (int d = 0; d < nd; d++) v_diff[d] = v_a[d] - Dual(0.0);
                drho_particle += (rho_a / rho_b) * m_b * dot3(nd, v_diff, grad_kernel);
            }
            for (int d = 0; d < nd; d++) dv[a * (nd + 1) + d] += dv_particle[d];
            dv[a * (nd + 1) + nd] += drho_particle;
        }
    }
    // add_acceleration!
    for (int64_t a = 0; a < nf; a++)
        for (int d = 0; d < nd; d++) dv[a * (nd + 1) + d] += gvec[d];
}

const double CK_A[5] = {0.0, -567301805773.0 / 1357537059087.0,
                        -2404267990393.0 / 2016746695238.0,
                        -3550918686646.0 / 2091501179385.0,
                        -1275806237668.0 / 842570457699.0};
const double CK_B[5] = {1432997174477.0 / 9575080441755.0,
                        5161836677717.0 / 13612068292357.0,
                        1720146321549.0 / 2090206949498.0,
                        3134564353537.0 / 4481467310338.0,
                        2277821191437.0 / 14882151754819.0};

// |gravity|; along one axis exactly the component's magnitude, so that g * ghat reproduces
// the gravity vector bit for bit
double gravity_magnitude(const oracle_params *P)
{
    int nonzero = 0, axis = 0;
    for (int d = 0; d < P->ndims; d++)
        if (P->gravity[d] != 0.0) {
            nonzero++;
            axis = d;
        }
    if (nonzero == 1) return fabs(P->gravity[axis]);
    double gn = 0.0;
    for (int d = 0; d < P->ndims; d++) gn += P->gravity[d] * P->gravity[d];
    return sqrt(gn);
}

Consts consts(const oracle_params *P)
{
    Consts K;
    K.P = P;
    K.nd = P->ndims;
    K.h = P->smoothing_length;
    K.R = 2.0 * K.h;
    K.almostzero = sqrt(julia_eps(K.R * K.R));
    const double gn = gravity_magnitude(P);
    for (int d = 0; d < 3; d++) K.ghat[d] = (d < K.nd && gn > 0.0) ? P->gravity[d] / gn : 0.0;
    return K;
}

// params[3 * (NT + 1)]: value and NT tangents of alpha, delta, g
void seeded_params(const oracle_params *P, const double *seed, Dual &alpha, Dual &delta, Dual &g)
{
    alpha = Dual(P->alpha);
    delta = Dual(P->delta);
    g = Dual(gravity_magnitude(P));
    for (int k = 0; k < NT; k++) {
        alpha.d[k] = seed ? seed[0 * NT + k] : 0.0;
        delta.d[k] = seed ? seed[1 * NT + k] : 0.0;
        g.d[k] = seed ? seed[2 * NT + k] : 0.0;
    }
}

void load(std::vector<Dual> &out, const double *x, const double *seed, size_t n)
{
    out.assign(n, Dual(0.0));
    for (size_t i = 0; i < n; i++) {
        out[i].v = x[i];
        for (int k = 0; k < NT; k++) out[i].d[k] = seed ? seed[k * n + i] : 0.0;
    }
}

void store(const std::vector<Dual> &in, double *x, double *tangents)
{
    const size_t n = in.size();
    for (size_t i = 0; i < n; i++) {
        if (x) x[i] = in[i].v;
        if (tangents)
            for (int k = 0; k < NT; k++) tangents[k * n + i] = in[i].d[k];
    }
}

}  // namespace

extern "C" {

int oracle_dual_tangents(void) { return NT; }

/* One kick on duals.  seed_params: 3 x NT tangents of (alpha, delta, g); seed_v, seed_u:
 * NT x (len) tangents of v ((nd+1) x nf) and u (nd x nf), either may be NULL.  Outputs
 * dv (value) and dv_tangents (NT x len(dv)). */
int oracle_dual_kick(const oracle_params *P, int64_t nf, const double *v, const double *u,
                     const double *m, int64_t nw, const double *xw, const double *mw,
                     const double *seed_params, const double *seed_v, const double *seed_u,
                     double *dv, double *dv_tangents)
{
    const Consts K = consts(P);
    const int nd = K.nd;
    Dual alpha, delta, g;
    seeded_params(P, seed_params, alpha, delta, g);
    std::vector<Dual> vv, uu, out;
    load(vv, v, seed_v, (size_t)nf * (nd + 1));
    load(uu, u, seed_u, (size_t)nf * nd);
    kick(K, alpha, delta, g, nf, vv, uu, m, nw, xw, mw, out);
    store(out, dv, dv_tangents);
    return 0;
}

/* nsteps of CarpenterKennedy2N54 (oracle_solve_ck2n54) on duals from (v, u), seeded as in
 * oracle_dual_kick; v, u are advanced in place (values), v_tangents / u_tangents (NT x len,
 * may be NULL) receive the tangents of the final state.  With u_target (nd x nf) the loss
 * sum_a |u_a - target_a|^2 / nf and its NT tangents are returned in loss[0..NT]. */
int oracle_dual_solve_ck2n54(const oracle_params *P, int64_t nf, double *v, double *u,
                             const double *m, int64_t nw, const double *xw, const double *mw,
                             double dt, int64_t nsteps, const double *seed_params,
                             const double *seed_v, const double *seed_u, double *v_tangents,
                             double *u_tangents, const double *u_target, double *loss)
{
    const Consts K = consts(P);
    const int nd = K.nd;
    Dual alpha, delta, g;
    seeded_params(P, seed_params, alpha, delta, g);
    const size_t nv = (size_t)nf * (nd + 1), nu = (size_t)nf * nd;
    std::vector<Dual> vv, uu, kv, ku(nu), tv(nv), tu(nu);
    load(vv, v, seed_v, nv);
    load(uu, u, seed_u, nu);
    for (int64_t s = 0; s < nsteps; s++) {
        for (int i = 0; i < 5; i++) {
            kick(K, alpha, delta, g, nf, vv, uu, m, nw, xw, mw, kv);
            for (int64_t a = 0; a < nf; a++)  // oracle_drift
                for (int d = 0; d < nd; d++) ku[a * nd + d] = vv[a * (nd + 1) + d] + Dual(0.0);
            if (i == 0) {
                for (size_t k = 0; k < nv; k++) { tv[k] = Dual(dt) * kv[k]; vv[k] = dfma(CK_B[0], tv[k], vv[k]); }
                for (size_t k = 0; k < nu; k++) { tu[k] = Dual(dt) * ku[k]; uu[k] = dfma(CK_B[0], tu[k], uu[k]); }
            } else {
                for (size_t k = 0; k < nv; k++) {
                    tv[k] = dfma(dt, kv[k], Dual(CK_A[i]) * tv[k]);
                    vv[k] = dfma(CK_B[i], tv[k], vv[k]);
                }
                for (size_t k = 0; k < nu; k++) {
                    tu[k] = dfma(dt, ku[k], Dual(CK_A[i]) * tu[k]);
                    uu[k] = dfma(CK_B[i], tu[k], uu[k]);
                }
            }
        }
    }
    store(vv, v, v_tangents);
    store(uu, u, u_tangents);
    if (u_target && loss) {
        Dual L(0.0);
        for (size_t k = 0; k < nu; k++) {
            Dual e = uu[k] - Dual(u_target[k]);
            L += e * e;
        }
        L = L / Dual((double)nf);
        loss[0] = L.v;
        for (int k = 0; k < NT; k++) loss[1 + k] = L.d[k];
    }
    return 0;
}

}  // extern "C"
