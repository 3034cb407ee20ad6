/*
 * sph_oracle.c -- CPU FP64 restatement of TrixiParticles.jl's weakly compressible SPH
 * right-hand side (WCSPH fluid + dummy-particle walls with Adami pressure extrapolation).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU baseline.  It is
 * loaded by the test-suite, by the smoke check and by the bench's cpu_baseline leg, and
 * nowhere else.  The product path (the HIP library) never calls into it.
 *
 * Every expression follows the reference expression by expression, evaluated left to right
 * with no FMA contraction (build with -ffp-contract=off), as Julia evaluates it:
 *   - `x^2`, `x^3` with literal exponents are repeated multiplications (Julia literal_pow);
 *     `x^4` under `@fastpow` is (x*x)*(x*x);
 *   - `div_fast(a, b)` is the exact IEEE quotient a / b for Float64 on the CPU
 *     (test/examples/gpu.jl:31-37);
 *   - `dot(a, b)` of static vectors is ((a1*b1 + a2*b2) + a3*b3);
 *   - `Float64^Float64` is the C library pow().
 *
 * Neighbour iteration: for target a, all points b of the neighbour set with
 * |x_a - x_b|^2 <= R^2, enumerated through a uniform cell grid of cell size R (larger only
 * for boxes far wider than the point cloud; cells walked z-outer, y, x-inner; inside a
 * cell in ascending index).  The squared
 * distance is ((dx*dx + dy*dy) + dz*dz) without FMA; the HIP path evaluates the radius
 * test with the same operation order so neighbour sets agree bit for bit.
 *
 * Threading: OpenMP static schedule over target particles, the analogue of the
 * reference's default PolyesterBackend static chunking (semidiscretization.jl:24-26).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Parameters of one fluid system + one dummy-particle wall system.  The layout is mirrored
 * by a ctypes Structure in oracle/oracle.py. */
typedef struct {
    int32_t ndims;
    int32_t kernel;               /* 0 = WendlandC2Kernel, 1 = SchoenbergCubicSplineKernel */
    int32_t clip_fluid_eos;       /* StateEquationCole(clip_negative_pressure=...) */
    int32_t clip_wall_pressure;   /* BoundaryModelDummyParticles(clip_negative_pressure=...) */
    int32_t use_viscosity;        /* fluid viscosity = ArtificialViscosityMonaghan, else nothing */
    int32_t use_density_diffusion;/* 0 nothing, 1 DensityDiffusionMolteniColagrossi,
                                     2 DensityDiffusionAntuono */
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
    double pressure_offset;       /* AdamiPressureExtrapolation(pressure_offset=...) */
    int32_t adaptive_sound_speed; /* StateEquationAdaptiveCole: sound_speed set per kick */
    int32_t density_calculator;   /* 0 = ContinuityDensity, 1 = SummationDensity */
    double mach_number_target;
    double min_sound_speed;
    double max_sound_speed;
    int32_t wall_viscosity;       /* wall model viscosity: 0 nothing, 1 Monaghan, 2 Adami */
    int32_t reserved1;
    double wall_alpha;
    double wall_beta;
    double wall_epsilon;
    double wall_nu;
} oracle_params;

/* ---------------------------------------------------------------------------------------
 * Smoothing kernels (src/general/smoothing_kernels.jl)
 * ------------------------------------------------------------------------------------- */

/* WendlandC2 normalisation, smoothing_kernels.jl:453-458 */
static double wc2_normalization(int nd, double h_inv)
{
    if (nd == 2) return (7.0 / (4.0 * M_PI)) * (h_inv * h_inv);
    return (21.0 / (16.0 * M_PI)) * (h_inv * h_inv * h_inv);
}

/* SchoenbergCubicSpline normalisation, smoothing_kernels.jl:217-227 */
static double cubic_normalization(int nd, double h_inv)
{
    if (nd == 1) return (2.0 / 3.0) * h_inv;
    if (nd == 2) return (10.0 / (M_PI * 7.0)) * (h_inv * h_inv);
    return (1.0 / M_PI) * (h_inv * h_inv * h_inv);
}

/* kernel_unsafe(::WendlandC2Kernel), smoothing_kernels.jl:436-441 (under @fastpow) */
static double wc2_kernel_unsafe(int nd, double r, double h)
{
    double h_inv = 1.0 / h;
    double q = r * h_inv;
    double w = 1.0 - q / 2.0;
    double w2 = w * w;
    return wc2_normalization(nd, h_inv) * (w2 * w2) * (2.0 * q + 1.0);
}

/* kernel_deriv_div_r_unsafe(::WendlandC2Kernel), smoothing_kernels.jl:443-448 */
static double wc2_deriv_div_r_unsafe(int nd, double r, double h)
{
    double h_inv = 1.0 / h;
    double q = r * h_inv;
    double w = 1.0 - q / 2.0;
    return -5.0 * wc2_normalization(nd, h_inv) * (w * w * w) * (h_inv * h_inv);
}

/* kernel_unsafe(::SchoenbergCubicSplineKernel), smoothing_kernels.jl:193-200 */
static double cubic_kernel_unsafe(int nd, double r, double h)
{
    double h_inv = 1.0 / h;
    double q = r * h_inv;
    double a = 2.0 - q, b = 1.0 - q;
    double result = 1.0 * (a * a * a) / 4.0 - (double)(q < 1.0) * (b * b * b);
    return cubic_normalization(nd, h_inv) * result;
}

/* kernel_deriv_div_r_unsafe(::SchoenbergCubicSplineKernel), smoothing_kernels.jl:202-213 */
static double cubic_deriv_div_r_unsafe(int nd, double r, double h)
{
    double h_inv = 1.0 / h;
    double q = r * h_inv;
    double a = 2.0 - q, b = 1.0 - q;
    double result = -3.0 * (a * a) / 4.0 + 3.0 * (double)(q < 1.0) * (b * b);
    double kernel_deriv = cubic_normalization(nd, h_inv) * result * h_inv;
    return kernel_deriv / r;
}

static double kernel_unsafe(int which, int nd, double r, double h)
{
    return which == 1 ? cubic_kernel_unsafe(nd, r, h) : wc2_kernel_unsafe(nd, r, h);
}

static double deriv_div_r_unsafe(int which, int nd, double r, double h)
{
    return which == 1 ? cubic_deriv_div_r_unsafe(nd, r, h) : wc2_deriv_div_r_unsafe(nd, r, h);
}

/* kernel(kernel, r, h) with the compact-support guard r < 2h, smoothing_kernels.jl:30-34 */
double oracle_kernel(int which, int nd, double r, double h)
{
    double cs = 2.0 * h;
    return r < cs ? kernel_unsafe(which, nd, r, h) : 0.0;
}

/* kernel_deriv(kernel, r, h), smoothing_kernels.jl:36-45 */
double oracle_kernel_deriv(int which, int nd, double r, double h)
{
    double cs = 2.0 * h;
    double eps_h2 = nextafter(h * h, INFINITY) - h * h; /* Julia eps(h^2) */
    if (r < cs && r * r > eps_h2) return deriv_div_r_unsafe(which, nd, r, h) * r;
    return 0.0;
}

/* kernel_deriv_div_r_unsafe, exposed for tests */
double oracle_kernel_deriv_div_r(int which, int nd, double r, double h)
{
    return deriv_div_r_unsafe(which, nd, r, h);
}

/* ---------------------------------------------------------------------------------------
 * StateEquationCole (src/schemes/fluid/weakly_compressible_sph/state_equations.jl:115-139)
 * ------------------------------------------------------------------------------------- */
double oracle_state_equation(double c, double rho0, double exponent, double p_bg, int clip,
                             double density)
{
    double B = rho0 * (c * c) / exponent;
    double pressure = B * (pow(density / rho0, exponent) - 1.0) + p_bg;
    if (clip) return pressure > 0.0 ? pressure : 0.0; /* max(0, pressure) */
    return pressure;
}

double oracle_inverse_state_equation(double c, double rho0, double exponent, double p_bg,
                                     double pressure)
{
    double B = rho0 * (c * c) / exponent;
    double tmp = (pressure - p_bg) / B + 1.0;
    return rho0 * pow(tmp, 1.0 / exponent);
}

static double eos(const oracle_params *P, double rho)
{
    return oracle_state_equation(P->sound_speed, P->reference_density, P->exponent,
                                 P->background_pressure, P->clip_fluid_eos, rho);
}

static double inverse_eos(const oracle_params *P, double p)
{
    return oracle_inverse_state_equation(P->sound_speed, P->reference_density, P->exponent,
                                         P->background_pressure, p);
}

/* Julia eps(x) for positive normal x: the spacing to the next larger Float64 */
static double julia_eps(double x) { return nextafter(x, INFINITY) - x; }

static double dot3(int nd, const double *a, const double *b)
{
    double s = a[0] * b[0] + a[1] * b[1];
    if (nd == 3) s = s + a[2] * b[2];
    return s;
}

/* ArtificialViscosityMonaghan functor, viscosity.jl:96-132.  Adds into dv[0..nd). */
static void av_monaghan(int nd, double alpha, double beta, double epsilon, double c,
                        double hl, const double *pos_diff, double distance, double m_b,
                        double rho_a, double rho_b, const double *v_a, const double *v_b,
                        const double *grad_kernel, double *dv)
{
    double v_diff[3] = {0, 0, 0};
    for (int i = 0; i < nd; i++) v_diff[i] = v_a[i] - v_b[i];
    double vr = dot3(nd, v_diff, pos_diff);
    if (vr < 0.0) {
        double h = (hl + hl) / 2.0;
        double rho_mean = (rho_a + rho_b) / 2.0;
        double mu = (h * vr) / (distance * distance + epsilon * (h * h));
        double s = (m_b * alpha * c * mu + m_b * beta * (mu * mu)) / rho_mean;
        for (int i = 0; i < nd; i++) dv[i] += 1.0 * (s * grad_kernel[i]);
    }
}

void oracle_av_monaghan(const oracle_params *P, const double *pos_diff, double distance,
                        double m_b, double rho_a, double rho_b, const double *v_a,
                        const double *v_b, const double *grad_kernel, double *dv)
{
    av_monaghan(P->ndims, P->alpha, P->beta, P->epsilon, P->sound_speed, P->smoothing_length,
                pos_diff, distance, m_b, rho_a, rho_b, v_a, v_b, grad_kernel, dv);
}

/* ViscosityAdami (viscosity.jl:219-280): adami_viscosity_force! with the kinematic
 * viscosities nu_a, nu_b of the two sides, h the mean smoothing length. */
void oracle_av_adami(int nd, double nu_a, double nu_b, double epsilon, double h,
                     const double *pos_diff, double distance, const double *grad_kernel,
                     double m_a, double m_b, double rho_a, double rho_b, const double *v_a,
                     const double *v_b, double *dv)
{
    double v_diff[3] = {0, 0, 0};
    for (int i = 0; i < nd; i++) v_diff[i] = v_a[i] - v_b[i];
    double eta_a = nu_a * rho_a;
    double eta_b = nu_b * rho_b;
    double volume_a = m_a / rho_a;
    double volume_b = m_b / rho_b;
    double tmp = (2 * eta_a * eta_b) /
                 ((eta_a + eta_b) * (distance * distance + epsilon * (h * h)) * m_a);
    double visc = (volume_a * volume_a + volume_b * volume_b) *
                  dot3(nd, grad_kernel, pos_diff) * tmp;
    for (int i = 0; i < nd; i++) dv[i] += 1.0 * visc * v_diff[i];
}

/* kinematic_viscosity(system, ::ArtificialViscosityMonaghan, h, c) (viscosity.jl:82-87) */
static double kinematic_viscosity_monaghan(const oracle_params *P)
{
    return P->alpha * P->smoothing_length * P->sound_speed / (2 * P->ndims + 4);
}

/* ---------------------------------------------------------------------------------------
 * Cell grid (stand-in for PointNeighbors' GridNeighborhoodSearch; cell size R, larger
 * only for boxes that would need more cells than memory holds).
 * ------------------------------------------------------------------------------------- */
typedef struct {
    int nd;
    double R, R2;
    double cs;      /* cell size: R, or larger where the box would need too many cells */
    double lo[3];
    int64_t n[3];
    int64_t ncells;
    int64_t *start; /* ncells + 1 */
    int64_t *items; /* point indices sorted by cell, ascending inside a cell */
} grid_t;

static int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

static void cell_of(const grid_t *g, const double *x, int64_t *c)
{
    for (int d = 0; d < g->nd; d++) {
        double f = floor((x[d] - g->lo[d]) / g->cs);
        int64_t ci = (f < -1.0e15) ? 0 : (f > 1.0e15 ? g->n[d] - 1 : (int64_t)f);
        c[d] = clampi(ci, 0, g->n[d] - 1);
    }
    if (g->nd == 2) c[2] = 0;
}

static int64_t lin(const grid_t *g, const int64_t *c) { return c[0] + g->n[0] * (c[1] + g->n[1] * c[2]); }

/* Build a grid over `np` points x (column-major nd x np), bounding box given by lo/hi. */
static int grid_build(grid_t *g, int nd, double R, const double *lo, const double *hi,
                      int64_t np, const double *x)
{
    g->nd = nd;
    g->R = R;
    g->R2 = R * R;
    /* Cells of size >= R keep the 3^D sweep exact.  A box far larger than the point cloud
     * (outliers) would need more cells than memory holds: then the cell size grows until
     * the grid has at most max(2^22, 8 np) cells.  Clamping an axis to 4096 cells is exact
     * as well: points beyond the last cell share it (see cell_of). */
    const double max_cells = fmax(4194304.0, 8.0 * (double)np);
    g->cs = R;
    for (;;) {
        double total = 1.0;
        for (int d = 0; d < nd; d++) {
            double span = (hi[d] + R) - (lo[d] - R);
            double n = ceil(span / g->cs) + 1.0;
            total *= fmin(fmax(n, 1.0), 4096.0);
        }
        if (total <= max_cells) break;
        g->cs *= 1.25;
    }
    g->ncells = 1;
    for (int d = 0; d < 3; d++) {
        if (d < nd) {
            g->lo[d] = lo[d] - R;
            double span = (hi[d] + R) - g->lo[d];
            int64_t n = (int64_t)ceil(span / g->cs) + 1;
            if (n < 1) n = 1;
            if (n > 4096) n = 4096;
            g->n[d] = n;
        } else {
            g->lo[d] = 0.0;
            g->n[d] = 1;
        }
        g->ncells *= g->n[d];
    }
    g->start = (int64_t *)calloc((size_t)g->ncells + 1, sizeof(int64_t));
    g->items = (int64_t *)malloc(sizeof(int64_t) * (size_t)(np > 0 ? np : 1));
    int64_t *cellid = (int64_t *)malloc(sizeof(int64_t) * (size_t)(np > 0 ? np : 1));
    if (!g->start || !g->items || !cellid) {
        free(g->start); free(g->items); free(cellid);
        return -1;
    }
    for (int64_t i = 0; i < np; i++) {
        int64_t c[3];
        cell_of(g, x + i * nd, c);
        cellid[i] = lin(g, c);
        g->start[cellid[i] + 1]++;
    }
    for (int64_t k = 0; k < g->ncells; k++) g->start[k + 1] += g->start[k];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)g->ncells);
    if (!fill) { free(g->start); free(g->items); free(cellid); return -1; }
    memcpy(fill, g->start, sizeof(int64_t) * (size_t)g->ncells);
    for (int64_t i = 0; i < np; i++) g->items[fill[cellid[i]]++] = i; /* stable */
    free(fill);
    free(cellid);
    return 0;
}

static void grid_free(grid_t *g)
{
    free(g->start);
    free(g->items);
}

static void bbox(int nd, int64_t n1, const double *x1, int64_t n2, const double *x2,
                 double *lo, double *hi)
{
    for (int d = 0; d < nd; d++) { lo[d] = INFINITY; hi[d] = -INFINITY; }
    for (int64_t i = 0; i < n1; i++)
        for (int d = 0; d < nd; d++) {
            double v = x1[i * nd + d];
            if (v < lo[d]) lo[d] = v;
            if (v > hi[d]) hi[d] = v;
        }
    for (int64_t i = 0; i < n2; i++)
        for (int d = 0; d < nd; d++) {
            double v = x2[i * nd + d];
            if (v < lo[d]) lo[d] = v;
            if (v > hi[d]) hi[d] = v;
        }
    for (int d = 0; d < nd; d++)
        if (!(lo[d] <= hi[d])) { lo[d] = 0.0; hi[d] = 0.0; }
}

/* squared distance in the fixed evaluation order shared with the HIP path */
static double dist2(int nd, const double *pd)
{
    double s = pd[0] * pd[0] + pd[1] * pd[1];
    if (nd == 3) s = s + pd[2] * pd[2];
    return s;
}

/* Visit every neighbour b (points of grid g, coordinates xb) of position xa. */
#define FOREACH_NEIGHBOR(g, xa, xb, ...)                                                    \
    do {                                                                                   \
        int64_t c0_[3];                                                                    \
        cell_of((g), (xa), c0_);                                                           \
        int64_t zlo_ = (g)->nd == 3 ? c0_[2] - 1 : 0, zhi_ = (g)->nd == 3 ? c0_[2] + 1 : 0; \
        for (int64_t cz_ = zlo_; cz_ <= zhi_; cz_++) {                                      \
            if (cz_ < 0 || cz_ >= (g)->n[2]) continue;                                     \
            for (int64_t cy_ = c0_[1] - 1; cy_ <= c0_[1] + 1; cy_++) {                      \
                if (cy_ < 0 || cy_ >= (g)->n[1]) continue;                                 \
                for (int64_t cx_ = c0_[0] - 1; cx_ <= c0_[0] + 1; cx_++) {                  \
                    if (cx_ < 0 || cx_ >= (g)->n[0]) continue;                             \
                    int64_t cc_[3] = {cx_, cy_, cz_};                                      \
                    int64_t l_ = lin((g), cc_);                                            \
                    for (int64_t k_ = (g)->start[l_]; k_ < (g)->start[l_ + 1]; k_++) {     \
                        int64_t b = (g)->items[k_];                                        \
                        double pos_diff[3] = {0, 0, 0};                                    \
                        for (int d_ = 0; d_ < (g)->nd; d_++)                               \
                            pos_diff[d_] = (xa)[d_] - (xb)[b * (g)->nd + d_];              \
                        double d2_ = dist2((g)->nd, pos_diff);                             \
                        if (d2_ > (g)->R2) continue;                                       \
                        double distance = sqrt(d2_);                                       \
                        __VA_ARGS__                                                        \
                    }                                                                      \
                }                                                                          \
            }                                                                              \
        }                                                                                  \
    } while (0)

/* ---------------------------------------------------------------------------------------
 * Neighbour sets (exported for parity tests)
 * ------------------------------------------------------------------------------------- */

/* Brute-force O(na*nb) neighbour counts / lists: b with |xa - xb|^2 <= R^2.
 * If idx is NULL only counts are written. offsets has na+1 entries when idx != NULL. */
void oracle_neighbors_bruteforce(int nd, double R, int64_t na, const double *xa, int64_t nb,
                                 const double *xb, int64_t *counts, const int64_t *offsets,
                                 int64_t *idx)
{
    double R2 = R * R;
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < na; a++) {
        int64_t n = 0;
        for (int64_t b = 0; b < nb; b++) {
            double pd[3] = {0, 0, 0};
            for (int d = 0; d < nd; d++) pd[d] = xa[a * nd + d] - xb[b * nd + d];
            if (dist2(nd, pd) <= R2) {
                if (idx) idx[offsets[a] + n] = b;
                n++;
            }
        }
        counts[a] = n;
    }
}

/* Cell-list neighbour counts / lists (same sets as brute force, cell order). */
int oracle_neighbors_celllist(int nd, double R, int64_t na, const double *xa, int64_t nb,
                              const double *xb, int64_t *counts, const int64_t *offsets,
                              int64_t *idx)
{
    double lo[3], hi[3];
    bbox(nd, na, xa, nb, xb, lo, hi);
    grid_t g;
    if (grid_build(&g, nd, R, lo, hi, nb, xb)) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < na; a++) {
        int64_t n = 0;
        const double *x = xa + a * nd;
        FOREACH_NEIGHBOR(&g, x, xb, {
            (void)distance;
            if (idx) idx[offsets[a] + n] = b;
            n++;
        });
        counts[a] = n;
    }
    grid_free(&g);
    return 0;
}

/* ---------------------------------------------------------------------------------------
 * Fluid pressure: compute_pressure! (wcsph/system.jl:426-438)
 * ------------------------------------------------------------------------------------- */
void oracle_compute_pressure(const oracle_params *P, int64_t n, const double *v,
                             double *pressure)
{
    int nd = P->ndims;
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < n; a++) pressure[a] = eos(P, v[a * (nd + 1) + nd]);
}

/* ---------------------------------------------------------------------------------------
 * Adami pressure extrapolation for the wall system:
 * compute_pressure!(model, ::AdamiPressureExtrapolation) dummy_particles.jl:489-542,
 * boundary_pressure_extrapolation!(Val(true)) :586-606, boundary_pressure_inner! :637-672,
 * compute_adami_density! :547-569.
 *   xw: nd x nw wall coordinates; x: nd x nf fluid coordinates; v: (nd+1) x nf fluid state;
 *   pf: fluid pressures.  Outputs wall pressure, wall density and the volume scratch.
 * ------------------------------------------------------------------------------------- */
int oracle_adami(const oracle_params *P, int64_t nw, const double *xw, int64_t nf,
                 const double *x, const double *v, const double *pf, double *pw,
                 double *rhow, double *volume)
{
    int nd = P->ndims;
    double h = P->smoothing_length;
    double R = 2.0 * h; /* compact_support(model) = 2h, neighborhood_search.jl:134-140 */
    double lo[3], hi[3];
    bbox(nd, nw, xw, nf, x, lo, hi);
    grid_t g;
    if (grid_build(&g, nd, R, lo, hi, nf, x)) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t w = 0; w < nw; w++) {
        double pressure = 0.0, vol = 0.0;
        const double *xa = xw + w * nd;
        FOREACH_NEIGHBOR(&g, xa, x, {
            double density_neighbor = v[b * (nd + 1) + nd];
            double fluid_pressure = pf[b];
            double resulting_acceleration[3] = {0, 0, 0};
            double scaled[3] = {0, 0, 0};
            for (int d = 0; d < nd; d++) {
                resulting_acceleration[d] = P->gravity[d] - 0.0;
                scaled[d] = density_neighbor * pos_diff[d];
            }
            double hydrostatic_pressure = dot3(nd, resulting_acceleration, scaled);
            double sum_pressures = P->pressure_offset + fluid_pressure + 0.0 +
                                   hydrostatic_pressure;
            double kernel_weight = oracle_kernel(P->kernel, nd, distance, h);
            pressure += sum_pressures * kernel_weight;
            vol += kernel_weight;
        });
        if (vol > DBL_EPSILON) pressure /= vol;
        if (P->clip_wall_pressure) pressure = pressure > 0.0 ? pressure : 0.0;
        pw[w] = pressure;
        rhow[w] = inverse_eos(P, pressure);
        if (volume) volume[w] = vol;
    }
    grid_free(&g);
    return 0;
}

/* No-slip wall velocity of every wall particle (a wall model with a viscosity):
 * interpolate_fluid_velocity! sums W v_f over the fluid neighbours
 * (dummy_particles.jl:716-731) and compute_wall_velocity! sets 2 v_wall - sum / volume with
 * v_wall = 0 where volume > eps (dummy_particles.jl:586-600, 738-756).  wvel is nd x nw. */
int oracle_wall_velocity(const oracle_params *P, int64_t nw, const double *xw, int64_t nf,
                         const double *x, const double *v, double *wvel)
{
    int nd = P->ndims;
    double h = P->smoothing_length;
    double R = 2.0 * h;
    double lo[3], hi[3];
    bbox(nd, nw, xw, nf, x, lo, hi);
    grid_t g;
    if (grid_build(&g, nd, R, lo, hi, nf, x)) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t w = 0; w < nw; w++) {
        double vol = 0.0, wv[3] = {0, 0, 0};
        const double *xa = xw + w * nd;
        FOREACH_NEIGHBOR(&g, xa, x, {
            double kernel_weight = oracle_kernel(P->kernel, nd, distance, h);
            vol += kernel_weight;
            for (int d = 0; d < nd; d++) wv[d] += kernel_weight * v[b * (nd + 1) + d];
        });
        for (int d = 0; d < nd; d++)
            wvel[w * nd + d] = (vol > DBL_EPSILON) ? 2 * 0.0 - wv[d] / vol : wv[d];
    }
    grid_free(&g);
    return 0;
}

/* ---------------------------------------------------------------------------------------
 * DensityDiffusionAntuono's update! (density_diffusion.jl:176-211) over the fluid's own
 * neighbours:
 *   1. compute_gradient_correction_matrix! (corrections.jl:329-354): per particle a,
 *      corr_a -= (V_b grad W_ab) * r_ab^T, V_b = m_b / rho_b, skipping zero gradients;
 *   2. correction_matrix_inversion_step! (corrections.jl:425-457): corr_a = inv(corr_a), or
 *      the identity where |det(corr_a)| < 1f-9;
 *   3. the normalised density gradient, summed pair by pair as the reference does:
 *      grad_a += ((-(rho_a - rho_b)) * L_a) * grad W_ab * V_b.
 * grad W is smoothing_kernel_grad (abstract_system.jl:142-155): zero at distance >= R and at
 * distance^2 < eps(h^2).  grad (nd x nf) receives the gradients and corr (nd*nd per
 * particle, column-major), if not NULL, the inverted matrices.
 * ------------------------------------------------------------------------------------- */
static void smoothing_kernel_grad(const oracle_params *P, int nd, const double *pos_diff,
                                  double distance, double *grad_kernel)
{
    double h = P->smoothing_length;
    for (int d = 0; d < 3; d++) grad_kernel[d] = 0.0;
    if (distance >= 2.0 * h || distance * distance < julia_eps(h * h)) return;
    double kdr = deriv_div_r_unsafe(P->kernel, nd, distance, h);
    for (int d = 0; d < nd; d++) grad_kernel[d] = kdr * pos_diff[d];
}

/* inv(L) of a 2x2 or 3x3 matrix (column-major) by its adjugate over the determinant,
 * which is returned (the caller tests it against the singularity bound). */
static double invert_small(int nd, const double *L, double *inv)
{
    if (nd == 2) {
        double det = L[0] * L[3] - L[2] * L[1];
        inv[0] = L[3] / det;
        inv[1] = -L[1] / det;
        inv[2] = -L[2] / det;
        inv[3] = L[0] / det;
        return det;
    }
    /* columns x0, x1, x2; det = x0 . (x1 x x2); the rows of the inverse are
     * (x1 x x2, x2 x x0, x0 x x1) / det */
    const double *x0 = L, *x1 = L + 3, *x2 = L + 6;
    double y0[3] = {x1[1] * x2[2] - x1[2] * x2[1], x1[2] * x2[0] - x1[0] * x2[2],
                    x1[0] * x2[1] - x1[1] * x2[0]};
    double y1[3] = {x2[1] * x0[2] - x2[2] * x0[1], x2[2] * x0[0] - x2[0] * x0[2],
                    x2[0] * x0[1] - x2[1] * x0[0]};
    double y2[3] = {x0[1] * x1[2] - x0[2] * x1[1], x0[2] * x1[0] - x0[0] * x1[2],
                    x0[0] * x1[1] - x0[1] * x1[0]};
    double det = (x0[0] * y0[0] + x0[1] * y0[1]) + x0[2] * y0[2];
    for (int j = 0; j < 3; j++) {
        inv[0 + 3 * j] = y0[j] / det;
        inv[1 + 3 * j] = y1[j] / det;
        inv[2 + 3 * j] = y2[j] / det;
    }
    return det;
}

int oracle_antuono_gradient(const oracle_params *P, int64_t nf, const double *x,
                            const double *v, const double *m, double *grad, double *corr_out)
{
    int nd = P->ndims, nn = nd * nd;
    double R = 2.0 * P->smoothing_length;
    double lo[3], hi[3];
    bbox(nd, nf, x, 0, NULL, lo, hi);
    grid_t g;
    if (grid_build(&g, nd, R, lo, hi, nf, x)) return -1;
    double *corr = corr_out ? corr_out : (double *)malloc(sizeof(double) * (size_t)(nf * nn + 1));
    if (!corr) {
        grid_free(&g);
        return -1;
    }
    const double singular = (double)1.0e-9f; /* the Float32 literal 1.0f-9 */
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < nf; a++) {
        const double *xa = x + a * nd;
        double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        FOREACH_NEIGHBOR(&g, xa, x, {
            double grad_kernel[3];
            smoothing_kernel_grad(P, nd, pos_diff, distance, grad_kernel);
            int zero = 1;
            for (int d = 0; d < nd; d++) zero &= grad_kernel[d] == 0.0;
            if (zero) continue;
            double volume = m[b] / v[b * (nd + 1) + nd];
            for (int j = 0; j < nd; j++)
                for (int i = 0; i < nd; i++)
                    L[i + nd * j] -= (volume * grad_kernel[i]) * pos_diff[j];
        });
        double inv[9];
        double det = invert_small(nd, L, inv);
        for (int k = 0; k < nn; k++)
            corr[a * nn + k] = fabs(det) < singular ? (k % (nd + 1) == 0 ? 1.0 : 0.0) : inv[k];
    }
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < nf; a++) {
        const double *xa = x + a * nd;
        const double *La = corr + a * nn;
        double rho_a = v[a * (nd + 1) + nd];
        double acc[3] = {0, 0, 0};
        FOREACH_NEIGHBOR(&g, xa, x, {
            double grad_kernel[3];
            smoothing_kernel_grad(P, nd, pos_diff, distance, grad_kernel);
            double rho_b = v[b * (nd + 1) + nd];
            double volume_b = m[b] / rho_b;
            double s = -(rho_a - rho_b);
            for (int i = 0; i < nd; i++) {
                double t = (s * La[i]) * grad_kernel[0];
                for (int j = 1; j < nd; j++) t = t + (s * La[i + nd * j]) * grad_kernel[j];
                acc[i] += t * volume_b;
            }
        });
        for (int d = 0; d < nd; d++) grad[a * nd + d] = acc[d];
    }
    if (!corr_out) free(corr);
    grid_free(&g);
    return 0;
}

/* The scalar of the pressure acceleration (dv_pressure = factor * grad W) of the density
 * calculator's default formulation (choose_pressure_acceleration_formulation,
 * pressure_acceleration.jl:139-154):
 *   ContinuityDensity  pressure_acceleration_continuity_density, -m_b (p_a + p_b)/(rho_a rho_b)
 *                      (pressure_acceleration.jl:34-41)
 *   SummationDensity   pressure_acceleration_summation_density,
 *                      -m_b (p_a / rho_a^2 + p_b / rho_b^2) (pressure_acceleration.jl:8-15) */
static double pressure_factor(const oracle_params *P, double m_b, double rho_a, double rho_b,
                              double p_a, double p_b)
{
    if (P->density_calculator == 1) return -m_b * (p_a / (rho_a * rho_a) + p_b / (rho_b * rho_b));
    return -m_b * ((p_a + p_b) / (rho_a * rho_b));
}

/* ---------------------------------------------------------------------------------------
 * WCSPH interact! (weakly_compressible_sph/rhs.jl:5-127) with ContinuityDensity, no
 * correction, pressure_acceleration_continuity_density (pressure_acceleration.jl:34-41),
 * ArtificialViscosityMonaghan (viscosity.jl:96-132), continuity_equation! (fluid.jl:160-186),
 * DensityDiffusionMolteniColagrossi (density_diffusion.jl:49-55, 213-240).
 *
 * Fluid-fluid: neighbour = the fluid itself.  dv ((nd+1) x nf) is accumulated into (+=).
 * ------------------------------------------------------------------------------------- */
int oracle_interact_fluid_fluid(const oracle_params *P, int64_t nf, const double *x,
                                const double *v, const double *m, const double *pf,
                                double *dv, const double *ddgrad)
{
    int nd = P->ndims;
    double h = P->smoothing_length;
    double R = 2.0 * h;
    double almostzero = sqrt(julia_eps(R * R));
    double sound_speed = P->sound_speed;
    double lo[3], hi[3];
    bbox(nd, nf, x, 0, NULL, lo, hi);
    grid_t g;
    if (grid_build(&g, nd, R, lo, hi, nf, x)) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < nf; a++) {
        const double *xa = x + a * nd;
        double p_a = pf[a];
        double v_a[3] = {0, 0, 0};
        for (int d = 0; d < nd; d++) v_a[d] = v[a * (nd + 1) + d];
        double rho_a = v[a * (nd + 1) + nd];
        double dv_particle[3] = {0, 0, 0};
        double drho_particle = 0.0;
        FOREACH_NEIGHBOR(&g, xa, x, {
            if (distance < almostzero) continue;
            double kdr = deriv_div_r_unsafe(P->kernel, nd, distance, h);
            double grad_kernel[3] = {0, 0, 0};
            for (int d = 0; d < nd; d++) grad_kernel[d] = kdr * pos_diff[d];
            double m_b = m[b];
            double v_b[3] = {0, 0, 0};
            for (int d = 0; d < nd; d++) v_b[d] = v[b * (nd + 1) + d];
            double rho_b = v[b * (nd + 1) + nd];
            double p_b = pf[b];
            double pfac = pressure_factor(P, m_b, rho_a, rho_b, p_a, p_b);
            for (int d = 0; d < nd; d++) dv_particle[d] += (pfac * grad_kernel[d]) * 1.0;
            /* dv_viscosity! -> ArtificialViscosityMonaghan (fluid neighbour) */
            if (P->use_viscosity)
                oracle_av_monaghan(P, pos_diff, distance, m_b, rho_a, rho_b, v_a, v_b,
                                   grad_kernel, dv_particle);
            /* continuity_equation!: nothing with SummationDensity (fluid.jl:133-139) */
            if (P->density_calculator == 1) continue;
            double v_diff[3] = {0, 0, 0};
            for (int d = 0; d < nd; d++) v_diff[d] = v_a[d] - v_b[d];
            drho_particle += (rho_a / rho_b) * m_b * dot3(nd, v_diff, grad_kernel);
            /* density_diffusion! (same system) */
            if (P->use_density_diffusion) {
                double volume_b = m_b / rho_b;
                double psi_s = (2.0 * (rho_a - rho_b)) / (distance * distance);
                if (P->use_density_diffusion == 2) {
                    /* DensityDiffusionAntuono psi (density_diffusion.jl:155-174) */
                    double result = 2.0 * (rho_a - rho_b);
                    double gsum[3] = {0, 0, 0};
                    for (int d = 0; d < nd; d++) gsum[d] = ddgrad[a * nd + d] + ddgrad[b * nd + d];
                    result -= dot3(nd, gsum, pos_diff);
                    psi_s = result / (distance * distance);
                }
                double psi[3] = {0, 0, 0};
                for (int d = 0; d < nd; d++) psi[d] = psi_s * pos_diff[d];
                double term = volume_b * dot3(nd, psi, grad_kernel);
                double h_avg = (h + h) / 2.0;
                drho_particle += P->delta * h_avg * sound_speed * term;
            }
        });
        for (int d = 0; d < nd; d++) dv[a * (nd + 1) + d] += dv_particle[d];
        dv[a * (nd + 1) + nd] += drho_particle;
    }
    grid_free(&g);
    return 0;
}

/* Fluid-wall: neighbour = WallBoundarySystem with BoundaryModelDummyParticles.  The wall
 * velocity is zero (wall_boundary/system.jl:107-123), rho_b / p_b come from the Adami pass,
 * the viscosity model is the wall model's (fluid.jl:109-112): nothing, or Monaghan / Adami
 * on the no-slip velocity wvel (viscous_velocity, wall_boundary/system.jl:148-163; nd x nw,
 * with the fluid masses m for m_a), and density diffusion only acts between identical
 * systems (fluid.jl:179-183). */
int oracle_interact_fluid_wall(const oracle_params *P, int64_t nf, const double *x,
                               const double *v, const double *pf, int64_t nw,
                               const double *xw, const double *mw, const double *pw,
                               const double *rhow, double *dv, const double *m,
                               const double *wvel)
{
    int nd = P->ndims;
    double h = P->smoothing_length;
    double R = 2.0 * h;
    double almostzero = sqrt(julia_eps(R * R));
    double lo[3], hi[3];
    bbox(nd, nf, x, nw, xw, lo, hi);
    grid_t g;
    if (grid_build(&g, nd, R, lo, hi, nw, xw)) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < nf; a++) {
        const double *xa = x + a * nd;
        double p_a = pf[a];
        double v_a[3] = {0, 0, 0};
        for (int d = 0; d < nd; d++) v_a[d] = v[a * (nd + 1) + d];
        double rho_a = v[a * (nd + 1) + nd];
        double dv_particle[3] = {0, 0, 0};
        double drho_particle = 0.0;
        FOREACH_NEIGHBOR(&g, xa, xw, {
            if (distance < almostzero) continue;
            double kdr = deriv_div_r_unsafe(P->kernel, nd, distance, h);
            double grad_kernel[3] = {0, 0, 0};
            for (int d = 0; d < nd; d++) grad_kernel[d] = kdr * pos_diff[d];
            double m_b = mw[b];
            double rho_b = rhow[b];
            double p_b = pw[b];
            double pfac = pressure_factor(P, m_b, rho_a, rho_b, p_a, p_b);
            for (int d = 0; d < nd; d++) dv_particle[d] += (pfac * grad_kernel[d]) * 1.0;
            if (P->wall_viscosity && wvel) {
                const double *v_w = wvel + b * nd;
                if (P->wall_viscosity == 1)
                    av_monaghan(nd, P->wall_alpha, P->wall_beta, P->wall_epsilon,
                                P->sound_speed, h, pos_diff, distance, m_b, rho_a, rho_b, v_a,
                                v_w, grad_kernel, dv_particle);
                else
                    oracle_av_adami(nd, kinematic_viscosity_monaghan(P), P->wall_nu,
                                    P->wall_epsilon, (h + h) / 2, pos_diff, distance,
                                    grad_kernel, m[a], m_b, rho_a, rho_b, v_a, v_w,
                                    dv_particle);
            }
            if (P->density_calculator == 1) continue;
            double v_diff[3] = {0, 0, 0};
            for (int d = 0; d < nd; d++) v_diff[d] = v_a[d] - 0.0;
            drho_particle += (rho_a / rho_b) * m_b * dot3(nd, v_diff, grad_kernel);
        });
        for (int d = 0; d < nd; d++) dv[a * (nd + 1) + d] += dv_particle[d];
        dv[a * (nd + 1) + nd] += drho_particle;
    }
    grid_free(&g);
    return 0;
}

/* ---------------------------------------------------------------------------------------
 * kick! (semidiscretization.jl:589-612) for the system pair (fluid, wall):
 *   set_zero!(dv); fluid EOS; Adami wall pressure; fluid-fluid; fluid-wall; gravity.
 * Optional outputs (may be NULL): fluid pressure, wall pressure, wall density.
 * ------------------------------------------------------------------------------------- */
/* update_speed_of_sound! (wcsph/system.jl:307-320): the sound speed of an adaptive state
 * equation for the fluid velocities v, min(c_max, max(c_min, sqrt(max_a v_a.v_a) / Ma)).
 * v_a.v_a is summed in component order, without fused multiply-adds. */
double oracle_adaptive_sound_speed(const oracle_params *P, int64_t nf, const double *v)
{
    int nd = P->ndims;
    double vmax2 = 0.0;
    for (int64_t a = 0; a < nf; a++) {
        const double *va = v + a * (nd + 1);
        double d = va[0] * va[0];
        for (int k = 1; k < nd; k++) d += va[k] * va[k];
        if (d > vmax2) vmax2 = d;
    }
    double c = sqrt(vmax2) / P->mach_number_target;
    if (c < P->min_sound_speed) c = P->min_sound_speed;
    if (c > P->max_sound_speed) c = P->max_sound_speed;
    return c;
}

/* summation_density! (density_calculators.jl:26-48) of the fluid, with compute_density!
 * (fluid.jl:133-139) under SummationDensity: set_zero!(density); then for every neighbour
 * system the fluid interacts with, in the semidiscretization's order (foreach_system: the
 * fluid itself, then the wall), every neighbour b within the compact support -- the particle
 * itself included, whose distance 0 gives W(0) -- adds hydrodynamic_mass(b) * W(distance)
 * (smoothing_kernel with the compact-support guard, smoothing_kernels.jl:30-34).  The wall
 * neighbours' mass is boundary_model.hydrodynamic_mass (wall_boundary/system.jl:173-175).
 * x: nd x nf, xw: nd x nw (nw may be 0), rho: nf out. */
int oracle_summation_density(const oracle_params *P, int64_t nf, const double *x,
                             const double *m, int64_t nw, const double *xw, const double *mw,
                             double *rho)
{
    int nd = P->ndims;
    double h = P->smoothing_length;
    double R = 2.0 * h;
    double lo[3], hi[3];
    for (int64_t a = 0; a < nf; a++) rho[a] = 0.0;
    bbox(nd, nf, x, nw, xw, lo, hi);
    grid_t g;
    if (grid_build(&g, nd, R, lo, hi, nf, x)) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < nf; a++) {
        const double *xa = x + a * nd;
        double density = rho[a];
        FOREACH_NEIGHBOR(&g, xa, x, { density += m[b] * oracle_kernel(P->kernel, nd, distance, h); });
        rho[a] = density;
    }
    grid_free(&g);
    if (nw > 0) {
        if (grid_build(&g, nd, R, lo, hi, nw, xw)) return -1;
#pragma omp parallel for schedule(static)
        for (int64_t a = 0; a < nf; a++) {
            const double *xa = x + a * nd;
            double density = rho[a];
            FOREACH_NEIGHBOR(&g, xa, xw, { density += mw[b] * oracle_kernel(P->kernel, nd, distance, h); });
            rho[a] = density;
        }
        grid_free(&g);
    }
    return 0;
}

/* Rows of v per fluid particle: velocity and density (ContinuityDensity), or the velocity
 * alone (SummationDensity: v_nvariables, wcsph/system.jl:232-242). */
static int v_rows(const oracle_params *P) { return P->ndims + (P->density_calculator == 1 ? 0 : 1); }

/* kick! for ContinuityDensity: v and dv hold nd+1 rows per particle. */
static int kick_continuity(const oracle_params *P0, int64_t nf, const double *v,
                           const double *x, const double *m, int64_t nw, const double *xw,
                           const double *mw, double *dv, double *pf_out, double *pw_out,
                           double *rhow_out)
{
    /* update_quantities! runs first in the kick (semidiscretization.jl:616-651): the
     * adaptive sound speed of this kick, used by the state equation (fluid and walls), the
     * viscosity and the density diffusion alike */
    oracle_params Q = *P0;
    if (P0->adaptive_sound_speed) Q.sound_speed = oracle_adaptive_sound_speed(P0, nf, v);
    const oracle_params *P = &Q;
    int nd = P->ndims;
    int rc = 0;
    double *pf = (double *)malloc(sizeof(double) * (size_t)(nf > 0 ? nf : 1));
    double *pw = (double *)malloc(sizeof(double) * (size_t)(nw > 0 ? nw : 1));
    double *rhow = (double *)malloc(sizeof(double) * (size_t)(nw > 0 ? nw : 1));
    double *wvel = NULL, *ddgrad = NULL;
    if (!pf || !pw || !rhow) { rc = -1; goto out; }
    memset(dv, 0, sizeof(double) * (size_t)(nf * (nd + 1)));
    oracle_compute_pressure(P, nf, v, pf);
    if (nw > 0 && (rc = oracle_adami(P, nw, xw, nf, x, v, pf, pw, rhow, NULL))) goto out;
    if (nw > 0 && P->wall_viscosity) {
        wvel = (double *)malloc(sizeof(double) * (size_t)(nw * nd));
        if (!wvel) { rc = -1; goto out; }
        if ((rc = oracle_wall_velocity(P, nw, xw, nf, x, v, wvel))) goto out;
    }
    if (P->use_density_diffusion == 2 && P->density_calculator != 1) {
        /* update_quantities! -> update!(::DensityDiffusionAntuono) */
        ddgrad = (double *)malloc(sizeof(double) * (size_t)(nf * nd + 1));
        if (!ddgrad) { rc = -1; goto out; }
        if ((rc = oracle_antuono_gradient(P, nf, x, v, m, ddgrad, NULL))) goto out;
    }
    if ((rc = oracle_interact_fluid_fluid(P, nf, x, v, m, pf, dv, ddgrad))) goto out;
    if (nw > 0 && (rc = oracle_interact_fluid_wall(P, nf, x, v, pf, nw, xw, mw, pw, rhow, dv,
                                                   m, wvel)))
        goto out;
    /* add_acceleration!, semidiscretization.jl:720-728 */
    for (int64_t a = 0; a < nf; a++)
        for (int d = 0; d < nd; d++) dv[a * (nd + 1) + d] += P->gravity[d];
    if (pf_out) memcpy(pf_out, pf, sizeof(double) * (size_t)nf);
    if (pw_out && nw > 0) memcpy(pw_out, pw, sizeof(double) * (size_t)nw);
    if (rhow_out && nw > 0) memcpy(rhow_out, rhow, sizeof(double) * (size_t)nw);
out:
    free(pf);
    free(pw);
    free(rhow);
    free(wvel);
    free(ddgrad);
    return rc;
}

/* kick! (semidiscretization.jl:589-612).  ContinuityDensity: v, dv are (nd+1) x nf.
 * SummationDensity: v, dv are nd x nf; compute_density! runs first in update_quantities!
 * (wcsph/system.jl:291-301), and the pass order is otherwise the same, with the summed
 * density in the state equation, the Adami pass and the pair terms (their density rows carry
 * it here; there is no density rate).  rho_out (may be NULL): the density the kick used
 * (system.cache.density under SummationDensity, the density row of v otherwise). */
int oracle_kick_rho(const oracle_params *P, int64_t nf, const double *v, const double *x,
                    const double *m, int64_t nw, const double *xw, const double *mw, double *dv,
                    double *pf_out, double *pw_out, double *rhow_out, double *rho_out)
{
    int nd = P->ndims;
    if (P->density_calculator != 1) {
        int rc = kick_continuity(P, nf, v, x, m, nw, xw, mw, dv, pf_out, pw_out, rhow_out);
        if (rc == 0 && rho_out)
            for (int64_t a = 0; a < nf; a++) rho_out[a] = v[a * (nd + 1) + nd];
        return rc;
    }
    size_t ne = (size_t)(nf > 0 ? nf : 1) * (size_t)(nd + 1);
    double *vx = (double *)malloc(sizeof(double) * ne);
    double *dvx = (double *)malloc(sizeof(double) * ne);
    double *rho = (double *)malloc(sizeof(double) * (size_t)(nf > 0 ? nf : 1));
    int rc = -1;
    if (vx && dvx && rho && oracle_summation_density(P, nf, x, m, nw, xw, mw, rho) == 0) {
        for (int64_t a = 0; a < nf; a++) {
            for (int d = 0; d < nd; d++) vx[a * (nd + 1) + d] = v[a * nd + d];
            vx[a * (nd + 1) + nd] = rho[a];
        }
        rc = kick_continuity(P, nf, vx, x, m, nw, xw, mw, dvx, pf_out, pw_out, rhow_out);
        if (rc == 0) {
            for (int64_t a = 0; a < nf; a++)
                for (int d = 0; d < nd; d++) dv[a * nd + d] = dvx[a * (nd + 1) + d];
            if (rho_out) memcpy(rho_out, rho, sizeof(double) * (size_t)nf);
        }
    }
    free(vx);
    free(dvx);
    free(rho);
    return rc;
}

int oracle_kick(const oracle_params *P, int64_t nf, const double *v, const double *x,
                const double *m, int64_t nw, const double *xw, const double *mw, double *dv,
                double *pf_out, double *pw_out, double *rhow_out)
{
    return oracle_kick_rho(P, nf, v, x, m, nw, xw, mw, dv, pf_out, pw_out, rhow_out, NULL);
}

/* drift! -> set_velocity!(::AbstractFluidSystem), semidiscretization.jl:561-571: du = the
 * velocity rows of v (vrows rows per particle: nd + 1 with ContinuityDensity, nd with
 * SummationDensity). */
void oracle_drift_rows(int nd, int vrows, int64_t nf, const double *v, double *du)
{
    for (int64_t a = 0; a < nf; a++)
        for (int d = 0; d < nd; d++) du[a * nd + d] = v[a * vrows + d] + 0.0;
}

void oracle_drift(int nd, int64_t nf, const double *v, double *du)
{
    oracle_drift_rows(nd, nd + 1, nf, v, du);
}

/* ---------------------------------------------------------------------------------------
 * CarpenterKennedy2N54 (OrdinaryDiffEqLowStorageRK, williamson_condition=false), with the
 * 2N-storage coefficients of Carpenter & Kennedy (1994), NASA TM-109112, solution 3.
 * Stage: tmp = A_i*tmp + dt*k ; y = y + B_i*tmp, applied to the partition (v, u) with
 * k = (kick!, drift!) evaluated at the current stage state.  The integrator's stage update is
 * written under `@muladd`, so `y + B*tmp` and `tmp + dt*k` are fused multiply-adds there; the
 * restatement uses fma() for exactly those two updates.
 * ------------------------------------------------------------------------------------- */
static const double CK_A[5] = {0.0,
                               -567301805773.0 / 1357537059087.0,
                               -2404267990393.0 / 2016746695238.0,
                               -3550918686646.0 / 2091501179385.0,
                               -1275806237668.0 / 842570457699.0};
static const double CK_B[5] = {1432997174477.0 / 9575080441755.0,
                               5161836677717.0 / 13612068292357.0,
                               1720146321549.0 / 2090206949498.0,
                               3134564353537.0 / 4481467310338.0,
                               2277821191437.0 / 14882151754819.0};
static const double CK_C[5] = {0.0,
                               1432997174477.0 / 9575080441755.0,
                               2526269341429.0 / 6820363962896.0,
                               2006345519317.0 / 3224310063776.0,
                               2802321613138.0 / 2924317926251.0};

void oracle_ck2n54_coefficients(double *A, double *B, double *C)
{
    for (int i = 0; i < 5; i++) { A[i] = CK_A[i]; B[i] = CK_B[i]; C[i] = CK_C[i]; }
}

/* Advance (v, u) by nsteps fixed steps of size dt.  v: (nd+1) x nf (nd x nf with
 * SummationDensity), u: nd x nf. */
int oracle_solve_ck2n54(const oracle_params *P, int64_t nf, double *v, double *u,
                        const double *m, int64_t nw, const double *xw, const double *mw,
                        double dt, int64_t nsteps)
{
    int nd = P->ndims;
    int64_t nv = nf * v_rows(P), nu = nf * nd;
    double *kv = (double *)malloc(sizeof(double) * (size_t)(nv > 0 ? nv : 1));
    double *ku = (double *)malloc(sizeof(double) * (size_t)(nu > 0 ? nu : 1));
    double *tv = (double *)calloc((size_t)(nv > 0 ? nv : 1), sizeof(double));
    double *tu = (double *)calloc((size_t)(nu > 0 ? nu : 1), sizeof(double));
    int rc = 0;
    if (!kv || !ku || !tv || !tu) { rc = -1; goto out; }
    for (int64_t s = 0; s < nsteps; s++) {
        for (int i = 0; i < 5; i++) {
            if ((rc = oracle_kick(P, nf, v, u, m, nw, xw, mw, kv, NULL, NULL, NULL))) goto out;
            oracle_drift_rows(nd, v_rows(P), nf, v, ku);
            if (i == 0) {
#pragma omp parallel for schedule(static)
                for (int64_t k = 0; k < nv; k++) { tv[k] = dt * kv[k]; v[k] = fma(CK_B[0], tv[k], v[k]); }
#pragma omp parallel for schedule(static)
                for (int64_t k = 0; k < nu; k++) { tu[k] = dt * ku[k]; u[k] = fma(CK_B[0], tu[k], u[k]); }
            } else {
#pragma omp parallel for schedule(static)
                for (int64_t k = 0; k < nv; k++) {
                    tv[k] = fma(dt, kv[k], CK_A[i] * tv[k]);
                    v[k] = fma(CK_B[i], tv[k], v[k]);
                }
#pragma omp parallel for schedule(static)
                for (int64_t k = 0; k < nu; k++) {
                    tu[k] = fma(dt, ku[k], CK_A[i] * tu[k]);
                    u[k] = fma(CK_B[i], tu[k], u[k]);
                }
            }
        }
    }
out:
    free(kv);
    free(ku);
    free(tv);
    free(tu);
    return rc;
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
