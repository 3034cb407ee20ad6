/*
 * tpsph_batch.h -- C ABI of the ensemble runner: many independent WCSPH dam breaks of
 * uneven size advanced together, one set of kernels per stage over all of them, each
 * CarpenterKennedy2N54 step captured in a HIP graph.
 *
 * What it replaces: a loop of the reference over an ensemble of set-ups, each
 *   semi = Semidiscretization(fluid_system, boundary_system)  semidiscretization.jl:106-155
 *   ode  = semidiscretize(semi, tspan)                        semidiscretization.jl:293-396
 *   solve(ode, CarpenterKennedy2N54(williamson_condition=false), dt=dt, adaptive=false)
 * as examples/fluid/dam_break_2d.jl:18-127 builds one; in SciML terms an EnsembleProblem of
 * those ODE problems solved with EnsembleSerial.  The cases share everything but their
 * geometry (particles, walls, grid box), their sound speed, their Monaghan alpha and their
 * diffusion delta: one tpsph_params gives the shared fields (as for tpsph_create; its
 * sound_speed, alpha, delta, domain_min/max, skin and pair_search are ignored), one
 * tpsph_case per case the rest.
 *
 * State: the cases' v_ode / u_ode (the reference layout of tpsph.h) concatenated in case
 * order, fluid_mass likewise; wall_coordinates (ndims x n_wall per case) and wall_mass
 * concatenated the same way.  Device pointers only.
 *
 * Results: every case's result equals, bit for bit, that of the case alone -- in an ensemble
 * of one, and on a tpsph_handle with the same params, its own alpha, delta, sound speed and
 * box, pair_search TPSPH_SEARCH_CELLS and skin 0 -- whatever the other cases are: the cell
 * order is rebuilt per case at every kick and every sum follows the case's own cell order.
 *
 * Errors as in tpsph.h; tpsph_batch_last_error(NULL) after a failed create.
 */
#ifndef TPSPH_BATCH_H
#define TPSPH_BATCH_H

#include <stdint.h>

#include "tpsph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tpsph_case {
    int64_t n_fluid;       /* fluid particles of the case                              */
    int64_t n_wall;        /* wall particles of the case                               */
    double alpha;          /* ArtificialViscosityMonaghan(; alpha)                     */
    double delta;          /* DensityDiffusionMolteniColagrossi(; delta)               */
    double sound_speed;    /* StateEquationCole(; sound_speed)                         */
    double domain_min[3];  /* grid box of the case (the FullGridCellList box)          */
    double domain_max[3];
} tpsph_case;

typedef struct tpsph_batch tpsph_batch;

/* Launch shapes the autotuner sets (tpsph_batch_set_launch). */
enum tpsph_batch_launch {
    TPSPH_BATCH_FLUID_BLOCK = 0,  /* threads per block of the fluid pass: 64, 128, 256   */
    TPSPH_BATCH_WALL_BLOCK = 1    /* threads per block of the Adami pass: 64, 128, 256   */
};

/* The ensemble of n_cases cases; static data copied to the device, the static wall cell
 * lists of all cases built once.  stream == NULL: the ensemble creates a stream of its own
 * (graph capture needs one other than the null stream). */
int tpsph_batch_create(const tpsph_params *params, int32_t n_cases, const tpsph_case *cases,
                       const double *fluid_mass, const double *wall_coordinates,
                       const double *wall_mass, void *stream, tpsph_batch **out);

/* kick! (semidiscretization.jl:589-612) of every case at once. */
int tpsph_batch_kick(tpsph_batch *b, double *dv_ode, const double *v_ode, const double *u_ode,
                     double t);

/* nsteps fixed CarpenterKennedy2N54 steps of every case; use_graph != 0 replays one captured
 * step (recaptured when v_ode, u_ode or dt change).  Waits for the device before starting
 * and returns after the last step. */
int tpsph_batch_solve_ck2n54(tpsph_batch *b, double *v_ode, double *u_ode, double t0, double dt,
                             int64_t nsteps, int32_t use_graph);

/* Launch shape `which` (enum tpsph_batch_launch) = value; results do not depend on it. */
int tpsph_batch_set_launch(tpsph_batch *b, int32_t which, int32_t value);

int tpsph_batch_synchronize(tpsph_batch *b);
const char *tpsph_batch_last_error(const tpsph_batch *b);
void tpsph_batch_destroy(tpsph_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* TPSPH_BATCH_H */
