/*
 * tpsph.h -- C ABI of the MI355X-native WCSPH right-hand side (HIP, gfx950).
 *
 * The library replaces the per-stage work behind TrixiParticles.jl's ODE surface for one
 * fluid system (WeaklyCompressibleSPHSystem, ContinuityDensity, StateEquationCole or
 * StateEquationAdaptiveCole, ArtificialViscosityMonaghan, DensityDiffusionMolteniColagrossi,
 * WendlandC2Kernel) and one
 * static wall system (WallBoundarySystem + BoundaryModelDummyParticles +
 * AdamiPressureExtrapolation):
 *
 *   kick!(dv_ode, v_ode, u_ode, p, t)   src/general/semidiscretization.jl:589-612
 *   drift!(du_ode, v_ode, u_ode, p, t)  src/general/semidiscretization.jl:522-536
 *
 * State layout is the reference's (column-major, one column per particle, FP64):
 *   v_ode : (ndims+1) x n_fluid  = [v_1 .. v_ndims, rho] per particle
 *           (wcsph/system.jl:240-242, 258-279; the wall block of v_ode is empty);
 *           ndims x n_fluid = [v_1 .. v_ndims] with SummationDensity (wcsph/system.jl:236)
 *   u_ode :  ndims    x n_fluid  = [x_1 .. x_ndims]       per particle
 *   dv_ode, du_ode: same shapes as v_ode, u_ode.
 * Results are always in the caller's particle order; the library's cell-sorted internal
 * order is invisible at this boundary.
 *
 * Pointers are device pointers when params.memory == TPSPH_MEM_DEVICE and host pointers when
 * params.memory == TPSPH_MEM_HOST (the library then stages through its own device buffers).
 * Every call is asynchronous on the handle's stream in device mode; host mode returns after
 * the results are in host memory.
 *
 * Errors: every entry point returns an int status (TPSPH_OK == 0).  The message of the last
 * failure is available from tpsph_last_error(handle) (or tpsph_last_error(NULL) for a
 * failed tpsph_create).  Kernels never abort; argument validation mirrors the reference's
 * ArgumentError checks at construction time.
 *
 * A handle is bound to one stream and is not re-entrant; distinct handles may run
 * concurrently on distinct streams.
 */
#ifndef TPSPH_H
#define TPSPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TPSPH_ABI_VERSION 6

enum tpsph_status {
    TPSPH_OK = 0,
    TPSPH_ERR_ARGUMENT = 1, /* invalid argument (ArgumentError in the reference) */
    TPSPH_ERR_HIP = 2,      /* HIP runtime failure */
    TPSPH_ERR_CAPACITY = 3, /* output buffer too small (see tpsph_neighbors_csr) */
    TPSPH_ERR_STATE = 4     /* call not valid in the handle's state */
};

enum tpsph_memory { TPSPH_MEM_DEVICE = 0, TPSPH_MEM_HOST = 1 };

enum tpsph_kernel_kind { TPSPH_KERNEL_WENDLAND_C2 = 0 };

/* Which neighbour relation tpsph_neighbors_csr reports. */
enum tpsph_pair {
    TPSPH_PAIR_FLUID_FLUID = 0, /* rows: fluid particles, columns: fluid particles */
    TPSPH_PAIR_FLUID_WALL = 1,  /* rows: fluid particles, columns: wall particles  */
    TPSPH_PAIR_WALL_FLUID = 2   /* rows: wall particles,  columns: fluid particles */
};

/* Pair search of the fluid kernel (tpsph_params.pair_search). */
enum tpsph_pair_search {
    TPSPH_SEARCH_AUTO = 0,   /* Verlet lists when skin > 0, else the cell walk           */
    TPSPH_SEARCH_LISTS = 1,  /* Verlet lists rebuilt with the cell list                  */
    TPSPH_SEARCH_CELLS = 2   /* cell walk on every kick                                  */
};

/* Density diffusion of the fluid (tpsph_params.use_density_diffusion). */
enum tpsph_density_diffusion {
    TPSPH_DIFFUSION_NONE = 0,      /* nothing                                              */
    TPSPH_DIFFUSION_MOLTENI = 1,   /* DensityDiffusionMolteniColagrossi(; delta)           */
    TPSPH_DIFFUSION_ANTUONO = 2    /* DensityDiffusionAntuono(; delta)                     */
};

/* Arithmetic of the pair terms (tpsph_params.precision). */
enum tpsph_precision {
    TPSPH_PRECISION_FP64 = 0,  /* FP64 throughout (the reference's Float64)                */
    TPSPH_PRECISION_MIXED = 1  /* FP32 pair arithmetic on brick-relative coordinates with  */
                               /* compensated density-rate sums; FP64 positions, state,    */
                               /* state equation, stage update and reductions.  The fluid  */
                               /* viscosity must have beta = 0 (TPSPH_ERR_ARGUMENT else)   */
};

/* Density calculator of the fluid (tpsph_params.density_calculator). */
enum tpsph_density_calculator {
    TPSPH_DENSITY_CONTINUITY = 0, /* ContinuityDensity(): rho is integrated, v_ode row ndims+1 */
    TPSPH_DENSITY_SUMMATION = 1   /* SummationDensity(): rho = sum_b m_b W_ab over the fluid */
                                  /* and wall neighbours each kick; v_ode holds the velocity */
                                  /* only (ndims rows) and dv_ode has no density rate        */
};

/* Kernel identifiers for tpsph_kernel_time. */
enum tpsph_timed_kernel {
    TPSPH_TIME_CELL_LIST = 0,   /* fluid cell list (counting sort) and Verlet lists        */
    TPSPH_TIME_GATHER = 1,      /* reorder into cell-sorted SoA + fluid state equation   */
    TPSPH_TIME_WALL = 2,        /* Adami pressure extrapolation + inverse state equation  */
    TPSPH_TIME_FLUID = 3,       /* fused fluid-fluid + fluid-wall interaction + gravity   */
    TPSPH_TIME_STAGE = 4,       /* CK2N54 stage update                                    */
    TPSPH_TIME_ADJOINT = 5,     /* reverse-mode passes (tpsph_kick_vjp, tpsph_adjoint_*)  */
    TPSPH_TIME_COUNT = 6
};

/*
 * Physical and numerical parameters.  Field meanings follow the reference constructors:
 *   smoothing_length, kernel           WeaklyCompressibleSPHSystem(; smoothing_kernel,
 *                                      smoothing_length) (wcsph/system.jl:90-174); the wall
 *                                      model uses the same kernel and length
 *   reference_density, sound_speed,    StateEquationCole(...) (state_equations.jl:100-139)
 *   exponent, background_pressure,
 *   clip_fluid_eos
 *   use_viscosity, alpha, beta,        ArtificialViscosityMonaghan (viscosity.jl:68-83)
 *   epsilon
 *   use_density_diffusion, delta       enum tpsph_density_diffusion with its delta:
 *                                      DensityDiffusionMolteniColagrossi
 *                                      (density_diffusion.jl:41-55) or DensityDiffusionAntuono
 *                                      (density_diffusion.jl:89-211), whose renormalised
 *                                      density gradients every kick computes in a pass of its
 *                                      own; Antuono runs with the Verlet lists (the brick
 *                                      kernel stages the gradients) or the cell walk
 *   acceleration                       `acceleration` of the fluid system (gravity)
 *   clip_wall_pressure                 BoundaryModelDummyParticles(; clip_negative_pressure)
 *   pressure_offset                    AdamiPressureExtrapolation(; pressure_offset)
 *   domain_min, domain_max             bounding box of the cell grid (the reference's
 *                                      FullGridCellList box); particles outside it are
 *                                      clamped into the border cells, which keeps the
 *                                      neighbour search exact but slower there
 *   cell_divisions                     cells per compact support along y and z (0 = 2)
 *   x_cell_divisions                   cells per compact support along x (0 = 2 cell_divisions);
 *                                      x is the fastest-varying cell index
 *   skin                               Verlet skin of the fluid cell list as a fraction of
 *                                      the compact support (< 0 = default 0.04): the
 *                                      cell-sorted order is rebuilt only once a particle has
 *                                      moved by more than skin/2 since the last rebuild.
 *                                      Every kick still tests all pairs exactly against the
 *                                      current positions, so the neighbour relation is the
 *                                      reference's one whatever the skin.
 *   pair_search                        how the fluid kernel finds candidate pairs
 *                                      (enum tpsph_pair_search); all give the same pairs
 *   adaptive_sound_speed,              StateEquationAdaptiveCole(; mach_number_target,
 *   mach_number_target,                min_sound_speed, max_sound_speed)
 *   min_sound_speed, max_sound_speed   (state_equations.jl:1-76): with adaptive_sound_speed
 *                                      = 1 every kick first sets the sound speed of the
 *                                      state equation, the viscosity and the density
 *                                      diffusion to clamp(max |v| / mach_number_target,
 *                                      min_sound_speed, max_sound_speed) over the fluid's
 *                                      current velocities (update_speed_of_sound!,
 *                                      wcsph/system.jl:307-320); `sound_speed` is then the
 *                                      value before the first kick (the reference starts
 *                                      at min_sound_speed)
 *   wall_viscosity, wall_alpha,        BoundaryModelDummyParticles(; viscosity): the wall
 *   wall_beta, wall_epsilon, wall_nu   model's ArtificialViscosityMonaghan(alpha, beta,
 *                                      epsilon) or ViscosityAdami(nu, epsilon) acting on the
 *                                      fluid-wall pairs with the no-slip wall velocity the
 *                                      Adami pass extrapolates (dummy_particles.jl:716-756);
 *                                      needs the fluid's ArtificialViscosityMonaghan
 *   density_calculator                 WeaklyCompressibleSPHSystem(; density_calculator):
 *                                      ContinuityDensity (density_calculators.jl:26) or
 *                                      SummationDensity (density_calculators.jl:11, 26-48;
 *                                      compute_density!, fluid.jl:133-139).  SummationDensity
 *                                      changes the state layout (v_ode and dv_ode hold
 *                                      ndims rows per particle, wcsph/system.jl:236-238,
 *                                      252-256), adds the summation pass before the state
 *                                      equation, takes the pressure acceleration
 *                                      -m_b (p_a/rho_a^2 + p_b/rho_b^2) grad W
 *                                      (pressure_acceleration.jl:8-15, 139-145) and has no
 *                                      continuity or density-diffusion term (fluid.jl:
 *                                      133-139); it runs in FP64, with the Verlet lists
 *                                      (summation pass over the lists, fluid pass on the
 *                                      brick kernel; AUTO selects them when skin > 0) or the
 *                                      cell walk
 */
typedef struct tpsph_params {
    int32_t ndims;                 /* 2 or 3 */
    int32_t kernel;                /* TPSPH_KERNEL_WENDLAND_C2 */
    int32_t memory;                /* TPSPH_MEM_DEVICE or TPSPH_MEM_HOST */
    int32_t clip_fluid_eos;        /* 0/1 */
    int32_t clip_wall_pressure;    /* 0/1 */
    int32_t use_viscosity;         /* 0/1 */
    int32_t use_density_diffusion; /* enum tpsph_density_diffusion */
    int32_t cell_divisions;        /* 0 = default (2) */
    int32_t x_cell_divisions;      /* 0 = default (2 * cell_divisions) */
    int32_t pair_search;           /* enum tpsph_pair_search, 0 = auto */
    double smoothing_length;
    double reference_density;
    double sound_speed;
    double exponent;
    double background_pressure;
    double alpha;
    double beta;
    double epsilon;
    double delta;
    double acceleration[3];
    double pressure_offset;
    double domain_min[3];
    double domain_max[3];
    double skin;                   /* < 0 = default (0.04) */
    int32_t adaptive_sound_speed;  /* 0/1: StateEquationAdaptiveCole */
    int32_t precision;             /* enum tpsph_precision, 0 = FP64 */
    double mach_number_target;
    double min_sound_speed;
    double max_sound_speed;
    int32_t wall_viscosity;        /* 0 nothing, 1 ArtificialViscosityMonaghan, 2 ViscosityAdami */
    int32_t density_calculator;    /* enum tpsph_density_calculator, 0 = ContinuityDensity */
    double wall_alpha;             /* ArtificialViscosityMonaghan(; alpha, beta) of the walls */
    double wall_beta;
    double wall_epsilon;           /* epsilon of the wall model */
    double wall_nu;                /* ViscosityAdami(; nu) of the walls */
} tpsph_params;

typedef struct tpsph_handle tpsph_handle;

/* Library ABI version (TPSPH_ABI_VERSION of the build). */
int tpsph_abi_version(void);

/* Create a handle for one fluid + one static wall system.
 *   fluid_mass:       n_fluid masses            (hydrodynamic_mass, fluid.jl:44-46)
 *   wall_coordinates: ndims x n_wall            (WallBoundarySystem.coordinates)
 *   wall_mass:        n_wall masses             (boundary_model.hydrodynamic_mass)
 *   stream:           hipStream_t to run on (NULL = the null stream)
 * Replaces semidiscretize(semi, tspan) (semidiscretization.jl:293-396) for this pair: the
 * static data is copied to the device once and the static wall cell list is built here. */
int tpsph_create(const tpsph_params *params, int64_t n_fluid, int64_t n_wall,
                 const double *fluid_mass, const double *wall_coordinates,
                 const double *wall_mass, void *stream, tpsph_handle **out);

/* kick!(dv_ode, v_ode, u_ode, p, t) (semidiscretization.jl:589-612): zero dv, rebuild the
 * fluid cell list (update_nhs!, neighborhood_search.jl:479-519), fluid pressure
 * (wcsph/system.jl:426-438), Adami wall pressure/density (dummy_particles.jl:489-569),
 * fluid-fluid and fluid-wall interact! (wcsph/rhs.jl:5-127), gravity
 * (semidiscretization.jl:668-731). */
int tpsph_kick(tpsph_handle *h, double *dv_ode, const double *v_ode, const double *u_ode,
               double t);

/* drift!(du_ode, v_ode, u_ode, p, t) (semidiscretization.jl:522-571): du = v. */
int tpsph_drift(tpsph_handle *h, double *du_ode, const double *v_ode, const double *u_ode,
                double t);

/* Quantities of the last kick, in the caller's particle order (any pointer may be NULL):
 * system.pressure (n_fluid), boundary_model.pressure (n_wall),
 * boundary_model.cache.density (n_wall).  tpsph_solve_ck2n54 ends with a state no kick has
 * seen yet, so after it this returns TPSPH_ERR_STATE until the next tpsph_kick (the
 * callbacks of the reference evaluate kick! on the saved state first, callbacks.jl:28-40). */
int tpsph_get_pressures(tpsph_handle *h, double *fluid_pressure, double *wall_pressure,
                        double *wall_density);

/* current_density(v, system) of the last kick (n_fluid values, caller order): the summed
 * density (system.cache.density, wcsph/system.jl:269-273) under SummationDensity, the density
 * row of v_ode under ContinuityDensity.  TPSPH_ERR_STATE as tpsph_get_pressures. */
int tpsph_get_density(tpsph_handle *h, double *fluid_density);

/* system_sound_speed(system) (wcsph/system.jl:285) as of the last kick: the sound speed the
 * state equation, viscosity and density diffusion used (with adaptive_sound_speed the value
 * update_speed_of_sound! set, else params.sound_speed).  The StepsizeCallback reads it
 * (fluid.jl:199-231).  Synchronises the handle's stream. */
int tpsph_get_sound_speed(tpsph_handle *h, double *sound_speed);

/* Neighbour sets of the current positions as CSR (PointNeighbors' foreach_point_neighbor
 * relation, neighborhood_search.jl:5-13: all b with |x_a - x_b|^2 <= (2h)^2, self
 * included), rows and column indices in the caller's order, columns ascending per row.
 * offsets has n_rows+1 entries.  With idx == NULL (or capacity too small) only offsets are
 * written and the call returns TPSPH_OK (resp. TPSPH_ERR_CAPACITY); offsets[n_rows] is the
 * total.  u_ode is the fluid position array as in tpsph_kick. */
int tpsph_neighbors_csr(tpsph_handle *h, const double *u_ode, int32_t pair, int64_t *offsets,
                        int32_t *idx, int64_t capacity);

/* The pairs the kernels of the last kick evaluated, as CSR in the caller's order (rows and
 * columns as in tpsph_neighbors_csr, columns ascending per row): the candidates of the lists
 * the kick read -- the brick lists of LDS indices, the Verlet lists, or the cell walk of an
 * overflowed list or of the cell search -- that passed the kernels' exact radius test on
 * the kick's positions.  That is PointNeighbors' foreach_point_neighbor relation
 * (neighborhood_search.jl:5-13) without the pairs the pair terms skip at distance^2 below
 * eps(R^2) (wcsph/rhs.jl:27-28, 52) for TPSPH_PAIR_FLUID_FLUID and TPSPH_PAIR_FLUID_WALL,
 * and with them for the Adami sums of TPSPH_PAIR_WALL_FLUID (dummy_particles.jl:637-672).
 * Exposes what the kick summed, for the parity tests of the pair search; TPSPH_ERR_STATE
 * before the first kick.  Capacity semantics as tpsph_neighbors_csr. */
int tpsph_kick_pairs_csr(tpsph_handle *h, int32_t pair, int64_t *offsets, int32_t *idx,
                         int64_t capacity);

/* One CarpenterKennedy2N54 stage over y (n values) with register tmp and slope k:
 *   first stage: tmp = dt*k;              y = y + b*tmp
 *   otherwise:   tmp = a*tmp + dt*k;      y = y + b*tmp      (2N-storage form)
 * The building block a device-array type needs to run OrdinaryDiffEq's low-storage
 * update without leaving the device. */
int tpsph_rk_stage(tpsph_handle *h, double *y, double *tmp, const double *k, double a,
                   double b, double dt, int64_t n, int32_t first_stage);

/* y = a*y + b*x over n values: the single broadcasts OrdinaryDiffEq's low-storage step
 * issues one at a time (LowStorageRK2N perform_step! with williamson_condition = false:
 * tmp = dt*k, u = u + B*tmp, tmp = A*tmp, tmp = tmp + dt*k), for a device-array type that
 * runs them on the device.  b == 0 reads only y (y = a*y), a == 0 reads only x (y = b*x),
 * a == b == 0 writes zeros (fill!); otherwise y = fma(b, x, a*y), muladd's rounding, so the
 * sequence reproduces tpsph_rk_stage and tpsph_solve_ck2n54 bit for bit. */
int tpsph_axpby(tpsph_handle *h, int64_t n, double *y, double a, const double *x, double b);

/* nsteps fixed steps of CarpenterKennedy2N54(williamson_condition=false) from t0, fully on
 * the device: per stage kick! + drift! + the fused stage update of (v_ode, u_ode). */
int tpsph_solve_ck2n54(tpsph_handle *h, double *v_ode, double *u_ode, double t0, double dt,
                       int64_t nsteps);

/* kick! without any device allocation, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops).  No counterpart in the reference: it is tpsph_create +
 * tpsph_kick + tpsph_destroy in one call, with every intermediate array (cell lists,
 * cell-sorted records, wall pressures) in a fixed 64 MiB device array of the library, one
 * per GPU, held until the process ends.  Results equal those of tpsph_kick on a handle
 * created with the same params and pair_search TPSPH_SEARCH_CELLS.
 *   - device pointers only (params.memory == TPSPH_MEM_DEVICE), asynchronous on `stream`;
 *   - params.domain_min < domain_max on every axis (the grid box; it is not derived from the
 *     wall coordinates, which would need a copy to the host);
 *   - StateEquationCole, Monaghan viscosity, Molteni-Colagrossi or no diffusion, walls
 *     without viscosity, FP64;
 *   - sizes whose arrays exceed the 64 MiB (about 120 bytes per fluid and 100 per wall
 *     particle plus 17 per cell) are rejected with TPSPH_ERR_ARGUMENT;
 *   - calls on one GPU share its array: calls that overlap on one GPU (different streams or
 *     threads) race and must be serialised by the caller.
 * Errors go to tpsph_last_error(NULL). */
int tpsph_kick_noalloc(const tpsph_params *params, int64_t n_fluid, int64_t n_wall,
                       const double *fluid_mass, const double *wall_coordinates,
                       const double *wall_mass, double *dv_ode, const double *v_ode,
                       const double *u_ode, double t, void *stream);

/*
 * Reverse mode.  The reference has no automatic differentiation through its right-hand side;
 * these entry points are what a ChainRulesCore.rrule of kick! and of a trajectory would call
 * (the Julia side is in INTEGRATION.md).  Differentiated inputs: the fluid state and the
 * scalars alpha (ArtificialViscosityMonaghan), delta (DensityDiffusionMolteniColagrossi) and
 * g (gravity = g * acceleration / |acceleration|); param_bar = {dL/dalpha, dL/ddelta, dL/dg}
 * (3 doubles, host or device memory; the call synchronises when it is given).  Supported:
 * device pointers, StateEquationCole with a fixed sound speed, Monaghan viscosity,
 * Molteni-Colagrossi or no density diffusion, walls without viscosity, FP64; anything else
 * returns TPSPH_ERR_ARGUMENT.  Non-smooth points take the branch the forward pass takes (the
 * viscosity switch, the wall-pressure clipping).  Deterministic (gathers only).
 */

/* The pullback of kick! at (v_ode, u_ode) (semidiscretization.jl:589-612, the terms of
 * wcsph/rhs.jl:5-127 and dummy_particles.jl:489-672): for the adjoint dv_bar of dv_ode,
 * v_bar = (d dv / d v)^T dv_bar, u_bar = (d dv / d u)^T dv_bar (shapes of v_ode, u_ode,
 * overwritten) and param_bar as above (may be NULL). */
int tpsph_kick_vjp(tpsph_handle *h, const double *dv_bar, const double *v_ode,
                   const double *u_ode, double t, double *v_bar, double *u_bar,
                   double *param_bar);

/* The forward half of a trajectory's rrule: nsteps of CarpenterKennedy2N54 from
 * (v0_ode, u0_ode) into (v_ode, u_ode) (which may alias the inputs), as tpsph_solve_ck2n54,
 * keeping the state every checkpoint_every steps in the handle for tpsph_adjoint_reverse. */
int tpsph_adjoint_forward(tpsph_handle *h, const double *v0_ode, const double *u0_ode,
                          double t0, double dt, int64_t nsteps, int64_t checkpoint_every,
                          double *v_ode, double *u_ode);

/* The pullback of the last tpsph_adjoint_forward: given the adjoints (v_bar, u_bar) of the
 * final state, writes those of the initial state (v0_bar, u0_bar) and param_bar.  Each
 * segment between checkpoints is recomputed from its checkpoint with every stage state kept
 * (5 checkpoint_every states of device memory), then swept backwards.  TPSPH_ERR_STATE
 * without a stored trajectory. */
int tpsph_adjoint_reverse(tpsph_handle *h, const double *v_bar, const double *u_bar,
                          double *v0_bar, double *u0_bar, double *param_bar);

/* Integrals of the fluid state (v_ode, u_ode) in FP64, summed in a fixed order on the
 * device (deterministic), the global reductions a run reports beside its state:
 *   out[0]    total mass sum m_a
 *   out[1..3] momentum sum m_a v_a (out[3] = 0 in 2D)
 *   out[4]    kinetic energy sum m_a |v_a|^2 / 2
 *   out[5]    potential energy -sum m_a g.x_a (g = params.acceleration)
 *   out[6]    fluid volume sum m_a / rho_a
 *   out[7]    sum rho_a
 * out holds 8 doubles (host or device memory).  Synchronises the handle's stream. */
int tpsph_integrals(tpsph_handle *h, const double *v_ode, const double *u_ode, double *out);

/* Number of rebuilds of the fluid cell-sorted order (and of the Verlet lists) since
 * tpsph_create: a kick rebuilds when a particle has moved by more than half the skin since
 * the last rebuild.  Synchronises the handle's stream. */
int tpsph_rebuild_count(tpsph_handle *h, int64_t *count);

/* SortingCallback's affect! for the fluid system (sort_particles! and sort_system!,
 * src/callbacks/sorting.jl:118-158): puts the caller's fluid particles into the cell-sorted
 * order of their current positions, in place.  Row j of the sorted v_ode / u_ode is row
 * perm[j] (0-based) of the state passed in; the handle's fluid masses are permuted with the
 * state (the reference leaves the masses unsorted, sorting.jl:131, which mixes the
 * non-uniform masses of a hydrostatic set-up up: the caller permutes its own copy of the
 * masses with perm too).  The order is the library's cell order (cells of the fluid grid,
 * x fastest, by caller index within a cell), so after the sort the library's internal
 * permutation is the identity and the gathers and scatters of every kick are contiguous.
 * Each particle's rates equal, bit for bit, those of the unsorted state on a handle that
 * rebuilds its cell list at the same positions (the sums run in the same order).  perm may
 * be NULL;
 * otherwise it holds n_fluid int64 values in the memory kind of the handle.  The pressures
 * of the sorted state are ready for tpsph_get_pressures afterwards.  A trajectory recorded
 * by tpsph_adjoint_forward is dropped (its states are in the old order). */
int tpsph_sort_particles(tpsph_handle *h, double *v_ode, double *u_ode, int64_t *perm);

/* Wait for all work queued on the handle's stream. */
int tpsph_synchronize(tpsph_handle *h);

/* Per-kernel HIP-event timing (events recorded on the handle's stream).  enable != 0
 * starts (and resets) timing; tpsph_kernel_time synchronises and reports the number of
 * launches and their summed duration in milliseconds since timing was enabled. */
int tpsph_enable_timing(tpsph_handle *h, int32_t enable);
int tpsph_kernel_time(tpsph_handle *h, int32_t which, int64_t *launches, double *total_ms);

/* Message of the last failure on h (or of the last failed tpsph_create when h == NULL). */
const char *tpsph_last_error(const tpsph_handle *h);

void tpsph_destroy(tpsph_handle *h);

#ifdef __cplusplus
}
#endif

#endif /* TPSPH_H */
