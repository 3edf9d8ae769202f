/*
 * ps.h — C-ABI of the MI355X (gfx950) unified particle solver hot path.
 *
 * One opaque handle (ps_system) owns every device buffer and one HIP stream.
 * It replaces the reference's process-global thrust::device_vectors and the
 * per-stage extern "C" functions in gpu/src/cuda/wrappers.cuh:12-97,
 * gpu/src/cuda/util.cuh:6-25 and gpu/src/cuda/shared_variables.cuh:13-36,
 * whose only caller is ParticleSystem (gpu/src/particlesystem.cpp:88-245).
 *
 * Conventions
 *  - Every function returns 0 (PS_OK) or a negative ps_error code; nothing
 *    calls exit().  ps_last_error() gives a message for the last failure.
 *  - Host arrays passed in are copied; none is retained after the call.
 *  - Particles are addressed by insertion index ("original index"), as the
 *    reference's VBO order is (gpu/src/particlesystem.cpp:338-342).  The
 *    solver keeps its own storage in cell-sorted order; readbacks return
 *    insertion order unless the name says "sorted".
 *  - Work is enqueued on the handle's stream; ps_step and the stage calls
 *    return before the GPU finishes.  Readbacks synchronise the stream.
 *  - One handle per host thread.  Two handles never share device state.
 *  - Phases follow gpu/src/cuda/shared_variables.cuh:4-9: FLUID 0, GAS 1,
 *    CLOTH 2, SOLID 3, RIGID+k (k >= 0) for bodies that must not self-collide.
 */
#ifndef PS_H
#define PS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PS_ABI_VERSION 1u

/* Particle phases (gpu/src/cuda/shared_variables.cuh:4-9). */
#define PS_PHASE_FLUID 0
#define PS_PHASE_GAS 1
#define PS_PHASE_CLOTH 2
#define PS_PHASE_SOLID 3
#define PS_PHASE_RIGID 4

/* Error codes. */
typedef enum ps_error {
    PS_OK = 0,
    PS_ERR_INVALID_ARGUMENT = -1, /* bad pointer, size or parameter value      */
    PS_ERR_CAPACITY = -2,         /* would exceed max_particles or a table size */
    PS_ERR_DEVICE = -3,           /* a HIP runtime call failed                  */
    PS_ERR_NO_DEVICE = -4,        /* no usable GPU                              */
    PS_ERR_STATE = -5,            /* call out of order (e.g. iterate w/o begin) */
    PS_ERR_OUT_OF_MEMORY = -6,    /* device allocation failed                   */
    PS_ERR_NOT_FINITE = -7,       /* non-finite input values                    */
    PS_ERR_UNSUPPORTED = -8,      /* call not available in this solver mode     */
    PS_ERR_IO = -9                /* a file could not be opened or written      */
} ps_error;

/* Flags for ps_params.flags. */
#define PS_FLAG_REBUILD_EVERY_ITERATION 0x1u /* reference order: grid per iteration
                                                (gpu/src/particlesystem.cpp:171-197) */
#define PS_FLAG_SLEEPING 0x2u /* zero v and keep x when |x_new - x| < 1e-4
                                 (cpu/src/particle.h:60-65)                     */
/* The reference's CPU solver instead of the 3D path: Simulation::tick
 * (cpu/src/simulation.cpp:115-369, ITERATIVE).  Box boundaries are
 * constraints made at prediction and counted in the Jacobi divisor; each
 * iteration projects the boundaries, then every fluid object in creation
 * order (Gauss-Seidel across objects, Jacobi inside one,
 * totalfluidconstraint.cpp:41-116): lambda_j = 0 for neighbours outside the
 * object, dp = sum / rho0 / (neighbours + self + boundary count), tensile
 * term -k_corr (W/W(dq h))^n with k_corr unscaled.  Neighbours come from an
 * all-pairs scan summed in index order; no XSPH.  Particles enter through
 * ps_add_fluid_object, ps_add_gas_object (2D), emitters (2D) and, for
 * immovable solids only (phase PS_PHASE_SOLID, inverse mass 0),
 * ps_add_particles; grid and max_neighbors are unused. Set PS_FLAG_SLEEPING
 * too for the reference's behaviour (it always sleeps). */
#define PS_FLAG_CPU_REFERENCE 0x4u
/* Read neighbours from global memory even where the kernels could stage
 * each workgroup's neighbour cells in LDS (every 3D composition; in a batch
 * the workgroups whose particles lie in one scene).  Results are identical
 * either way. */
#define PS_FLAG_NO_LDS_STAGING 0x8u
/* With PS_FLAG_CPU_REFERENCE: the stabilisation pre-pass the reference
 * compiles under USE_STABILIZATION (cpu/src/simulation.h:17-18, off there):
 * after the constraints are made, STABILIZATION_ITERATIONS = 2 passes
 * project every boundary constraint with stabile = true
 * (simulation.cpp:247-270), which moves the start position p with ep and
 * applies no friction (boundaryconstraint.cpp:14-71); the fluid-and-gas
 * jitter of these projections is drawn from the counter stream
 * 2^63 | (iteration at the substep start mod 2^27) << 36 | pass << 35 |
 * index << 3 | axis.  The passes count towards no particle's divisor.  In
 * this mode's scope (immovable solids) no contact enters the pre-pass. */
#define PS_FLAG_STABILIZATION 0x10u
/* The reference's tensile term, s_corr = -k_corr (W(r)/W(dq h))^n with k_corr
 * taken as it is (lambdaCorr = -K_P pow(W/W(DQ_P H), E_P),
 * gpu/src/cuda/integration_kernel.cuh:628-633; K_P in
 * cpu/src/constraint/totalfluidconstraint.cpp:106).  Without it the 3D path
 * scales k_corr by h^2, the units of lambda (a length^2), so that one
 * dimensionless k works at any h: at h = 2 (the reference's H) the default
 * form is 4x the reference's term, at h = 0.2 it is 1/25 of it.
 * ps_params_default sets this flag. */
#define PS_FLAG_UNSCALED_TENSILE 0x20u

/* Status bits raised by the kernels, read with ps_get_status. */
#define PS_STATUS_NEIGHBOR_OVERFLOW 0x1u /* a particle had more than max_neighbors
                                            candidates within the search radius */
#define PS_STATUS_NONFINITE 0x2u         /* a predicted position was NaN/Inf      */
#define PS_STATUS_OUTSIDE_GRID 0x4u      /* a predicted position hashed outside the
                                            grid and was clamped to an edge cell   */
#define PS_STATUS_INTERNAL 0x80000000u   /* a kernel's own bounds check tripped
                                            (inconsistent grid tables); the write
                                            was skipped                             */

/*
 * Simulation parameters.  The reference splits these between SimParams
 * (gpu/src/cuda/kernel.cuh:9-22, uploaded by setParameters
 * gpu/src/cuda/integration.cu:96-100), the ParticleSystem constructor
 * (gpu/src/particlesystem.cpp:36-69) and compile-time #defines
 * (gpu/src/cuda/integration_kernel.cuh:20-42).  Here they are all runtime.
 */
typedef struct ps_params {
    uint32_t abi_version;   /* must be PS_ABI_VERSION                              */
    uint32_t dims;          /* 3, or 2 (z held fixed, grid[2] == 1)                */
    float radius;           /* particle radius r                                   */
    float h;                /* SPH smoothing radius                                */
    float cell;             /* grid cell edge; must be >= max(h, collide_dist)     */
    float collide_dist;     /* particle contact distance (reference 2.001 r)       */
    float gravity[3];
    float box_min[3];       /* world box; particles are kept r inside it           */
    float box_max[3];
    float origin[3];        /* grid origin (world position of cell (0,0,0))        */
    uint32_t grid[3];       /* cells per axis; positions outside clamp to the edge */
    uint32_t iterations;    /* solver iterations per substep                       */
    float relaxation;       /* lambda relaxation epsilon                           */
    float k_corr;           /* tensile correction s_corr = -k h^2 (W(r)/W(dq h))^n,
                               -k (W(r)/W(dq h))^n with PS_FLAG_UNSCALED_TENSILE  */
    uint32_t n_corr;
    float dq_corr;          /* dq as a fraction of h                               */
    float xsph_c;           /* XSPH viscosity c (0 disables)                       */
    float s_solid;          /* density weight of non-fluid neighbours of a fluid   */
    float mu_s;             /* particle-particle static friction                   */
    float mu_k;             /* particle-particle kinetic friction                  */
    float wall_mu_s;        /* wall friction (reference uses sqrt of these)        */
    float wall_mu_k;
    float wall_jitter;      /* wall push-out jitter as a fraction of r             */
    float shape_stiffness;  /* default shape-matching stiffness                    */
    uint32_t max_particles; /* capacity fixed at ps_create                         */
    uint32_t max_neighbors; /* neighbour-table width per particle                  */
    uint64_t seed;          /* counter-based RNG seed for wall jitter              */
    uint32_t flags;         /* PS_FLAG_*                                           */
    uint32_t reserved[7];
} ps_params;

/* Constants derived from ps_params on the host (FP64, rounded once to FP32). */
typedef struct ps_derived {
    float h2;          /* h^2                                                     */
    float poly6;       /* 315 / (64 pi h^9)        (integration_kernel.cuh:29)    */
    float spiky;       /* 45 / (pi h^6)            (integration_kernel.cuh:30)    */
    float w_zero;      /* poly6 * h^6 = W(0)                                      */
    float w_dq;        /* poly6 * (h^2 - (dq h)^2)^3 = W(dq h)                    */
    float inv_w_dq;    /* 1 / W(dq h)                                             */
    float tensile_k;   /* k_corr * h^2 (the tensile term in lambda's units, length^2),
                          or k_corr with PS_FLAG_UNSCALED_TENSILE                   */
    float search_r2;   /* neighbour search radius^2 = max(h, collide_dist)^2      */
    float collide_r2;  /* collide_dist^2                                          */
    float wall_s;      /* sqrt(wall_mu_s) * r                                     */
    float wall_k;      /* sqrt(wall_mu_k) * r                                     */
    float jitter;      /* wall_jitter * r                                         */
    uint32_t num_cells;
    float w_scale;     /* 1 / (h^2 - (dq h)^2)                                    */
    float w_scale_h2;  /* h^2 / (h^2 - (dq h)^2): s = w_scale_h2 - w_scale r^2 gives
                          W(r)/W(dq h) = s^3                                      */
    uint32_t reserved[1];
} ps_derived;

typedef struct ps_status {
    uint32_t flags;            /* PS_STATUS_* bits since the last ps_clear_status */
    uint32_t max_neighbors;    /* largest candidate count seen                    */
    uint32_t num_particles;
    uint32_t num_cells;
    uint64_t substeps;         /* substeps completed                              */
    uint64_t iterations;       /* solver iterations completed                     */
    uint32_t units;            /* workgroups of the last neighbour build          */
    uint32_t staged_units;     /* of those, read neighbours from LDS-staged cells */
} ps_status;

/* One scene of a batch (ps_create_batch): box (also its grid origin; the
 * grid is dense with params.cell), gravity = params.gravity * gravity_scale,
 * and the seed of its wall jitter.  A scene evolves exactly as a
 * single-scene system with these values would. */
typedef struct ps_scene_desc {
    float box_min[3];
    float box_max[3];
    float gravity_scale;
    uint32_t reserved0;
    uint64_t seed;
    uint32_t reserved[4];
} ps_scene_desc;

typedef struct ps_system ps_system;

/* ---- parameters (host only; no GPU needed) ------------------------------ */

/* Fill with the reference's GPU-demo constants (radius 0.25, 64^3 grid,
 * box (-50,0,-50)-(50,200,50), 5 iterations: gpu/src/particleapp.cpp:65-80;
 * fluid constants gpu/src/cuda/integration_kernel.cuh:20-42). */
int ps_params_default(ps_params *params);
/* Check a parameter set; returns PS_OK or PS_ERR_INVALID_ARGUMENT. */
int ps_params_validate(const ps_params *params);
/* Derived constants exactly as the kernels use them. */
int ps_derive(const ps_params *params, ps_derived *out);

/* ---- lifecycle ---------------------------------------------------------- */

/* Replaces ParticleSystem::ParticleSystem/_init (gpu/src/particlesystem.cpp:36-112),
 * initIntegration (gpu/src/cuda/integration.cu:43-48), allocateArray
 * (gpu/src/cuda/util.cu:12-15) and setParameters (integration.cu:96-100).
 * Allocates every buffer for max_particles up front. */
int ps_create(const ps_params *params, ps_system **out);
/* Many independent scenes in one handle (no reference counterpart: the
 * reference runs one scene per process).  Every kernel launch covers all
 * scenes; each scene's cells are a disjoint range of one global grid.
 * params supplies everything except the boxes, grids, gravity scales and
 * seeds, which come from scenes[0..nscenes). */
int ps_create_batch(const ps_params *params, uint32_t nscenes, const ps_scene_desc *scenes,
                    ps_system **out);
uint32_t ps_num_scenes(const ps_system *sys);

/* Replaces _finalize (gpu/src/particlesystem.cpp:115-135) and the
 * free*Vectors calls (integration.cu:74-94, solver.cu:158-178,
 * shared_variables.cu:12-23). */
int ps_destroy(ps_system *sys);
/* Replaces setParameters (integration.cu:96-100).  Capacities, grid and
 * dims cannot change after ps_create. */
int ps_set_params(ps_system *sys, const ps_params *params);
int ps_get_params(const ps_system *sys, ps_params *out);
const char *ps_last_error(const ps_system *sys);

/* ---- scene building ------------------------------------------------------ */

/* Replaces addParticleMultiple (gpu/src/particlesystem.cpp:333-348), i.e.
 * appendIntegrationParticle (integration.cu:50-72), appendPhaseAndMass
 * (shared_variables.cu:32-50) and appendSolverParticle (solver.cu:64-70).
 * pos, vel: n*3 floats (vel may be NULL for zero); inv_mass, rest_density:
 * n floats; phase: n ints.  *first_out (may be NULL) receives the index of
 * the first new particle.  Exceeding max_particles is PS_ERR_CAPACITY (the
 * reference silently drops the batch, particlesystem.cpp:335-336). */
int ps_add_particles(ps_system *sys, uint32_t n, const float *pos, const float *vel,
                     const float *inv_mass, const int32_t *phase,
                     const float *rest_density, uint32_t *first_out);

/* Append particles to scene `scene` of a batch (ps_add_particles adds to
 * scene 0).  A scene's particles must form one contiguous insertion range:
 * fill scenes in increasing order.  Constraint and body calls take the
 * global insertion indices. */
int ps_add_scene_particles(ps_system *sys, uint32_t scene, uint32_t n, const float *pos,
                           const float *vel, const float *inv_mass, const int32_t *phase,
                           const float *rest_density, uint32_t *first_out);

/* Add one fluid object: n particles (pos, vel and inverse mass as in
 * ps_add_particles, vel may be NULL) of one rest density.  Replaces
 * Simulation::createFluid (cpu/src/simulation.cpp:431-449).  In
 * PS_FLAG_CPU_REFERENCE mode the particles form one TotalFluidConstraint;
 * otherwise this is ps_add_particles with phase PS_PHASE_FLUID. */
int ps_add_fluid_object(ps_system *sys, uint32_t n, const float *pos, const float *vel,
                        const float *inv_mass, float rest_density, uint32_t *first_out);

/* Add one gas object (PS_FLAG_CPU_REFERENCE mode, dims = 2): n particles of
 * phase PS_PHASE_GAS forming one GasConstraint of the given rest density,
 * open (drag towards the rest density) or closed.  Replaces
 * Simulation::createGas (cpu/src/simulation.cpp:408-428) and the
 * GasConstraint it creates (cpu/src/constraint/gasconstraint.cpp:3-114):
 * the fluid constraint with the gas constants (gasconstraint.h:10-19), a
 * per-particle force from the open-gas drag and the vorticity term applied
 * at the next prediction, and gas gravity scaled by ALPHA = -0.2
 * (simulation.cpp:143-154).  PS_ERR_UNSUPPORTED outside that mode or in 3D. */
int ps_add_gas_object(ps_system *sys, uint32_t n, const float *pos, const float *vel,
                      const float *inv_mass, float rest_density, int open, uint32_t *first_out);

/* Add a smoke emitter (PS_FLAG_CPU_REFERENCE mode, dims = 2) at pos[0..3),
 * emitting particles_per_sec markers and, when object >= 0, as many gas
 * particles into gas object `object` (the last object added; no object may
 * follow it).  Replaces Simulation::createSmokeEmitter
 * (cpu/src/simulation.cpp:454-457) and OpenSmokeEmitter
 * (cpu/src/opensmokeemitter.cpp:3-46): at the end of every substep, in
 * creation order, the emitter's double-precision timer gains dt and pays
 * 1/particles_per_sec per emission; each emission adds a marker and a gas
 * particle of inverse mass 1 at rest at pos (at the next free index); then
 * every marker moves by dt times the poly6-weighted mean velocity of all
 * particles and is clamped to the box in x and y (simulation.cpp:348-362).
 * Markers are not simulated; at most max_markers per emitter.  A ps_step
 * whose emissions would exceed max_markers or max_particles fails with
 * PS_ERR_CAPACITY before its first substep.  ps_step_graph is
 * PS_ERR_UNSUPPORTED on a system with emitters. */
int ps_add_smoke_emitter(ps_system *sys, const float *pos, double particles_per_sec, int32_t object,
                         uint32_t max_markers, uint32_t *emitter_out);

/* Add a fluid emitter (PS_FLAG_CPU_REFERENCE mode, dims = 2) at pos[0..3)
 * feeding fluid object `object`, the last object added (nothing may be
 * added after it).  Replaces Simulation::createFluidEmitter
 * (cpu/src/simulation.cpp:459-461) and FluidEmitter::tick
 * (cpu/src/fluidemitter.cpp:13-79), run at the end of every substep after
 * the smoke emitters, in creation order:
 *  - every particle of the object slower than .06 at height <= 5 counts
 *    its timer t (4 at creation) down by 1 when lambdas[k] <= 0 and turns
 *    into an immovable solid at t <= 0 (inverse mass 0, phase SOLID, v = 0,
 *    out of the object); otherwise t gains dt, up to 3.  As in the
 *    reference, k is the particle's position in the object's list and
 *    `lambdas` is keyed by particle index: the lambda of particle k if that
 *    one was in the object at its last projection, else 0;
 *  - while the emitter's total time (double, this substep included) is
 *    below 5 s, one fluid particle of inverse mass 1 per 1/particles_per_sec
 *    of its timer, at pos with v = (u, 1, 0).  u replaces frand(): the top
 *    24 bits of SplitMix64(seed + golden (2^62 | emitter << 40 | emission
 *    index) + 1), divided by 2^24.
 * Solids (made here or added with ps_add_particles) make the substep find
 * contacts: ContactConstraint (contactconstraint.cpp:14-40) between a solid
 * and any other particle closer than 2 radius - 1e-4 after prediction,
 * projected with the height-scaled mass of Particle::scaleMass
 * (particle.h:67-73) before the boundaries of each iteration.  ps_step
 * fails with PS_ERR_CAPACITY before its first substep when emissions would
 * exceed max_particles; ps_step_graph is PS_ERR_UNSUPPORTED. */
int ps_add_fluid_emitter(ps_system *sys, const float *pos, double particles_per_sec, int32_t object,
                         uint32_t *emitter_out);

/* Markers of an emitter in emission order (OpenSmokeEmitter::getParticles,
 * opensmokeemitter.cpp:48-51): *count receives their number and, unless
 * xyz is NULL, xyz[3k..3k+2] their positions (room for *count). */
int ps_read_markers(ps_system *sys, uint32_t emitter, uint32_t *count, float *xyz);

/* Replaces addDistanceConstraint + updateOccurences (solver.cu:72-106,125-156).
 * pairs: 2*m particle indices; rest: m lengths; stiffness: m values in (0,1]
 * (NULL means 1). */
int ps_add_distance_constraints(ps_system *sys, uint32_t m, const uint32_t *pairs,
                                const float *rest, const float *stiffness);

/* Distance constraints added so far (the reference keeps their count in
 * its global constraint vectors, solver.cu:32-41); ps_step_backward writes
 * this many dl_dstiffness entries.  A ps_add_distance_constraints call that
 * fails adds none. */
uint32_t ps_num_distance_constraints(const ps_system *sys);

/* Replaces addPointConstraint (solver.cu:108-123): particle idx[k] is held
 * at xyz[3k..3k+2] after every iteration. */
int ps_add_point_constraints(ps_system *sys, uint32_t m, const uint32_t *idx,
                             const float *xyz);

/* Shape-matching cluster over particles idx[0..n) (rest shape = their
 * current positions).  The GPU reference has only a stub
 * (gpu/src/cuda/solver_kernel.cuh:289-312); the semantics follow the CPU
 * createRigidBody + TotalShapeConstraint (cpu/src/simulation.cpp:371-406,
 * cpu/src/constraint/totalshapeconstraint.cpp:14-24) in 3D.
 * stiffness <= 0 uses params.shape_stiffness. */
int ps_add_rigid_body(ps_system *sys, uint32_t n, const uint32_t *idx, float stiffness,
                      uint32_t *body_out);
/* The same for nbody clusters at once: members of body b are
 * idx[offsets[b] .. offsets[b+1]); stiffness: nbody values or NULL (default).
 * *first_body_out (may be NULL) receives the id of the first new body. */
int ps_add_rigid_bodies(ps_system *sys, uint32_t nbody, const uint32_t *offsets,
                        const uint32_t *idx, const float *stiffness, uint32_t *first_body_out);

/* Overwrite velocities of particles [first, first+n) (n*3 floats). */
int ps_set_velocities(ps_system *sys, uint32_t first, uint32_t n, const float *vel);

/* ---- the substep (hot path) --------------------------------------------- */

/* ParticleSystem::update (gpu/src/particlesystem.cpp:144-246) for
 * `substeps` substeps of dt each: predict, grid, neighbours, `iterations`
 * Jacobi iterations, velocity + XSPH.  No dt clamp (the reference clamps to
 * 0.05, :149).  Asynchronous. */
int ps_step(ps_system *sys, float dt, uint32_t substeps);

/* The same substep in its stages, for callers that interleave work and for
 * stage-level parity tests:
 *   begin   = integrateSystem (integration.cu:122-135) + calcHash,
 *             sortParticles, reorderDataAndFindCellStart (integration.cu:161-275)
 *             + neighbour collection (integration_kernel.cuh:303-356,481-559)
 *   iterate = one pass of the loop body (particlesystem.cpp:171-235): particle
 *             contacts, PBF lambda/delta-p, distance, shape, walls, pins
 *   end     = calcVelocity (integration.cu:416-426) + XSPH, positions committed */
int ps_substep_begin(ps_system *sys, float dt);
int ps_substep_iterate(ps_system *sys);
int ps_substep_end(ps_system *sys);

/* Finer stages (begin = predict + build_grid + find_neighbors). */
int ps_predict(ps_system *sys, float dt);
int ps_build_grid(ps_system *sys);
int ps_find_neighbors(ps_system *sys);

/* ps_step through a captured HIP graph: the first call (or the first after
 * any scene, parameter or dt change) records `substeps` substeps into a
 * graph; later calls replay it.  substeps must be a multiple of the buffer
 * rotation period (4 for an even iteration count, 6 for odd; 12 works for
 * both). */
int ps_step_graph(ps_system *sys, float dt, uint32_t substeps);

/* ---- reverse mode ------------------------------------------------------- */

/* Gradient of a loss L(x_T, v_T) of the state T = `substeps` substeps ahead
 * (the reference has no gradient path: this is the reverse of ps_step,
 * ParticleSystem::update, gpu/src/particlesystem.cpp:144-246).  Runs the
 * substeps from the current state, then walks them backwards.  Each substep
 * is reversed from a tape of its forward pass: its neighbour table, sorted
 * start state, every iterate (with lambda), the coordinates walls and pins
 * set and the wall-friction branches; the tables, contact sets and branches
 * are held fixed, so the result is the exact gradient of the smooth map they
 * select.  checkpoint_every = 0 (or >= substeps) tapes the whole horizon
 * during the forward pass; k > 0 keeps (x, v) every k substeps and one
 * segment of k substeps on tape: the last segment is taped by the forward
 * pass, each earlier one is recomputed from its checkpoint when the reverse
 * pass reaches it.  Inputs, host arrays in insertion order, NULL meaning
 * zero: dl_dx, dl_dv (3n).  Outputs (any may be NULL): dl_dx0, dl_dv0 (3n),
 * dl_drho0 (n: per-particle rest density), dl_dstiffness (one per distance
 * constraint, in the order they were added), dl_dmu_k (1: wall_mu_k).
 * Leaves the system in the state after the substeps, as ps_step would.
 * Covers fluid, gas and cloth particles, distance constraints, pins, walls
 * with friction and XSPH; PS_ERR_UNSUPPORTED for solid or rigid particles,
 * rigid bodies, sleeping, PS_FLAG_REBUILD_EVERY_ITERATION, batches and the
 * CPU-reference mode; PS_ERR_CAPACITY if a neighbour table overflowed (the
 * transposed sums need every pair).  Device memory, kept by the handle:
 * (4 max_neighbors + 56 + 69 iterations) bytes per particle per taped
 * substep, plus 52 per particle per checkpoint.  Synchronous.  dl_dmu_k
 * needs wall_mu_k > 0 (PS_ERR_INVALID_ARGUMENT otherwise: wall_k = sqrt(mu_k) r
 * has no finite derivative at 0). */
int ps_step_backward(ps_system *sys, float dt, uint32_t substeps, uint32_t checkpoint_every,
                     const float *dl_dx, const float *dl_dv, float *dl_dx0, float *dl_dv0,
                     float *dl_drho0, float *dl_dstiffness, float *dl_dmu_k);

int ps_synchronize(ps_system *sys);

/* ---- readback (synchronous) ---------------------------------------------- */

uint32_t ps_num_particles(const ps_system *sys); /* getNumParticles, particlesystem.h:48 */
/* Positions / velocities of particles [first, first+n) in insertion order,
 * n*3 floats.  The reference keeps positions in a GL VBO
 * (particlesystem.h:47); copyArrayFromDevice (util.cu:64-67) is its readback. */
int ps_read_positions(ps_system *sys, uint32_t first, uint32_t n, float *xyz);
int ps_read_velocities(ps_system *sys, uint32_t first, uint32_t n, float *xyz);

/* ---- export to device memory and trajectory files ------------------------ */

enum ps_field { PS_FIELD_POSITIONS = 1, PS_FIELD_VELOCITIES = 2, PS_FIELD_PHASES = 4 };

/* One field of particles [first, first+n) in insertion order, written to
 * DEVICE memory dst without a host round trip: PS_FIELD_POSITIONS as float4
 * (x, y, z, 1), the layout of the reference's position VBO
 * (particlesystem.cpp:95-96, mapped for the solver by mapGLBufferObject,
 * util.cu:39-62); PS_FIELD_VELOCITIES as float4 (vx, vy, vz, 0);
 * PS_FIELD_PHASES as int32.  dst is any device-accessible pointer, e.g. a GL
 * buffer mapped with hipGraphicsMapResources +
 * hipGraphicsResourceGetMappedPointer.  Enqueued on the system's stream
 * without waiting: ps_synchronize (or any readback) orders the caller after
 * it. */
int ps_export_device(ps_system *sys, uint32_t field, uint32_t first, uint32_t n, void *dst);

/* Binary trajectory files (the reference has no on-disk format).
 * Little-endian.  File header, 32 bytes: "PSTRAJ01" (8 bytes), uint32
 * version = 1, uint32 fields (PS_FIELD_* bits), uint32 dims, 12 zero bytes.
 * Each frame: uint64 substeps completed (ps_status.substeps), uint32 n,
 * uint32 fields, then per field in bit order, n records in insertion order:
 * positions and velocities as 3 float32, phases as int32.  The particle
 * count may change between frames (emitters).  ps_trajectory_write appends
 * the system's current state (it synchronises the system's stream). */
typedef struct ps_trajectory ps_trajectory;
int ps_trajectory_open(const char *path, uint32_t fields, uint32_t dims, ps_trajectory **out);
int ps_trajectory_write(ps_trajectory *traj, ps_system *sys);
int ps_trajectory_close(ps_trajectory *traj);

/* Internal sorted-order state, for parity tests and tools.
 * sorted_index[t] = insertion index of the particle in sorted slot t;
 * cell[t] = its cell key; cell_offsets: num_cells+1 entries, particles of
 * cell c occupy slots [off[c], off[c+1]) (reference cellStart/cellEnd,
 * integration_kernel.cuh:229-299). Any pointer may be NULL. */
int ps_read_grid(ps_system *sys, uint32_t *sorted_index, uint32_t *cell,
                 uint32_t *cell_offsets);
/* counts: n entries; table: n*max_neighbors entries, row t holds the sorted
 * slots of slot t's neighbours in ascending order (entries past the count
 * are unspecified). */
int ps_read_neighbors(ps_system *sys, uint32_t *counts, uint32_t *table);
/* Current iterate in sorted order: n*4 floats (x, y, z, lambda).
 * start: positions at the start of the substep, sorted order, n*3 floats. */
int ps_read_sorted_state(ps_system *sys, float *iterate_xyzl, float *start_xyz);

int ps_get_status(ps_system *sys, ps_status *out);
int ps_clear_status(ps_system *sys);

/* ---- per-kernel timing with HIP events on the handle's stream ----------- */

typedef enum ps_kernel_id {
    PS_K_PREDICT_HASH = 0,
    PS_K_SCAN,
    PS_K_SCATTER,
    PS_K_REORDER,
    PS_K_NEIGHBORS,
    PS_K_LAMBDA,
    PS_K_SOLVE,
    PS_K_SHAPE,
    PS_K_FINALIZE,
    /* reverse pass (ps_step_backward) */
    PS_K_ADJ_XSPH,
    PS_K_ADJ_FINALIZE,
    PS_K_ADJ_PRE,       /* no launches: the pre-step ends adj_finalize and adj_position */
    PS_K_ADJ_LAMBDA,
    PS_K_ADJ_POSITION,
    PS_K_ADJ_STIFFNESS, /* no launches: adj_position forms dL/dstiffness */
    PS_K_ADJ_PREDICT,
    /* insertion-order export (ps_export_device, trajectory files) */
    PS_K_EXPORT,
    PS_K_COUNT
} ps_kernel_id;

/* kernel_mask: bit k brackets every launch of kernel k with two events
 * (PS_TIMING_ALL for every kernel, 0 to stop).  Each event pair costs a few
 * microseconds of stream time, so time only the kernels you report. */
#define PS_TIMING_ALL 0xffffffffu
int ps_timing_enable(ps_system *sys, uint32_t kernel_mask);
int ps_timing_reset(ps_system *sys);
/* Total milliseconds and launch count for a kernel since the last reset. */
int ps_timing_read(ps_system *sys, int kernel_id, double *total_ms, uint64_t *launches);
/* Bracket only one launch in `period` of each enabled kernel (1, the
 * default: every launch), counting from the call or the last
 * ps_timing_reset; ps_timing_read then reports the bracketed launches.  A
 * period coprime with the iteration count samples every iteration alike. */
int ps_timing_sample(ps_system *sys, uint32_t period);
const char *ps_kernel_name(int kernel_id);

/* ---- stateless grid sort over caller device arrays --------------------- */

/* The grid of a stateless sort: the grid fields of the reference's SimParams
 * (worldOrigin, cellSize, gridSize; kernel.cuh:9-22). */
typedef struct ps_grid_desc {
    float origin[3];
    float cell;        /* edge of a cubic cell, > 0                          */
    uint32_t grid[3];  /* cells per axis, each > 0, product < 2^31           */
    uint32_t reserved;
} ps_grid_desc;

/* Hash, stable sort and cell bounds of n positions, all in device memory,
 * in one call: the reference's calcHash + sortParticles +
 * reorderDataAndFindCellStart sequence (wrappers.cuh:30-46; integration.cu:
 * 161-275) over the caller's arrays, with the cells and order of ps_step:
 *   pos          n records (x, y, z, w) of float4, the reference's layout;
 *   cell_key     out, n: the cell of each sorted slot, (z gy + y) gx + x with
 *                each coordinate floor((p - origin) / cell) clamped into
 *                the grid (the reference wraps, calcGridHash :197-205);
 *   sorted_index out, n: the input index at each sorted slot, ordered by
 *                (cell, input index) (a stable sort of the keys);
 *   cell_start, cell_end  out, one per cell: sorted slots [start, end) of the
 *                cell, both 0xffffffff for an empty cell (the reference sets
 *                cellStart so and leaves cellEnd unset, integration.cu:198);
 *   sorted_pos   out, n float4 records in sorted order, or NULL.
 * stream is a hipStream_t (NULL: the null stream); the call returns after
 * enqueueing its work.  Its workspace (12 n + 8 cells bytes and a little)
 * is allocated and freed on the stream for each call, as the reference's
 * reorderDataAndFindCellStart allocates a copy of the positions per call
 * (integration.cu:216-220).  PS_ERR_INVALID_ARGUMENT for a null pointer
 * (sorted_pos aside) or a bad grid; PS_ERR_OUT_OF_MEMORY if the workspace
 * cannot be had. */
int ps_grid_sort(const ps_grid_desc *grid, uint32_t n, const float *pos, uint32_t *cell_key,
                 uint32_t *sorted_index, uint32_t *cell_start, uint32_t *cell_end,
                 float *sorted_pos, void *stream);

/* ps_grid_sort without any allocation: the workspace is a fixed device array
 * inside the library (one per GPU, held until the process ends), for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops).
 * No counterpart in the reference.  Limits: n <= PS_GRID_NOALLOC_MAX_PARTICLES
 * and grid cells <= PS_GRID_NOALLOC_MAX_CELLS, else PS_ERR_CAPACITY.  Every
 * call on one GPU shares that GPU's array, so calls that overlap on one GPU
 * (on different streams, or from different threads) race; calls ordered on
 * one stream do not. */
#define PS_GRID_NOALLOC_MAX_PARTICLES (1u << 20)
#define PS_GRID_NOALLOC_MAX_CELLS (1u << 21)
int ps_grid_sort_noalloc(const ps_grid_desc *grid, uint32_t n, const float *pos, uint32_t *cell_key,
                         uint32_t *sorted_index, uint32_t *cell_start, uint32_t *cell_end,
                         float *sorted_pos, void *stream);

/* ---- autotuning -------------------------------------------------------- */

/* Choose the workgroup size (64, 128, 256 or 512) of each per-particle
 * kernel (predict_hash, reorder, neighbors, lambda, solve, finalize) by
 * timing `substeps` substeps of this system per candidate with events, and
 * whether the lambda, solve and finalize kernels stage each workgroup's
 * neighbour cells in LDS (whole substeps
 * with staging off and on, alternately twice; the choice shows in
 * ps_status.staged_units, zero when off).  Neither choice changes any
 * result.  The simulation state is restored afterwards.  Choices are
 * cached per (kernel, log2 particle count, scene composition) for the
 * process and, if the environment names a file in PS_AUTOTUNE_CACHE,
 * across processes.  blocks_out (PS_K_COUNT entries, may be NULL) receives
 * the sizes in use. */
int ps_autotune(ps_system *sys, float dt, uint32_t substeps, uint32_t *blocks_out);
int ps_get_tuning(const ps_system *sys, uint32_t *blocks_out);
int ps_set_tuning(ps_system *sys, const uint32_t *blocks);

#ifdef __cplusplus
}
#endif

#endif /* PS_H */
