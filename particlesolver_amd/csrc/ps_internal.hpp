// ps_internal.hpp — shared between the host API (ps_api.hip) and the
// kernels (ps_kernels.hip).  Device-state layout and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/ps.h"

namespace ps {

// Per-particle attributes that travel with a particle through every reorder.
struct __align__(16) Attr {
    float w;        // inverse mass
    int32_t phase;  // PS_PHASE_*
    float rho0;     // rest density
    uint32_t orig;  // insertion index
};

// Device-resident status word (read by ps_get_status).
struct DevStatus {
    uint32_t flags;
    uint32_t max_neighbors;
    uint64_t iter;  // solver iterations completed (keys the wall-jitter RNG)
};

// One scene of a batch: its own box, dense grid (cells cell_base ..
// cell_base + gx*gy*gz - 1 of the global cell index), gravity and RNG seed.
// Particles of a scene are one contiguous insertion range starting at first.
struct __align__(16) SceneDev {
    float box_min[3];
    float cell;
    float box_max[3];
    uint32_t cell_base;
    float origin[3];
    uint32_t first;
    float gravity[3];
    uint32_t pad;
    uint32_t grid[3];
    uint32_t pad2;
    uint64_t seed;
    uint64_t pad3;
};

// Constants handed to every kernel by value (kernarg segment).
struct KConst {
    float h, h2, poly6, spiky, w_zero, w_dq, inv_w_dq, search_r2, collide_r2, collide_dist;
    // 1/(h^2 - (dq h)^2) and h^2 times it: with s = fma(-w_scale, r^2,
    // w_scale_h2), W(r)/W(dq h) = s^3
    float w_scale, w_scale_h2;
    float wall_s, wall_k, jitter, radius, relax, k_corr, xsph_c, s_solid, mu_s, mu_k;
    float gravity[3];
    float box_min[3], box_max[3];
    float origin[3];
    float cell;
    uint32_t grid[3];
    uint32_t n_corr;
    uint32_t dims;
    uint64_t seed;
    uint32_t nscenes;  // > 1: per-particle scene table lookups
};

// Constants of the CPU-reference mode (ps_cpuref.hip), rounded once to
// float on the host in the order the reference evaluates them.
struct CrConst {
    float h, h2, poly6, spiky, w_dq, relax, k_corr, jitter, radius, sf, kf;
    float touch;  // contact test: closer than a diameter less EPSILON (1e-4)
    float g_wdq;  // gas objects: poly6(DQ_P^2 H^2) of GasConstraint
    float gravity[3];
    float box_min[3], box_max[3];
    uint32_t n_corr;
    uint32_t dims;
    uint64_t seed;
};

// Scene-composition switches chosen on the host per launch; each selects a
// separately compiled variant of the solve kernel, so a pure-fluid scene runs
// without the contact/distance/shape code paths.
enum SolveFeature : uint32_t {
    F_FLUID = 1u,
    F_CONTACT = 2u,
    F_DIST = 4u,
    F_SHAPE = 8u,
    F_PIN = 16u,
};

// LDS staging of neighbour cells (ps_kernels.hip).  A unit is one
// workgroup of kStageUnit consecutive sorted particles; its particles lie in
// at most kStageRows grid rows, and the candidates of a particle in row q
// and stencil row t = (dz+1)*3 + (dy+1) form a slot range inside interval
// 9q + t: the slots of cells x in [xa_q - 1, xb_q + 1] of the target row.
// The intervals are laid out in LDS back to back, each padded to a whole
// number of kStageChunk-slot chunks; a staged unit's table holds LDS
// indices, and src[c] is the global slot of the first element of chunk c.
constexpr uint32_t kStageUnit = 256;
constexpr uint32_t kStageRows = 3;
constexpr uint32_t kStageIv = 9 * kStageRows;
constexpr uint32_t kStageSlots = 3072;
// the staged neighbour build: slots of one stencil row a wavefront copies
// into its LDS (x, y, z: 3 KB)
constexpr uint32_t kWaveRowSlots = 256;
constexpr uint32_t kStageChunk = 16;
constexpr uint32_t kStageChunks = kStageSlots / kStageChunk;
// LDS index of a staged unit's sentinel record, one past the staged slots:
// a position at +infinity, whose squared distance to every particle fails
// the pair tests.  A staged unit's table pads its last chunk with it, so
// the staged kernels visit whole chunks without bound checks.
constexpr uint32_t kStageSentinel = kStageSlots;

// Contacts a particle of the global-slot solve lists in LDS during its
// counting walk; the corrections visit the listed ones, or walk the whole
// table again when there are more.
constexpr uint32_t kContactList = 16;
// the same per lane in the staged solve (two lanes per particle, each over
// one half of the table)
constexpr uint32_t kHalfContactList = 8;

struct StageDesc {
    uint32_t staged;              // 1: the unit's table entries are LDS indices
    uint32_t nchunks;             // chunks in use
    uint32_t walk;                // 1: a particle of the unit walks its table for contacts
    uint32_t src[kStageChunks];   // global slot of each chunk's first element
};

struct DevBuffers {
    // particle state, sorted storage order
    float4 *X, *X2;  // positions at substep start (x*); .w: in scenes that are not
                     // a uniform fluid, the sort's XSPH mass (1/w fluid, -1 other)
    float4 *P, *Q;   // iterate (x, y, z, lambda) ping-pong
    float4 *V, *V2;  // velocities
    Attr *A, *A2;
    uint32_t *key, *key2;  // cell of each slot
    uint32_t *rank;        // arrival rank inside the cell (from the hash atomics)
    uint32_t *tmp_orig;    // insertion index by provisional sorted slot
    uint32_t *slot_of;     // insertion index -> current slot (kept only while
                           // track_slots: distance constraints or bodies exist)
    uint32_t track_slots;
    uint32_t *sid, *sid2;  // scene of each slot (batches only)
    float *mj;             // per slot, written by the sort for scenes that are not
                           // a uniform fluid: the particle's density mass as a
                           // neighbour, 1/w (times s_solid if not fluid), or -1
                           // for w = 0 (left out of the fluid sums)
    int32_t *phs;          // per slot, written with mj: the phase, its sign bit
                           // set for w = 0 (the staged solve's contact records)
    uint32_t contacts;     // the scene has contact phases (then phs is written)
    uint32_t solve_units;  // staged solve launches: 0 every unit, 1 the units without a
                           // contact-walking particle, 2 the units with one
    uint32_t split_solve;  // such particles are rare: launch the two classes apart
    SceneDev *scenes;
    // grid
    uint32_t *cell_count;  // num_cells, zero between builds
    uint32_t *cell_off;    // num_cells + 1
    uint32_t *scan_part;   // per scan tile
    // neighbours: sliced ELL in chunks of four (see ps_kernels.hip)
    uint32_t *nbr;
    uint32_t *ncnt;
    uint32_t *nbr_glob;    // recording a substep for the reverse pass with staging:
                           // the same table with global slots (else null)
    uint32_t kq;  // max_neighbors / 4
    // distance constraints (insertion-index CSR)
    uint32_t *dist_off;    // max_particles + 1
    uint32_t *dist_adj;    // 2*m: constraint ids
    uint2 *dist_pairs;     // m
    float2 *dist_rs;       // m: (rest, stiffness)
    uint32_t *occ;         // per particle incidence count (distance + pins)
    // pins
    uint32_t *pin_of;      // per particle: pin + 1 (0 = none)
    float4 *pin_xyz;
    // shape matching
    uint32_t *body_off;    // nbody + 1
    uint32_t *body_idx;    // insertion indices
    float4 *body_rest;     // rest offsets
    float *body_stiff;
    float *body_mass;
    float4 *body_q;        // rotation quaternion (w, x, y, z)
    float4 *body_goal;     // per member
    uint32_t *body_of;     // per particle: body + 1
    uint32_t *member_of;   // per particle: member index
    DevStatus *status;
    // LDS staging of neighbour cells: one descriptor per unit
    StageDesc *stage;
    uint32_t staging;      // 1: the last neighbour build staged its units
    // adjoint recompute only (else null): per particle, the coordinates the
    // walls and pins set and the wall-friction branch (walls() bits)
    uint8_t *clamp_rec;
    float4 *wall_rec;      // with clamp_rec: position before wall friction,
                           // normal code (n_x+1)*9 + (n_y+1)*3 + (n_z+1) in .w
    float4 *tape_lq;       // lambda pass: S = sum_j a d (xyz), D (w)
    float4 *tape_pit;      // lambda pass: the iterate after it (x, y, z, lambda)
    float4 *tape_x;        // reorder: the sorted start state, its attributes and
    Attr *tape_a;          // (with track_slots) the slot map, also to tape
    uint32_t *tape_s;
    float4 *tape_q;        // last solve of a taped substep: its result, also to tape
    float *tape_rho;       // lambda pass: density
    float4 *tape_f;        // solve: fluid sum f (xyz), contact count bits (w)
    // CPU-reference mode only
    float4 *force;         // per particle, applied at the next prediction (gas)
    float *lam;            // lambda per particle
    uint32_t *bc;          // boundary flags | no-jitter bit | constraint count << 8
    float *tmass;          // height-scaled mass of the contacts (per substep)
    float *ttimer;         // FluidEmitter solidification timer t
    uint32_t *cnum;        // contacts of a movable particle (<= kCrMaxContacts)
    uint32_t *clist;       // their immovable partners, kCrMaxContacts per particle
    uint32_t *sflag;       // scratch of the solidification pass
};

// Contacts kept per movable particle in the CPU-reference mode; more raise
// PS_STATUS_NEIGHBOR_OVERFLOW and the rest are left out.
constexpr uint32_t kCrMaxContacts = 32;

// Per-particle attributes the reverse pass gathers from neighbours.
struct __align__(16) AdjAttr {
    float w;
    int32_t phase;
    uint32_t nc;  // contact count
    float mass;   // density multiplier as a neighbour: 1/w, times s_solid if not fluid
};

// Work buffers of the reverse pass (ps_adjoint.hip), storage order unless
// noted.
struct AdjBuffers {
    const float4 *Pit;     // iterate of the iteration being reversed (x, y, z, lambda)
    float4 *pb;            // dL/dp before it (finalize: of the final iterate)
    const float4 *SD;      // recorded lambda pass: S = sum_j a d (xyz), D (w)
    const float *rho;      // recorded lambda pass: density
    float4 *Y;             // masked dL/dy (xyz); finalize: dL/dS
    float4 *M;             // mu = dL/dy / rho0 (fluid) or dL/dy (xyz), 1/rho0 (w)
    AdjAttr *AB;           // attributes with contact counts (non-uniform scenes)
    float4 *MC;            // mu as in M (xyz), Cr = dL/dC / rho0 (w); Cr = 0 if not fluid
    float4 *KV;            // V = dL/dD sk S (xyz), u = dL/dD sk^2 (w), sk = spiky / rho0
    uint32_t *nc;          // contact count
    float *r0b, *wkb;      // dL/drho0 and dL/dwall_k of this substep
    float4 *xsb;           // dL/dx* of this substep
    float *sb;             // dL/dstiffness, per distance constraint
    float4 *gx, *gv;       // insertion order: dL/dx, dL/dv of the state
    float *drho, *dwk;     // insertion order: accumulated dL/drho0, dL/dwall_k
    uint32_t contacts;     // the scene has contact phases
    uint32_t dist;         // the scene has distance constraints
    // The pre-step of the next reversed iteration (pins, walls and wall
    // friction backwards), run per particle by adj_finalize_k and
    // adj_position_k on the dL/dp they form, with that iteration's records;
    // it writes Yn and ABn (the other buffer of a pair, as the kernel's own
    // neighbours still read Y and AB), M, nc, and updates xsb, wkb, r0b.
    // clampn null: no next iteration, pb is written instead.
    const uint8_t *clampn;
    const float4 *wrecn, *tfn;
    float4 *Yn;
    AdjAttr *ABn;
};

// Timing of one launch by the dispatch itself: the host puts an event pair
// here before calling a launcher, and the launcher's kernel, started with
// PS_LAUNCH, takes it (hipExtLaunchKernelGGL records the pair at the
// kernel's own start and end, so no stream interval outside the kernel lies
// between them) and clears it.  Launchers of two kernels leave it (the scan)
// or split it (the solve of a scene with contacts: the first launch records
// the start, the second the end).
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchEvents g_launch_events;
}  // namespace ps

#include <hip/hip_ext.h>

#define PS_LAUNCH(kern, grid, block, lds, st, ...)                                               \
    do {                                                                                         \
        const ps::LaunchEvents ev_ = ps::g_launch_events;                                        \
        ps::g_launch_events = ps::LaunchEvents{};                                                \
        if (ev_.start || ev_.stop)                                                               \
            hipExtLaunchKernelGGL(kern, grid, block, lds, st, ev_.start, ev_.stop, 0u, __VA_ARGS__); \
        else                                                                                     \
            hipLaunchKernelGGL(kern, grid, block, lds, st, __VA_ARGS__);                         \
    } while (0)

namespace ps {

// Kernel launchers (ps_kernels.hip).  All enqueue on `st` and return the
// launch error.
// bs: workgroup size (64..512, a multiple of 64) of the tunable kernels
hipError_t launch_predict_hash(hipStream_t st, const DevBuffers &b, const KConst &k, uint32_t n,
                               float dt, bool predict, bool hash, uint32_t bs);
hipError_t launch_scan(hipStream_t st, const DevBuffers &b, uint32_t num_cells);
hipError_t launch_scatter(hipStream_t st, const DevBuffers &b, uint32_t n);
hipError_t launch_reorder(hipStream_t st, const DevBuffers &b, uint32_t n, uint32_t bs, float mj_solid);
hipError_t launch_neighbors(hipStream_t st, const DevBuffers &b, const KConst &k, uint32_t n,
                            uint32_t kmax, uint32_t bs);
hipError_t launch_lambda(hipStream_t st, const DevBuffers &b, const KConst &k, uint32_t n,
                         bool uniform, float wu, uint32_t bs);
hipError_t launch_shape(hipStream_t st, const DevBuffers &b, uint32_t nbody);
// it: iteration index inside the current substep (the RNG key is the device
// iteration counter + it); finalize advances the counter by `its`.
hipError_t launch_solve(hipStream_t st, const DevBuffers &b, const KConst &k, uint32_t n,
                        uint32_t features, bool uniform, float wu, uint32_t it, uint32_t bs);
hipError_t launch_finalize(hipStream_t st, const DevBuffers &b, const KConst &k, uint32_t n,
                           float dt, bool sleeping, bool uniform, float wu, float wf, uint32_t its,
                           uint32_t bs);

uint32_t scan_tiles(uint32_t num_cells);
// cells per scan tile (launch_scan's partials: one per tile)
constexpr uint32_t kScanTileCells = 4096;

// Records msg as the calling thread's last error (ps_last_error(NULL)) and
// returns code: for the entry points without a system handle.
int fail_stateless(int code, const std::string &msg);

// Reverse pass (ps_adjoint.hip), one kernel per call: velocity + XSPH
// backwards (ADJ_XSPH then ADJ_FINALIZE; b.P the final iterate, b.X the
// substep start), one iteration backwards (ADJ_LAMBDA, ADJ_POSITION, which
// also adds dL/dstiffness per distance constraint; a.Pit, a.SD, a.rho of
// that iteration), predict backwards (ADJ_PREDICT).  ADJ_FINALIZE and
// ADJ_POSITION end with the next iteration's pre-step (AdjBuffers::clampn),
// or write a.pb.  count: particles.
enum AdjKernel : uint32_t {
    ADJ_XSPH,
    ADJ_FINALIZE,
    ADJ_LAMBDA,
    ADJ_POSITION,
    ADJ_PREDICT,
};
// uniform: every particle fluid with one inverse mass (no attribute gathers)
hipError_t launch_adj(hipStream_t st, AdjKernel which, const DevBuffers &b, const KConst &k,
                      const AdjBuffers &a, uint32_t count, float dt, bool uniform);

// CPU-reference mode (ps_cpuref.hip); storage in insertion order, iterate
// ping-pong b.P -> b.Q per fluid object.
hipError_t launch_cr_predict(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n,
                             float dt, bool boundaries);
// lambda, then delta + Jacobi update (b.P -> b.Q) of one fluid or gas
// object [first, first + count); gas: 0 fluid, 1 closed gas, 2 open gas;
// next_iter >= 0 also projects the boundaries of that iteration of the
// substep
hipError_t launch_cr_lambda(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n,
                            uint32_t first, uint32_t count, float rho0, int gas);
hipError_t launch_cr_delta(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n,
                           uint32_t first, uint32_t count, float rho0, int gas, int next_iter);
hipError_t launch_cr_boundary(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n,
                              uint32_t it);
hipError_t launch_cr_end(hipStream_t st, const DevBuffers &b, uint32_t n, float dt, bool sleeping,
                         uint32_t its);
// smoke emitters: `count` (<= 64) emissions at pos into markers[first_mark..]
// and, unless first_part == ~0u, gas particles of rest density rho0 at
// [first_part, +count);
// then the advection of markers[0, nmark) through the n particles
hipError_t launch_cr_emit(hipStream_t st, const DevBuffers &b, float4 *markers, uint32_t first_mark,
                          uint32_t first_part, uint32_t count, const float *pos, float rho0);
// insertion-order export of field PS_FIELD_* of records [first, first+count)
hipError_t launch_export(hipStream_t st, const DevBuffers &b, uint32_t n, uint32_t field,
                         uint32_t first, uint32_t count, void *dst);
hipError_t launch_cr_markers(hipStream_t st, const DevBuffers &b, float4 *markers, const CrConst &c,
                             uint32_t nmark, uint32_t n, float dt);
// immovable solids present: contacts found after prediction (which then
// leaves the boundaries to the contact pass), projected with the
// boundaries at the start of every iteration, E -> E2 (b.P -> b.Q)
hipError_t launch_cr_find_contacts(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n);
// PS_FLAG_STABILIZATION: the two stabilisation passes over the boundaries
// (b.P and b.X both move)
hipError_t launch_cr_stabilize(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n);
hipError_t launch_cr_contacts(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n,
                              uint32_t it);
// FluidEmitter: solidification over the object [first, first+count), then
// `count` emissions of emitter `eidx` (emission counter `emitted`) at
// first_part
hipError_t launch_cr_solidify(hipStream_t st, const DevBuffers &b, uint32_t first, uint32_t count,
                              float dt);
hipError_t launch_cr_emit_fluid(hipStream_t st, const DevBuffers &b, const CrConst &c,
                                uint32_t first_part, uint32_t count, const float *pos, float rho0,
                                uint32_t eidx, uint64_t emitted);

}  // namespace ps
