// ps_api.hip — host implementation of the C-ABI in include/ps.h.
//
// Owns the device buffers of one particle system and enqueues the substep
// kernels on the system's own stream.  No process globals: every handle is
// independent (the reference's state is global thrust::device_vectors,
// gpu/src/cuda/integration.cu:23-34, solver.cu:32-41,
// shared_variables.cu:8-10).

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "ps_internal.hpp"

using ps::Attr;
using ps::DevBuffers;
using ps::KConst;

namespace {

struct TimedLaunch {
    int id;
    hipEvent_t start, stop;
    bool by_dispatch;  // the events ride on the kernel's dispatch (PS_LAUNCH)
};

}  // namespace

struct ps_system {
    ps_params prm;
    ps_derived der;
    KConst kc;
    DevBuffers b;
    hipStream_t stream = nullptr;
    // the reverse pass's own stream (created by the first ps_step_backward):
    // it reverses one segment while `stream` recomputes the one before
    hipStream_t rstream = nullptr;
    // the stream timed launches are enqueued on (null: `stream`)
    hipStream_t tstream = nullptr;
    uint32_t n = 0, cap = 0, ncell = 0, kmax = 0, tiles = 0;
    // scene composition
    bool uniform = true;     // every particle fluid with one inverse mass
    float wu = 1.0f;
    // every fluid particle of one inverse mass wf (any other phases): staged
    // XSPH then takes the fluid mass as a constant
    bool fluid_mass_uniform = true;
    bool any_fluid = false;
    float wf = 1.0f;
    bool any_contact_phase = false;
    uint64_t n_walk = 0;  // particles that walk their table for contacts (phase >= cloth, w != 0)
    bool any_solid_phase = false;  // SOLID or RIGID particles present
    // host copies of the constraint sets (insertion indices)
    std::vector<uint32_t> dist_pairs;
    std::vector<float> dist_rest, dist_stiff;
    std::vector<uint32_t> occ;
    std::vector<uint32_t> pin_of;
    std::vector<float> pin_xyz;
    std::vector<uint32_t> body_off{0}, body_idx, body_of, member_of;
    std::vector<float> body_rest, body_stiff, body_mass;
    size_t dist_cap = 0, pin_cap = 0, body_cap = 0, member_cap = 0;
    // scenes (one unless created with ps_create_batch)
    std::vector<ps::SceneDev> scenes;
    std::vector<uint32_t> scene_count;
    uint32_t last_scene = 0;
    // captured substep graph (ps_step_graph)
    hipGraphExec_t gexec = nullptr;
    float g_dt = 0.0f;
    uint32_t g_sub = 0;
    // substep state
    uint32_t it_sub = 0;    // iterations done in the open substep
    int stage = 0;          // 1 between predict and end
    bool hashed = false;    // keys/ranks of the current storage are valid
    bool nbr_valid = false; // neighbour table matches the current storage
    float dt = 0.0f;
    uint64_t substeps = 0, iters = 0;
    // CPU-reference mode (PS_FLAG_CPU_REFERENCE)
    bool cpuref = false;
    ps::CrConst cr;
    struct FluidObject {
        uint32_t first, count;
        float rho0;
        int gas;  // 0 fluid, 1 closed gas, 2 open gas
    };
    std::vector<FluidObject> objects;
    // smoke emitters (OpenSmokeEmitter), ticked in creation order at the end
    // of every substep; `object` indexes `objects` (-1: markers only) and is
    // the last object, so its particles stay one contiguous range
    struct Emitter {
        double timer, rate;
        float pos[3];
        int32_t object;
        uint32_t nmark, cap;
        float4 *mark;
    };
    std::vector<Emitter> emitters;
    // fluid emitters (FluidEmitter), ticked after the smoke emitters; each
    // feeds and solidifies fluid object `object` (the last object)
    struct FluidEmitter {
        double timer, total, rate;
        float pos[3];
        int32_t object;
        uint64_t emitted;
    };
    std::vector<FluidEmitter> femitters;
    bool cr_bnd_done = false;  // boundaries of iteration it_sub already projected
    bool cr_swapped = false;   // iterate lives in b.Q (odd number of object passes)
    // workgroup size of each kernel (ps_autotune)
    // LDS staging of neighbour cells as ps_autotune chose it for this
    // composition: -1 not tuned (staged where the kernels allow it), 0 off,
    // 1 on
    int stage_choice = -1;
    uint32_t block[PS_K_COUNT] = {256, 256, 256, 256, 256, 256, 256, 256, 256, 256,
                                  256, 256, 256, 256, 256, 256, 256};  // reverse pass, export: fixed 256
    // timing
    uint32_t timing = 0;  // kernel mask
    uint32_t timing_period = 1;     // bracket one launch in this many
    uint64_t seen[PS_K_COUNT] = {};  // launches of each enabled kernel
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> free_events;
    double total_ms[PS_K_COUNT] = {};
    uint64_t launches[PS_K_COUNT] = {};
    std::string err;
    // reverse pass (ps_step_backward): workspace grown on demand and kept
    struct AdjWork {
        size_t cap = 0, cks = 0, tslots = 0, iters = 0, ndist = 0, tab = 0;
        // checkpoints of the per-particle state, cks slots x cap
        float4 *ckX = nullptr, *ckV = nullptr;
        Attr *ckA = nullptr;
        uint32_t *ckS = nullptr;
        ps::DevStatus *ckSt = nullptr;
        // tape of one segment, tslots substeps: neighbour table (tab words
        // per slot) and counts, storage-order start state, attributes, slot
        // map and final iterate; per iteration the iterate, clamp bits and
        // wall-friction records
        uint32_t *tN = nullptr, *tC = nullptr, *tS = nullptr;
        float4 *tX = nullptr, *tP = nullptr, *tPit = nullptr, *tW = nullptr, *tLQ = nullptr, *tF = nullptr;
        float *tRho = nullptr;
        Attr *tA = nullptr;
        uint8_t *tClamp = nullptr;
        // reverse-pass work arrays
        float4 *pb = nullptr, *Y = nullptr, *Y2 = nullptr, *M = nullptr, *xsb = nullptr;
        ps::AdjAttr *AB = nullptr, *AB2 = nullptr;
        float4 *gx = nullptr, *gv = nullptr;
        float4 *MC = nullptr, *KV = nullptr;
        uint32_t *nc = nullptr;
        float *r0b = nullptr, *wkb = nullptr, *drho = nullptr, *dwk = nullptr, *sb = nullptr;
    } adj;
    bool adj_active = false;  // inside ps_step_backward: neighbours stay global
};

namespace {

thread_local std::string g_err;

int fail(ps_system *s, int code, const std::string &msg)
{
    if (s) s->err = msg;
    g_err = msg;
    return code;
}

#define HIP_TRY(sys, call)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((sys), e_ == hipErrorOutOfMemory ? PS_ERR_OUT_OF_MEMORY : PS_ERR_DEVICE, \
                        std::string(#call) + ": " + hipGetErrorString(e_));                \
    } while (0)

bool is_finite(float x) { return isfinite(x); }

void derive_constants(const ps_params &p, ps_derived &d)
{
    // POLY6 = 315/(64 pi h^9), SPIKEY = 45/(pi h^6)
    // (gpu/src/cuda/integration_kernel.cuh:29-30), in FP64, rounded once.
    const double pi = 3.14159265358979323846;
    const double h = p.h, h2 = h * h, h3 = h2 * h, h6 = h3 * h3, h9 = h6 * h3;
    const double poly6 = 315.0 / (64.0 * pi * h9);
    const double spiky = 45.0 / (pi * h6);
    const double dq = (double)p.dq_corr * h;
    const double tq = h2 - dq * dq;
    const double sr = p.h > p.collide_dist ? (double)p.h : (double)p.collide_dist;
    memset(&d, 0, sizeof(d));
    d.h2 = (float)h2;
    d.poly6 = (float)poly6;
    d.spiky = (float)spiky;
    d.w_zero = (float)(poly6 * h6);
    d.w_dq = (float)(poly6 * tq * tq * tq);
    d.inv_w_dq = (float)(1.0 / (poly6 * tq * tq * tq));
    d.w_scale = (float)(1.0 / tq);
    d.w_scale_h2 = (float)(h2 / tq);
    // The tensile term is added to lambda (units length^2); the paper's
    // dimensionless k measures lengths in units of h, hence k h^2.  The
    // reference takes K_P as it is (integration_kernel.cuh:633).
    d.tensile_k = (p.flags & PS_FLAG_UNSCALED_TENSILE) ? p.k_corr : (float)((double)p.k_corr * h2);
    d.search_r2 = (float)(sr * sr);
    d.collide_r2 = (float)((double)p.collide_dist * (double)p.collide_dist);
    d.wall_s = (float)(sqrt((double)p.wall_mu_s) * (double)p.radius);
    d.wall_k = (float)(sqrt((double)p.wall_mu_k) * (double)p.radius);
    d.jitter = (float)((double)p.wall_jitter * (double)p.radius);
    d.num_cells = p.grid[0] * p.grid[1] * p.grid[2];
}

void fill_kconst(const ps_params &p, const ps_derived &d, KConst &k)
{
    memset(&k, 0, sizeof(k));
    k.h = p.h;
    k.h2 = d.h2;
    k.poly6 = d.poly6;
    k.spiky = d.spiky;
    k.w_zero = d.w_zero;
    k.w_dq = d.w_dq;
    k.inv_w_dq = d.inv_w_dq;
    k.w_scale = d.w_scale;
    k.w_scale_h2 = d.w_scale_h2;
    k.search_r2 = d.search_r2;
    k.collide_r2 = d.collide_r2;
    k.collide_dist = p.collide_dist;
    k.wall_s = d.wall_s;
    k.wall_k = d.wall_k;
    k.jitter = d.jitter;
    k.radius = p.radius;
    k.relax = p.relaxation;
    k.k_corr = d.tensile_k;
    k.xsph_c = p.xsph_c;
    k.s_solid = p.s_solid;
    k.mu_s = p.mu_s;
    k.mu_k = p.mu_k;
    for (int a = 0; a < 3; a++) {
        k.gravity[a] = p.gravity[a];
        k.box_min[a] = p.box_min[a];
        k.box_max[a] = p.box_max[a];
        k.origin[a] = p.origin[a];
        k.grid[a] = p.grid[a];
    }
    k.cell = p.cell;
    k.n_corr = p.n_corr;
    k.dims = p.dims;
    k.seed = p.seed;
}

// Constants of the CPU-reference mode, in the order the reference evaluates
// them (totalfluidconstraint.cpp:107,123-137) in FP64, rounded once.
void fill_crconst(const ps_params &p, const ps_derived &d, ps::CrConst &c)
{
    memset(&c, 0, sizeof(c));
    const double pi = 3.14159265358979323846;
    const double h = p.h, h2 = h * h, h6 = h2 * h2 * h2, h9 = h6 * h2 * h;
    const double poly6 = 315.0 / (64.0 * pi * h9);
    const double dq = p.dq_corr, t = h2 - (((dq * dq) * h) * h);
    c.h = p.h;
    c.h2 = (float)h2;
    c.poly6 = (float)poly6;
    c.spiky = (float)(45.0 / (pi * h6));
    c.w_dq = (float)(poly6 * ((t * t) * t));
    // GasConstraint: poly6(DQ_P * DQ_P * H * H), DQ_P = .25 (gasconstraint.h:16, .cpp:94)
    const double tg = h2 - (((0.25 * 0.25) * h) * h);
    c.g_wdq = (float)(poly6 * ((tg * tg) * tg));
    c.relax = p.relaxation;
    c.k_corr = p.k_corr;
    c.jitter = d.jitter;
    c.radius = p.radius;
    c.sf = sqrtf(p.wall_mu_s);
    c.kf = sqrtf(p.wall_mu_k);
    // PARTICLE_DIAM - EPSILON (simulation.cpp:182) with the runtime radius
    c.touch = (float)(2.0 * (double)p.radius - 1e-4);
    for (int a = 0; a < 3; a++) {
        c.gravity[a] = p.gravity[a];
        c.box_min[a] = p.box_min[a];
        c.box_max[a] = p.box_max[a];
    }
    c.n_corr = p.n_corr;
    c.dims = p.dims;
    c.seed = p.seed;
}

int validate(const ps_params *p, std::string &why)
{
    if (!p) {
        why = "params is NULL";
        return PS_ERR_INVALID_ARGUMENT;
    }
    if (p->abi_version != PS_ABI_VERSION) {
        why = "abi_version mismatch";
        return PS_ERR_INVALID_ARGUMENT;
    }
    if (p->dims != 2 && p->dims != 3) {
        why = "dims must be 2 or 3";
        return PS_ERR_INVALID_ARGUMENT;
    }
    const float pos[] = {p->radius, p->h, p->cell, p->collide_dist};
    for (float v : pos)
        if (!(v > 0.0f) || !is_finite(v)) {
            why = "radius, h, cell and collide_dist must be positive and finite";
            return PS_ERR_INVALID_ARGUMENT;
        }
    if (p->cell < p->h || p->cell < p->collide_dist) {
        why = "cell must be >= max(h, collide_dist) (one-cell neighbour stencil)";
        return PS_ERR_INVALID_ARGUMENT;
    }
    for (int a = 0; a < 3; a++) {
        if (!is_finite(p->gravity[a]) || !is_finite(p->box_min[a]) || !is_finite(p->box_max[a]) ||
            !is_finite(p->origin[a])) {
            why = "gravity, box and origin must be finite";
            return PS_ERR_INVALID_ARGUMENT;
        }
        if (p->grid[a] == 0 || p->grid[a] > (1u << 20)) {
            why = "grid dimensions must be in [1, 2^20]";
            return PS_ERR_INVALID_ARGUMENT;
        }
    }
    if (p->dims == 2 && p->grid[2] != 1) {
        why = "2D systems need grid[2] == 1";
        return PS_ERR_INVALID_ARGUMENT;
    }
    const uint64_t cells = (uint64_t)p->grid[0] * p->grid[1] * p->grid[2];
    if (cells >= (1ull << 31)) {
        why = "too many grid cells";
        return PS_ERR_INVALID_ARGUMENT;
    }
    if (p->iterations == 0 || p->iterations > 1000) {
        why = "iterations must be in [1, 1000]";
        return PS_ERR_INVALID_ARGUMENT;
    }
    if (p->max_particles == 0 || p->max_particles > (1u << 30)) {
        why = "max_particles must be in [1, 2^30]";
        return PS_ERR_INVALID_ARGUMENT;
    }
    if (p->max_neighbors == 0 || p->max_neighbors > 4096 || (p->max_neighbors & 3u)) {
        why = "max_neighbors must be a multiple of 4 in [4, 4096]";
        return PS_ERR_INVALID_ARGUMENT;
    }
    const float nonneg[] = {p->relaxation, p->k_corr, p->dq_corr, p->xsph_c, p->s_solid,
                            p->mu_s, p->mu_k, p->wall_mu_s, p->wall_mu_k, p->wall_jitter,
                            p->shape_stiffness};
    for (float v : nonneg)
        if (!(v >= 0.0f) || !is_finite(v)) {
            why = "solver coefficients must be finite and non-negative";
            return PS_ERR_INVALID_ARGUMENT;
        }
    if (!(p->dq_corr < 1.0f)) {
        why = "dq_corr must be < 1";
        return PS_ERR_INVALID_ARGUMENT;
    }
    for (int a = 0; a < 3; a++)
        if (!(p->box_max[a] > p->box_min[a])) {
            why = "box_max must exceed box_min";
            return PS_ERR_INVALID_ARGUMENT;
        }
    if ((p->flags & PS_FLAG_STABILIZATION) && !(p->flags & PS_FLAG_CPU_REFERENCE)) {
        why = "PS_FLAG_STABILIZATION needs PS_FLAG_CPU_REFERENCE";
        return PS_ERR_INVALID_ARGUMENT;
    }
    return PS_OK;
}

template <typename T>
int dev_alloc(ps_system *s, T **ptr, size_t count)
{
    void *p = nullptr;
    HIP_TRY(s, hipMalloc(&p, sizeof(T) * (count ? count : 1)));
    *ptr = (T *)p;
    return PS_OK;
}

template <typename T>
void dev_free(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

void adj_free(ps_system *s)
{
    ps_system::AdjWork &w = s->adj;
    dev_free(w.ckX); dev_free(w.ckV); dev_free(w.ckA); dev_free(w.ckS); dev_free(w.ckSt);
    dev_free(w.tN); dev_free(w.tC); dev_free(w.tS); dev_free(w.tX); dev_free(w.tP); dev_free(w.tPit);
    dev_free(w.tW); dev_free(w.tA); dev_free(w.tClamp); dev_free(w.tLQ); dev_free(w.tF); dev_free(w.tRho);
    dev_free(w.pb); dev_free(w.Y); dev_free(w.Y2); dev_free(w.M); dev_free(w.AB); dev_free(w.AB2);
    dev_free(w.xsb); dev_free(w.gx); dev_free(w.gv); dev_free(w.MC); dev_free(w.KV); dev_free(w.nc); dev_free(w.r0b);
    dev_free(w.wkb); dev_free(w.drho); dev_free(w.dwk); dev_free(w.sb);
    w.cap = w.cks = w.tslots = w.iters = w.ndist = w.tab = 0;
}

void free_all(ps_system *s)
{
    DevBuffers &b = s->b;
    dev_free(b.X); dev_free(b.X2); dev_free(b.P); dev_free(b.Q); dev_free(b.V); dev_free(b.V2);
    dev_free(b.A); dev_free(b.A2); dev_free(b.key); dev_free(b.key2); dev_free(b.rank);
    dev_free(b.tmp_orig); dev_free(b.slot_of); dev_free(b.mj); dev_free(b.phs); dev_free(b.cell_count); dev_free(b.cell_off);
    dev_free(b.scan_part); dev_free(b.nbr); dev_free(b.ncnt); dev_free(b.dist_off);
    dev_free(b.dist_adj); dev_free(b.dist_pairs); dev_free(b.dist_rs); dev_free(b.occ);
    dev_free(b.pin_of); dev_free(b.pin_xyz); dev_free(b.body_off); dev_free(b.body_idx);
    dev_free(b.body_rest); dev_free(b.body_stiff); dev_free(b.body_mass); dev_free(b.body_q);
    dev_free(b.body_goal); dev_free(b.body_of); dev_free(b.member_of); dev_free(b.status);
    dev_free(b.sid); dev_free(b.sid2); dev_free(b.scenes); dev_free(b.lam); dev_free(b.bc);
    dev_free(b.force); dev_free(b.tmass); dev_free(b.ttimer); dev_free(b.cnum); dev_free(b.clist);
    dev_free(b.sflag);
    dev_free(b.stage);
    for (ps_system::Emitter &e : s->emitters) dev_free(e.mark);
    adj_free(s);
}

void drop_graph(ps_system *s)
{
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
}

// ---- timing -------------------------------------------------------------

int timed_begin(ps_system *s, int id, TimedLaunch &tl)
{
    tl.id = id;
    tl.start = tl.stop = nullptr;
    tl.by_dispatch = false;
    if (!(s->timing & (1u << id))) return PS_OK;
    if (s->seen[id]++ % s->timing_period) return PS_OK;
    for (hipEvent_t *e : {&tl.start, &tl.stop}) {
        if (!s->free_events.empty()) {
            *e = s->free_events.back();
            s->free_events.pop_back();
        } else {
            HIP_TRY(s, hipEventCreate(e));
        }
    }
    // single-kernel launchers take the pair on their dispatch; the scan's
    // two kernels are bracketed on the stream
    tl.by_dispatch = id != PS_K_SCAN;
    if (tl.by_dispatch)
        ps::g_launch_events = ps::LaunchEvents{tl.start, tl.stop};
    else
        HIP_TRY(s, hipEventRecord(tl.start, s->tstream ? s->tstream : s->stream));
    return PS_OK;
}

int timed_end(ps_system *s, TimedLaunch &tl)
{
    if (!tl.start) return PS_OK;
    if (tl.by_dispatch && ps::g_launch_events.start) {
        // no kernel took the pair (nothing to launch): no measurement
        ps::g_launch_events = ps::LaunchEvents{};
        s->free_events.push_back(tl.start);
        s->free_events.push_back(tl.stop);
        return PS_OK;
    }
    if (!tl.by_dispatch) HIP_TRY(s, hipEventRecord(tl.stop, s->tstream ? s->tstream : s->stream));
    s->pending.push_back(tl);
    return PS_OK;
}

int timing_collect(ps_system *s)
{
    if (s->pending.empty()) return PS_OK;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    if (s->rstream) HIP_TRY(s, hipStreamSynchronize(s->rstream));
    for (TimedLaunch &tl : s->pending) {
        float ms = 0.0f;
        HIP_TRY(s, hipEventElapsedTime(&ms, tl.start, tl.stop));
        s->total_ms[tl.id] += ms;
        s->launches[tl.id] += 1;
        s->free_events.push_back(tl.start);
        s->free_events.push_back(tl.stop);
    }
    s->pending.clear();
    return PS_OK;
}

#define LAUNCH(sys, id, expr)                                                              \
    do {                                                                                   \
        TimedLaunch tl_;                                                                   \
        int rc_ = timed_begin((sys), (id), tl_);                                           \
        if (rc_) return rc_;                                                               \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((sys), PS_ERR_DEVICE, std::string("launch ") + ps_kernel_name(id) + \
                                                  ": " + hipGetErrorString(e_));          \
        rc_ = timed_end((sys), tl_);                                                       \
        if (rc_) return rc_;                                                               \
    } while (0)

// ---- stages -------------------------------------------------------------

uint32_t features(const ps_system *s)
{
    uint32_t f = ps::F_FLUID;
    if (s->any_contact_phase) f |= ps::F_CONTACT;
    if (!s->dist_pairs.empty()) f |= ps::F_DIST;
    if (s->body_off.size() > 1) f |= ps::F_SHAPE;
    if (!s->pin_xyz.empty()) f |= ps::F_PIN;
    return f;
}

bool uniform_fluid(const ps_system *s) { return s->uniform && features(s) == ps::F_FLUID; }

int stage_hash(ps_system *s, bool predict, float dt)
{
    LAUNCH(s, PS_K_PREDICT_HASH, ps::launch_predict_hash(s->stream, s->b, s->kc, s->n, dt, predict, true,
                                                         s->block[PS_K_PREDICT_HASH]));
    s->hashed = true;
    return PS_OK;
}

// scan + scatter + reorder on keys already computed; swaps the buffers
int stage_sort(ps_system *s)
{
    LAUNCH(s, PS_K_SCAN, ps::launch_scan(s->stream, s->b, s->ncell));
    LAUNCH(s, PS_K_SCATTER, ps::launch_scatter(s->stream, s->b, s->n));
    // the neighbour masses for the fluid sums of scenes that are not a uniform fluid
    LAUNCH(s, PS_K_REORDER, ps::launch_reorder(s->stream, s->b, s->n, s->block[PS_K_REORDER],
                                               uniform_fluid(s) ? -1.0f : s->kc.s_solid));
    DevBuffers &b = s->b;
    std::swap(b.X, b.X2);
    std::swap(b.P, b.Q);
    std::swap(b.V, b.V2);
    std::swap(b.A, b.A2);
    std::swap(b.key, b.key2);
    s->hashed = false;
    s->nbr_valid = false;
    return PS_OK;
}

// Neighbour cells can be staged in LDS (in a batch, the units that lie in
// one scene), unless the caller turned it off by flag; the autotuner may
// turn it off too.  (Staged XSPH takes the fluid mass as a constant: with
// fluid masses that differ, finalize gathers from global memory through
// the units' chunk tables.)
bool can_stage(const ps_system *s) { return !s->cpuref && !(s->prm.flags & PS_FLAG_NO_LDS_STAGING); }

bool staging_ok(const ps_system *s) { return can_stage(s) && s->stage_choice != 0; }

int stage_neighbors(ps_system *s)
{
    s->b.staging = staging_ok(s) ? 1u : 0u;
    LAUNCH(s, PS_K_NEIGHBORS,
           ps::launch_neighbors(s->stream, s->b, s->kc, s->n, s->kmax, s->block[PS_K_NEIGHBORS]));
    s->nbr_valid = true;
    return PS_OK;
}

int stage_iterate(ps_system *s)
{
    int rc;
    if (s->prm.flags & PS_FLAG_REBUILD_EVERY_ITERATION) {
        if ((rc = stage_hash(s, false, 0.0f))) return rc;
        if ((rc = stage_sort(s))) return rc;
        if ((rc = stage_neighbors(s))) return rc;
    }
    if (!s->nbr_valid) return fail(s, PS_ERR_STATE, "neighbour table is stale: call ps_find_neighbors");
    const bool uni = uniform_fluid(s);
    LAUNCH(s, PS_K_LAMBDA,
           ps::launch_lambda(s->stream, s->b, s->kc, s->n, uni, s->wu, s->block[PS_K_LAMBDA]));
    const uint32_t nbody = (uint32_t)s->body_off.size() - 1;
    if (nbody) LAUNCH(s, PS_K_SHAPE, ps::launch_shape(s->stream, s->b, nbody));
    LAUNCH(s, PS_K_SOLVE,
           ps::launch_solve(s->stream, s->b, s->kc, s->n, features(s), uni, s->wu, s->it_sub,
                            s->block[PS_K_SOLVE]));
    std::swap(s->b.P, s->b.Q);
    s->iters++;
    s->it_sub++;
    return PS_OK;
}

int stage_end(ps_system *s)
{
    const bool uni = uniform_fluid(s);
    LAUNCH(s, PS_K_FINALIZE,
           ps::launch_finalize(s->stream, s->b, s->kc, s->n, s->dt,
                               (s->prm.flags & PS_FLAG_SLEEPING) != 0, uni, s->wu,
                               s->fluid_mass_uniform ? s->wf : -1.0f, s->it_sub,
                               s->block[PS_K_FINALIZE]));
    s->it_sub = 0;
    std::swap(s->b.X, s->b.P);
    s->substeps++;
    s->stage = 0;
    return PS_OK;
}

// ---- CPU-reference mode stages --------------------------------------------

// Immovable solids (or a FluidEmitter that makes them) switch on the
// contact pass: contacts are found after prediction and each iteration
// starts with them and the boundaries.
bool cr_contacts(const ps_system *s) { return s->any_solid_phase || !s->femitters.empty(); }

int cr_begin(ps_system *s, float dt)
{
    s->dt = dt;
    s->it_sub = 0;
    if (s->cr_swapped) {
        std::swap(s->b.P, s->b.Q);
        s->cr_swapped = false;
    }
    const bool contacts = cr_contacts(s);
    const bool stab = (s->prm.flags & PS_FLAG_STABILIZATION) != 0;
    LAUNCH(s, PS_K_PREDICT_HASH,
           ps::launch_cr_predict(s->stream, s->b, s->cr, s->n, dt, !contacts && !stab));
    if (contacts) LAUNCH(s, PS_K_NEIGHBORS, ps::launch_cr_find_contacts(s->stream, s->b, s->cr, s->n));
    if (stab) LAUNCH(s, PS_K_SOLVE, ps::launch_cr_stabilize(s->stream, s->b, s->cr, s->n));
    // with neither, prediction projects the boundaries of iteration 0
    s->cr_bnd_done = !contacts && !stab;
    s->stage = 1;
    return PS_OK;
}

int cr_iterate(ps_system *s)
{
    const uint32_t it = s->it_sub;
    const bool contacts = cr_contacts(s);
    if (contacts) {
        // the CONTACT group: contacts, then boundaries, per particle
        LAUNCH(s, PS_K_SOLVE, ps::launch_cr_contacts(s->stream, s->b, s->cr, s->n, it));
        std::swap(s->b.P, s->b.Q);
        s->cr_swapped = !s->cr_swapped;
    } else if (!s->cr_bnd_done) {
        LAUNCH(s, PS_K_SOLVE, ps::launch_cr_boundary(s->stream, s->b, s->cr, s->n, it));
    }
    // without contacts, the last object pass also projects the next
    // iteration's boundaries
    const int next = !contacts && it + 1 < s->prm.iterations ? (int)(it + 1) : -1;
    const size_t nobj = s->objects.size();
    for (size_t k = 0; k < nobj; k++) {
        const ps_system::FluidObject &o = s->objects[k];
        LAUNCH(s, PS_K_LAMBDA,
               ps::launch_cr_lambda(s->stream, s->b, s->cr, s->n, o.first, o.count, o.rho0, o.gas));
        LAUNCH(s, PS_K_SOLVE,
               ps::launch_cr_delta(s->stream, s->b, s->cr, s->n, o.first, o.count, o.rho0, o.gas,
                                   k + 1 == nobj ? next : -1));
        std::swap(s->b.P, s->b.Q);
        s->cr_swapped = !s->cr_swapped;
    }
    s->cr_bnd_done = nobj > 0 && next >= 0;
    s->iters++;
    s->it_sub++;
    return PS_OK;
}

// Emissions of every emitter over `substeps` substeps of dt from the current
// timers (OpenSmokeEmitter::tick, opensmokeemitter.cpp:18-29: the timer
// gains dt in double and pays 1/rate per emission); adds the particles and
// the markers of each emitter to what is already there and reports the
// first emitter to overflow.
int emission_check(ps_system *s, float dt, uint32_t substeps)
{
    uint64_t n = s->n;
    for (size_t k = 0; k < s->emitters.size(); k++) {
        const ps_system::Emitter &e = s->emitters[k];
        double timer = e.timer;
        uint64_t count = 0;
        for (uint32_t t = 0; t < substeps; t++) {
            timer += (double)dt;
            while (timer >= 1.0 / e.rate) {
                timer -= 1.0 / e.rate;
                count++;
            }
            if (e.nmark + count > e.cap)
                return fail(s, PS_ERR_CAPACITY,
                            "smoke emitter " + std::to_string(k) + " exceeds its marker capacity");
        }
        if (e.object >= 0) n += count;
    }
    for (const ps_system::FluidEmitter &e : s->femitters) {
        double timer = e.timer, total = e.total;
        for (uint32_t t = 0; t < substeps; t++) {
            timer += (double)dt;
            total += (double)dt;
            while (total < 5.0 && timer >= 1.0 / e.rate) {
                timer -= 1.0 / e.rate;
                n++;
            }
        }
    }
    if (n > s->cap) return fail(s, PS_ERR_CAPACITY, "emitters exceed max_particles");
    return PS_OK;
}

// OpenSmokeEmitter::tick and the markers' box clamp (simulation.cpp:348-362)
// of every emitter, in creation order, after the velocity update: the
// emissions of this substep, then the markers' advection through all
// particles (those just emitted included).
int cr_emitters(ps_system *s)
{
    for (ps_system::Emitter &e : s->emitters) {
        uint32_t count = 0;
        e.timer += (double)s->dt;
        while (e.timer >= 1.0 / e.rate) {
            e.timer -= 1.0 / e.rate;
            count++;
        }
        ps_system::FluidObject *o = e.object >= 0 ? &s->objects[e.object] : nullptr;
        for (uint32_t done = 0; done < count;) {
            const uint32_t c = std::min(count - done, 64u);
            LAUNCH(s, PS_K_FINALIZE,
                   ps::launch_cr_emit(s->stream, s->b, e.mark, e.nmark, o ? s->n : ~0u, c, e.pos,
                                      o ? o->rho0 : 0.0f));
            e.nmark += c;
            if (o) {
                o->count += c;
                s->n += c;
                s->scene_count[0] += c;
            }
            done += c;
        }
        LAUNCH(s, PS_K_FINALIZE,
               ps::launch_cr_markers(s->stream, s->b, e.mark, s->cr, e.nmark, s->n, s->dt));
    }
    return PS_OK;
}

// FluidEmitter::tick (fluidemitter.cpp:13-79) of every fluid emitter in
// creation order: solidification over its object, then its emissions (a
// double timer and the 5 s total, counted on the host).
int cr_fluid_emitters(ps_system *s)
{
    for (size_t k = 0; k < s->femitters.size(); k++) {
        ps_system::FluidEmitter &e = s->femitters[k];
        ps_system::FluidObject &o = s->objects[e.object];
        LAUNCH(s, PS_K_FINALIZE, ps::launch_cr_solidify(s->stream, s->b, o.first, o.count, s->dt));
        uint32_t count = 0;
        e.timer += (double)s->dt;
        e.total += (double)s->dt;
        while (e.total < 5.0 && e.timer >= 1.0 / e.rate) {
            e.timer -= 1.0 / e.rate;
            count++;
        }
        for (uint32_t done = 0; done < count;) {
            const uint32_t c = std::min(count - done, 64u);
            LAUNCH(s, PS_K_FINALIZE,
                   ps::launch_cr_emit_fluid(s->stream, s->b, s->cr, s->n, c, e.pos, o.rho0,
                                            (uint32_t)k, e.emitted));
            e.emitted += c;
            o.count += c;
            s->n += c;
            s->scene_count[0] += c;
            done += c;
        }
    }
    return PS_OK;
}

int cr_end(ps_system *s)
{
    if (!s->emitters.empty() || !s->femitters.empty())
        if (int rc = emission_check(s, s->dt, 1)) return rc;
    LAUNCH(s, PS_K_FINALIZE,
           ps::launch_cr_end(s->stream, s->b, s->n, s->dt, (s->prm.flags & PS_FLAG_SLEEPING) != 0,
                             s->it_sub));
    // the iterate is rebuilt by the next prediction: restore the buffer order
    if (s->cr_swapped) {
        std::swap(s->b.P, s->b.Q);
        s->cr_swapped = false;
    }
    if (int rc = cr_emitters(s)) return rc;
    if (int rc = cr_fluid_emitters(s)) return rc;
    s->it_sub = 0;
    s->substeps++;
    s->stage = 0;
    return PS_OK;
}

// mode dispatch of the substep stages
int any_iterate(ps_system *s) { return s->cpuref ? cr_iterate(s) : stage_iterate(s); }
int any_end(ps_system *s) { return s->cpuref ? cr_end(s) : stage_end(s); }

int unsupported_in_cpuref(ps_system *s, const char *what)
{
    return fail(s, PS_ERR_UNSUPPORTED,
                std::string(what) + " is not available with PS_FLAG_CPU_REFERENCE");
}

int check_sys(ps_system *s)
{
    if (!s) return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "system is NULL");
    return PS_OK;
}

// Upload the whole constraint incidence structure (insertion indices).
int upload_constraints(ps_system *s)
{
    DevBuffers &b = s->b;
    const size_t m = s->dist_rest.size();
    if (m > s->dist_cap) {
        size_t nc = m * 2;
        dev_free(b.dist_adj);
        dev_free(b.dist_pairs);
        dev_free(b.dist_rs);
        s->dist_cap = 0;
        int rc;
        if ((rc = dev_alloc(s, &b.dist_adj, 2 * nc))) return rc;
        if ((rc = dev_alloc(s, &b.dist_pairs, nc))) return rc;
        if ((rc = dev_alloc(s, &b.dist_rs, nc))) return rc;
        s->dist_cap = nc;
    }
    std::vector<uint32_t> off(s->cap + 1, 0), adj(2 * m), fill(s->cap, 0);
    for (size_t c = 0; c < m; c++) {
        off[s->dist_pairs[2 * c] + 1]++;
        off[s->dist_pairs[2 * c + 1] + 1]++;
    }
    for (uint32_t i = 0; i < s->cap; i++) off[i + 1] += off[i];
    for (size_t c = 0; c < m; c++) {
        uint32_t a = s->dist_pairs[2 * c], bb = s->dist_pairs[2 * c + 1];
        adj[off[a] + fill[a]++] = (uint32_t)c;
        adj[off[bb] + fill[bb]++] = (uint32_t)c;
    }
    std::vector<float> rs(2 * m);
    for (size_t c = 0; c < m; c++) {
        rs[2 * c] = s->dist_rest[c];
        rs[2 * c + 1] = s->dist_stiff[c];
    }
    HIP_TRY(s, hipMemcpy(b.dist_off, off.data(), sizeof(uint32_t) * off.size(), hipMemcpyHostToDevice));
    if (m) {
        HIP_TRY(s, hipMemcpy(b.dist_adj, adj.data(), sizeof(uint32_t) * adj.size(), hipMemcpyHostToDevice));
        HIP_TRY(s, hipMemcpy(b.dist_pairs, s->dist_pairs.data(), sizeof(uint32_t) * 2 * m,
                             hipMemcpyHostToDevice));
        HIP_TRY(s, hipMemcpy(b.dist_rs, rs.data(), sizeof(float) * 2 * m, hipMemcpyHostToDevice));
    }
    HIP_TRY(s, hipMemcpy(b.occ, s->occ.data(), sizeof(uint32_t) * s->cap, hipMemcpyHostToDevice));
    return PS_OK;
}

int upload_pins(ps_system *s)
{
    DevBuffers &b = s->b;
    const size_t np = s->pin_xyz.size() / 3;
    if (np > s->pin_cap) {
        dev_free(b.pin_xyz);
        s->pin_cap = 0;
        int rc = dev_alloc(s, &b.pin_xyz, np * 2);
        if (rc) return rc;
        s->pin_cap = np * 2;
    }
    std::vector<float> p4(4 * np);
    for (size_t k = 0; k < np; k++) {
        p4[4 * k] = s->pin_xyz[3 * k];
        p4[4 * k + 1] = s->pin_xyz[3 * k + 1];
        p4[4 * k + 2] = s->pin_xyz[3 * k + 2];
        p4[4 * k + 3] = 0.0f;
    }
    if (np)
        HIP_TRY(s, hipMemcpy(b.pin_xyz, p4.data(), sizeof(float) * 4 * np, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.pin_of, s->pin_of.data(), sizeof(uint32_t) * s->cap, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.occ, s->occ.data(), sizeof(uint32_t) * s->cap, hipMemcpyHostToDevice));
    return PS_OK;
}

// Start keeping slot_of (insertion index -> storage slot) current: rebuilt
// from the attributes once, then maintained by every reorder.
int enable_slot_tracking(ps_system *s)
{
    if (s->b.track_slots) return PS_OK;
    std::vector<Attr> A(s->n);
    std::vector<uint32_t> slot(s->cap, 0);
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    if (s->n) {
        HIP_TRY(s, hipMemcpy(A.data(), s->b.A, sizeof(Attr) * s->n, hipMemcpyDeviceToHost));
        for (uint32_t t = 0; t < s->n; t++) slot[A[t].orig] = t;
        HIP_TRY(s, hipMemcpy(s->b.slot_of, slot.data(), sizeof(uint32_t) * s->n, hipMemcpyHostToDevice));
    }
    s->b.track_slots = 1;
    drop_graph(s);
    return PS_OK;
}

// Copy the sorted positions and insertion indices to the host.
int read_sorted_positions(ps_system *s, std::vector<float4> &X, std::vector<Attr> &A)
{
    X.resize(s->n);
    A.resize(s->n);
    if (s->n == 0) return PS_OK;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    HIP_TRY(s, hipMemcpy(X.data(), s->b.X, sizeof(float4) * s->n, hipMemcpyDeviceToHost));
    HIP_TRY(s, hipMemcpy(A.data(), s->b.A, sizeof(Attr) * s->n, hipMemcpyDeviceToHost));
    return PS_OK;
}

}  // namespace

// ===========================================================================

namespace ps {
int fail_stateless(int code, const std::string &msg) { return fail(nullptr, code, msg); }
}  // namespace ps

extern "C" {

int ps_params_default(ps_params *p)
{
    if (!p) return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "params is NULL");
    memset(p, 0, sizeof(*p));
    p->abi_version = PS_ABI_VERSION;
    p->dims = 3;
    // gpu/src/particleapp.cpp:65-80: radius 0.25, 64^3 grid, 15000 particles,
    // box (-50,0,-50)..(50,200,50), 5 iterations; cell = 2r
    // (particlesystem.cpp:62).  H = 2 (integration_kernel.cuh:25) exceeds the
    // one-cell stencil at cell 0.5, so the default grid uses cell = h = 2.
    p->radius = 0.25f;
    p->h = 2.0f;
    p->cell = 2.0f;
    p->collide_dist = 0.25f * 2.001f;   // integration_kernel.cuh:319
    p->gravity[0] = 0.0f;
    p->gravity[1] = -9.8f;
    p->gravity[2] = 0.0f;
    p->box_min[0] = -50.0f; p->box_min[1] = 0.0f; p->box_min[2] = -50.0f;
    p->box_max[0] = 50.0f; p->box_max[1] = 200.0f; p->box_max[2] = 50.0f;
    p->origin[0] = -50.0f; p->origin[1] = 0.0f; p->origin[2] = -50.0f;
    p->grid[0] = 50; p->grid[1] = 100; p->grid[2] = 50;
    p->iterations = 5;
    p->relaxation = 0.01f;  // FLUID_RELAXATION :32
    p->k_corr = 0.1f;       // K_P :33
    p->n_corr = 4;          // E_P :34
    p->dq_corr = 0.2f;      // DQ_P :35
    p->xsph_c = 0.0f;       // the reference has no XSPH
    p->s_solid = 1.0f;      // solids count fully in density (:577)
    p->mu_s = 0.005f;       // S_FRICTION :39
    p->mu_k = 0.0002f;      // K_FRICTION :40
    p->wall_mu_s = 0.005f;
    p->wall_mu_k = 0.0002f;
    p->wall_jitter = 0.01f; // eps = d * 0.01 (:83,91)
    p->shape_stiffness = 1.0f;
    p->max_particles = 15000;
    p->max_neighbors = 128;
    p->seed = 1234;         // integration.cu:47
    p->flags = PS_FLAG_UNSCALED_TENSILE;  // lambdaCorr = -K_P (W/W(dq H))^E_P (:628-633)
    return PS_OK;
}

int ps_params_validate(const ps_params *p)
{
    std::string why;
    int rc = validate(p, why);
    if (rc) return fail(nullptr, rc, why);
    return PS_OK;
}

int ps_derive(const ps_params *p, ps_derived *out)
{
    std::string why;
    int rc = validate(p, why);
    if (rc) return fail(nullptr, rc, why);
    if (!out) return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "out is NULL");
    derive_constants(*p, *out);
    return PS_OK;
}

static int create_impl(const ps_params *params, uint32_t nscenes, const ps_scene_desc *descs,
                       ps_system **out)
{
    std::string why;
    int rc = validate(params, why);
    if (rc) return fail(nullptr, rc, why);
    if (!out) return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (nscenes == 0 || (nscenes > 1 && !descs))
        return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "a batch needs scene descriptions");
    const bool cpuref = (params->flags & PS_FLAG_CPU_REFERENCE) != 0;
    if (cpuref && (nscenes > 1 || descs))
        return fail(nullptr, PS_ERR_UNSUPPORTED, "the CPU-reference mode runs one scene per handle");
    // scene table: one scene from the parameters, or one per description
    std::vector<ps::SceneDev> scenes(nscenes);
    uint64_t cells = 0;
    for (uint32_t q = 0; q < nscenes; q++) {
        ps::SceneDev &sc = scenes[q];
        memset(&sc, 0, sizeof(sc));
        sc.cell = params->cell;
        sc.cell_base = (uint32_t)cells;
        if (nscenes == 1 && !descs) {
            for (int a = 0; a < 3; a++) {
                sc.box_min[a] = params->box_min[a];
                sc.box_max[a] = params->box_max[a];
                sc.origin[a] = params->origin[a];
                sc.gravity[a] = params->gravity[a];
                sc.grid[a] = params->grid[a];
            }
            sc.seed = params->seed;
        } else {
            const ps_scene_desc &d = descs[q];
            if (!is_finite(d.gravity_scale))
                return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "gravity_scale must be finite");
            for (int a = 0; a < 3; a++) {
                if (!is_finite(d.box_min[a]) || !is_finite(d.box_max[a]) || !(d.box_max[a] > d.box_min[a]))
                    return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "scene box must be finite and non-empty");
                sc.box_min[a] = d.box_min[a];
                sc.box_max[a] = d.box_max[a];
                sc.origin[a] = d.box_min[a];
                sc.gravity[a] = params->gravity[a] * d.gravity_scale;
                const double g = ceil(((double)d.box_max[a] - (double)d.box_min[a]) / (double)params->cell - 1e-9);
                sc.grid[a] = (params->dims == 2 && a == 2) ? 1u : (uint32_t)(g < 1.0 ? 1.0 : g);
                if (sc.grid[a] > (1u << 20))
                    return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "scene grid too large");
            }
            sc.seed = d.seed;
        }
        cells += (uint64_t)sc.grid[0] * sc.grid[1] * sc.grid[2];
        if (cells >= (1ull << 31)) return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "too many grid cells");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, PS_ERR_NO_DEVICE, "no HIP device available");
    ps_system *s = new ps_system();
    s->prm = *params;
    derive_constants(s->prm, s->der);
    s->der.num_cells = (uint32_t)cells;
    fill_kconst(s->prm, s->der, s->kc);
    s->kc.nscenes = nscenes;
    s->cpuref = cpuref;
    fill_crconst(s->prm, s->der, s->cr);
    s->scenes = scenes;
    s->scene_count.assign(nscenes, 0);
    s->cap = params->max_particles;
    s->ncell = (uint32_t)cells;
    s->kmax = params->max_neighbors;
    s->tiles = ps::scan_tiles(s->ncell);
    s->occ.assign(s->cap, 0);
    s->pin_of.assign(s->cap, 0);
    s->body_of.assign(s->cap, 0);
    s->member_of.assign(s->cap, 0);
    const size_t c = s->cap;
    DevBuffers &b = s->b;
    b.kq = s->kmax / 4;
    const size_t groups = (c + 63) / 64;
#define ALLOC(ptr, cnt)                                 \
    do {                                                \
        int r_ = dev_alloc(s, &(ptr), (cnt));           \
        if (r_) {                                       \
            free_all(s);                                \
            delete s;                                   \
            return r_;                                  \
        }                                               \
    } while (0)
    ALLOC(b.X, c); ALLOC(b.X2, c); ALLOC(b.P, c); ALLOC(b.Q, c); ALLOC(b.V, c); ALLOC(b.V2, c);
    ALLOC(b.A, c); ALLOC(b.A2, c); ALLOC(b.key, c); ALLOC(b.key2, c); ALLOC(b.rank, c);
    ALLOC(b.tmp_orig, c); ALLOC(b.slot_of, c); ALLOC(b.mj, c); ALLOC(b.phs, c);
    ALLOC(b.cell_count, s->ncell); ALLOC(b.cell_off, (size_t)s->ncell + 1);
    ALLOC(b.scan_part, s->tiles);
    ALLOC(b.nbr, groups * 64 * s->kmax); ALLOC(b.ncnt, c);
    ALLOC(b.dist_off, c + 1); ALLOC(b.occ, c); ALLOC(b.pin_of, c);
    ALLOC(b.body_of, c); ALLOC(b.member_of, c); ALLOC(b.status, 1);
    ALLOC(b.stage, (c + ps::kStageUnit - 1) / ps::kStageUnit);
    ALLOC(b.scenes, nscenes);
    if (nscenes > 1) {
        ALLOC(b.sid, c);
        ALLOC(b.sid2, c);
    }
    if (cpuref) {
        ALLOC(b.lam, c);
        ALLOC(b.bc, c);
        ALLOC(b.force, c);
        ALLOC(b.tmass, c);
        ALLOC(b.ttimer, c);
        ALLOC(b.cnum, c);
        ALLOC(b.clist, c * ps::kCrMaxContacts);
        ALLOC(b.sflag, c);
    }
#undef ALLOC
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemset(b.cell_count, 0, sizeof(uint32_t) * s->ncell);
    if (e == hipSuccess) e = hipMemset(b.cell_off, 0, sizeof(uint32_t) * ((size_t)s->ncell + 1));
    if (e == hipSuccess) e = hipMemset(b.ncnt, 0, sizeof(uint32_t) * c);
    if (e == hipSuccess) e = hipMemset(b.dist_off, 0, sizeof(uint32_t) * (c + 1));
    if (e == hipSuccess) e = hipMemset(b.occ, 0, sizeof(uint32_t) * c);
    if (e == hipSuccess) e = hipMemset(b.pin_of, 0, sizeof(uint32_t) * c);
    if (e == hipSuccess) e = hipMemset(b.body_of, 0, sizeof(uint32_t) * c);
    if (e == hipSuccess) e = hipMemset(b.member_of, 0, sizeof(uint32_t) * c);
    if (e == hipSuccess) e = hipMemset(b.status, 0, sizeof(ps::DevStatus));
    if (e == hipSuccess)
        e = hipMemset(b.stage, 0, sizeof(ps::StageDesc) * ((c + ps::kStageUnit - 1) / ps::kStageUnit));
    if (e == hipSuccess)
        e = hipMemcpy(b.scenes, scenes.data(), sizeof(ps::SceneDev) * nscenes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        std::string msg = std::string("device setup: ") + hipGetErrorString(e);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        free_all(s);
        delete s;
        return fail(nullptr, PS_ERR_DEVICE, msg);
    }
    *out = s;
    return PS_OK;
}

int ps_create(const ps_params *params, ps_system **out)
{
    return create_impl(params, 1, nullptr, out);
}

int ps_create_batch(const ps_params *params, uint32_t nscenes, const ps_scene_desc *scenes,
                    ps_system **out)
{
    if (nscenes == 0 || !scenes)
        return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "a batch needs at least one scene");
    return create_impl(params, nscenes, scenes, out);
}

uint32_t ps_num_scenes(const ps_system *s) { return s ? (uint32_t)s->scenes.size() : 0; }

int ps_destroy(ps_system *s)
{
    if (!s) return PS_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (TimedLaunch &tl : s->pending) {
        (void)hipEventDestroy(tl.start);
        (void)hipEventDestroy(tl.stop);
    }
    for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
    drop_graph(s);
    free_all(s);
    if (s->rstream) (void)hipStreamDestroy(s->rstream);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return PS_OK;
}

int ps_set_params(ps_system *s, const ps_params *p)
{
    if (int rc = check_sys(s)) return rc;
    std::string why;
    int rc = validate(p, why);
    if (rc) return fail(s, rc, why);
    if (p->max_particles != s->prm.max_particles || p->max_neighbors != s->prm.max_neighbors ||
        p->dims != s->prm.dims || memcmp(p->grid, s->prm.grid, sizeof(p->grid)) != 0)
        return fail(s, PS_ERR_INVALID_ARGUMENT,
                    "capacities, grid and dims are fixed at ps_create");
    if (s->scenes.size() > 1 && p->cell != s->prm.cell)
        return fail(s, PS_ERR_INVALID_ARGUMENT, "the cell size of a batch is fixed at creation");
    if (((p->flags ^ s->prm.flags) & PS_FLAG_CPU_REFERENCE) != 0)
        return fail(s, PS_ERR_INVALID_ARGUMENT, "PS_FLAG_CPU_REFERENCE is fixed at ps_create");
    if (s->stage) return fail(s, PS_ERR_STATE, "cannot change parameters inside a substep");
    const uint32_t ncells = s->der.num_cells, nsc = s->kc.nscenes;
    s->prm = *p;
    derive_constants(s->prm, s->der);
    s->der.num_cells = ncells;
    fill_kconst(s->prm, s->der, s->kc);
    s->kc.nscenes = nsc;
    fill_crconst(s->prm, s->der, s->cr);
    if (nsc == 1) {
        // the single scene follows the parameters' box, gravity and seed
        ps::SceneDev &sc = s->scenes[0];
        for (int a = 0; a < 3; a++) {
            sc.box_min[a] = p->box_min[a];
            sc.box_max[a] = p->box_max[a];
            sc.origin[a] = p->origin[a];
            sc.gravity[a] = p->gravity[a];
        }
        sc.seed = p->seed;
    }
    drop_graph(s);
    return PS_OK;
}

int ps_get_params(const ps_system *s, ps_params *out)
{
    if (!s || !out) return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = s->prm;
    return PS_OK;
}

const char *ps_last_error(const ps_system *s)
{
    if (s && !s->err.empty()) return s->err.c_str();
    return g_err.c_str();
}

static int add_particles_impl(ps_system *s, uint32_t scene, uint32_t n, const float *pos,
                              const float *vel, const float *inv_mass, const int32_t *phase,
                              const float *rest_density, uint32_t *first_out);

// An emitter feeding an object appends to the end of the particle range, so
// that object stays the last one.
static int objects_open(ps_system *s)
{
    for (const ps_system::Emitter &e : s->emitters)
        if (e.object >= 0)
            return fail(s, PS_ERR_STATE, "nothing can be added after a smoke emitter that feeds an object");
    if (!s->femitters.empty())
        return fail(s, PS_ERR_STATE, "nothing can be added after a fluid emitter");
    return PS_OK;
}

int ps_add_particles(ps_system *s, uint32_t n, const float *pos, const float *vel,
                     const float *inv_mass, const int32_t *phase, const float *rest_density,
                     uint32_t *first_out)
{
    return ps_add_scene_particles(s, 0, n, pos, vel, inv_mass, phase, rest_density, first_out);
}

int ps_add_scene_particles(ps_system *s, uint32_t scene, uint32_t n, const float *pos,
                           const float *vel, const float *inv_mass, const int32_t *phase,
                           const float *rest_density, uint32_t *first_out)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) {
        // immovable solids only (new Particle(pos, 0), simulation.cpp:1182-1185)
        for (uint32_t k = 0; k < n; k++)
            if (!phase || !inv_mass || phase[k] != PS_PHASE_SOLID || inv_mass[k] != 0.0f)
                return unsupported_in_cpuref(
                    s, "ps_add_particles other than immovable solids (use ps_add_fluid_object)");
        if (int rc = objects_open(s)) return rc;
    }
    return add_particles_impl(s, scene, n, pos, vel, inv_mass, phase, rest_density, first_out);
}

int ps_add_fluid_object(ps_system *s, uint32_t n, const float *pos, const float *vel,
                        const float *inv_mass, float rest_density, uint32_t *first_out)
{
    if (int rc = check_sys(s)) return rc;
    if (int rc = objects_open(s)) return rc;
    if (!(rest_density > 0.0f) || !is_finite(rest_density))
        return fail(s, PS_ERR_INVALID_ARGUMENT, "fluid rest density must be > 0");
    std::vector<int32_t> phase(n, PS_PHASE_FLUID);
    std::vector<float> rho(n, rest_density);
    uint32_t first = 0;
    int rc = add_particles_impl(s, 0, n, pos, vel, inv_mass, phase.data(), rho.data(), &first);
    if (rc) return rc;
    if (s->cpuref && n) s->objects.push_back({first, n, rest_density, 0});
    if (first_out) *first_out = first;
    return PS_OK;
}

int ps_add_gas_object(ps_system *s, uint32_t n, const float *pos, const float *vel,
                      const float *inv_mass, float rest_density, int open, uint32_t *first_out)
{
    if (int rc = check_sys(s)) return rc;
    if (int rc = objects_open(s)) return rc;
    if (!s->cpuref || s->prm.dims != 2)
        return fail(s, PS_ERR_UNSUPPORTED, "gas objects belong to the 2D CPU-reference mode");
    if (!(rest_density > 0.0f) || !is_finite(rest_density))
        return fail(s, PS_ERR_INVALID_ARGUMENT, "gas rest density must be > 0");
    for (uint32_t k = 0; k < n && inv_mass; k++)
        if (inv_mass[k] == 0.0f)  // simulation.cpp:418-421
            return fail(s, PS_ERR_INVALID_ARGUMENT, "a gas particle cannot have infinite mass");
    std::vector<int32_t> phase(n, PS_PHASE_GAS);
    std::vector<float> rho(n, rest_density);
    uint32_t first = 0;
    int rc = add_particles_impl(s, 0, n, pos, vel, inv_mass, phase.data(), rho.data(), &first);
    if (rc) return rc;
    if (n) s->objects.push_back({first, n, rest_density, open ? 2 : 1});
    if (first_out) *first_out = first;
    return PS_OK;
}

int ps_add_smoke_emitter(ps_system *s, const float *pos, double particles_per_sec, int32_t object,
                         uint32_t max_markers, uint32_t *emitter_out)
{
    if (int rc = check_sys(s)) return rc;
    if (!s->cpuref || s->prm.dims != 2)
        return fail(s, PS_ERR_UNSUPPORTED, "smoke emitters belong to the 2D CPU-reference mode");
    if (s->stage) return fail(s, PS_ERR_STATE, "cannot add an emitter inside a substep");
    if (!pos || !is_finite(pos[0]) || !is_finite(pos[1]) || !is_finite(pos[2]))
        return fail(s, PS_ERR_INVALID_ARGUMENT, "emitter position must be finite");
    if (!(particles_per_sec > 0.0) || !isfinite(particles_per_sec))
        return fail(s, PS_ERR_INVALID_ARGUMENT, "particles_per_sec must be > 0");
    if (object >= (int32_t)s->objects.size() || object < -1)
        return fail(s, PS_ERR_INVALID_ARGUMENT, "object out of range");
    if (object >= 0) {
        const ps_system::FluidObject &o = s->objects[object];
        if (!o.gas) return fail(s, PS_ERR_INVALID_ARGUMENT, "a smoke emitter feeds a gas object");
        if (object + 1 != (int32_t)s->objects.size() || o.first + o.count != s->n)
            return fail(s, PS_ERR_UNSUPPORTED, "a smoke emitter feeds the last object only");
        if (!s->femitters.empty())
            return fail(s, PS_ERR_UNSUPPORTED, "a fluid emitter feeds a different object");
    }
    ps_system::Emitter e{};
    e.timer = 0.0;
    e.rate = particles_per_sec;
    for (int a = 0; a < 3; a++) e.pos[a] = pos[a];
    e.object = object;
    e.cap = max_markers;
    HIP_TRY(s, hipMalloc((void **)&e.mark, sizeof(float4) * (max_markers ? max_markers : 1)));
    drop_graph(s);
    s->emitters.push_back(e);
    if (emitter_out) *emitter_out = (uint32_t)s->emitters.size() - 1;
    return PS_OK;
}

int ps_add_fluid_emitter(ps_system *s, const float *pos, double particles_per_sec, int32_t object,
                         uint32_t *emitter_out)
{
    if (int rc = check_sys(s)) return rc;
    if (!s->cpuref || s->prm.dims != 2)
        return fail(s, PS_ERR_UNSUPPORTED, "fluid emitters belong to the 2D CPU-reference mode");
    if (s->stage) return fail(s, PS_ERR_STATE, "cannot add an emitter inside a substep");
    if (!pos || !is_finite(pos[0]) || !is_finite(pos[1]) || !is_finite(pos[2]))
        return fail(s, PS_ERR_INVALID_ARGUMENT, "emitter position must be finite");
    if (!(particles_per_sec > 0.0) || !isfinite(particles_per_sec))
        return fail(s, PS_ERR_INVALID_ARGUMENT, "particles_per_sec must be > 0");
    if (object < 0 || object >= (int32_t)s->objects.size())
        return fail(s, PS_ERR_INVALID_ARGUMENT, "object out of range");
    const ps_system::FluidObject &o = s->objects[object];
    if (o.gas) return fail(s, PS_ERR_INVALID_ARGUMENT, "a fluid emitter feeds a fluid object");
    if (object + 1 != (int32_t)s->objects.size() || o.first + o.count != s->n)
        return fail(s, PS_ERR_UNSUPPORTED, "a fluid emitter feeds the last object only");
    for (const ps_system::Emitter &e : s->emitters)
        if (e.object >= 0 && e.object != object)
            return fail(s, PS_ERR_UNSUPPORTED, "another emitter feeds a different object");
    ps_system::FluidEmitter e{};
    e.timer = e.total = 0.0;
    e.rate = particles_per_sec;
    for (int a = 0; a < 3; a++) e.pos[a] = pos[a];
    e.object = object;
    e.emitted = 0;
    drop_graph(s);
    s->femitters.push_back(e);
    if (emitter_out) *emitter_out = (uint32_t)s->femitters.size() - 1;
    return PS_OK;
}

int ps_read_markers(ps_system *s, uint32_t emitter, uint32_t *count, float *xyz)
{
    if (int rc = check_sys(s)) return rc;
    if (emitter >= s->emitters.size()) return fail(s, PS_ERR_INVALID_ARGUMENT, "emitter out of range");
    const ps_system::Emitter &e = s->emitters[emitter];
    if (count) *count = e.nmark;
    if (!xyz || !e.nmark) return PS_OK;
    std::vector<float4> m(e.nmark);
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    HIP_TRY(s, hipMemcpy(m.data(), e.mark, sizeof(float4) * e.nmark, hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < e.nmark; k++) {
        xyz[3 * k] = m[k].x;
        xyz[3 * k + 1] = m[k].y;
        xyz[3 * k + 2] = m[k].z;
    }
    return PS_OK;
}

static int add_particles_impl(ps_system *s, uint32_t scene, uint32_t n, const float *pos,
                              const float *vel, const float *inv_mass, const int32_t *phase,
                              const float *rest_density, uint32_t *first_out)
{
    if (s->stage) return fail(s, PS_ERR_STATE, "cannot add particles inside a substep");
    if (scene >= s->scenes.size()) return fail(s, PS_ERR_INVALID_ARGUMENT, "scene out of range");
    if (n && s->n && scene < s->last_scene)
        return fail(s, PS_ERR_INVALID_ARGUMENT,
                    "fill scenes in increasing order (a scene is one contiguous index range)");
    if (n == 0) {
        if (first_out) *first_out = s->n;
        return PS_OK;
    }
    if (!pos || !inv_mass || !phase || !rest_density)
        return fail(s, PS_ERR_INVALID_ARGUMENT, "NULL input array");
    if ((uint64_t)s->n + n > s->cap)
        return fail(s, PS_ERR_CAPACITY, "max_particles exceeded");
    std::vector<float4> X(n), V(n);
    std::vector<Attr> A(n);
    std::vector<uint32_t> slot(n);
    bool uni = s->uniform;
    float wu = s->n ? s->wu : inv_mass[0];
    bool contact = s->any_contact_phase, solid = s->any_solid_phase;
    uint64_t nwalk = s->n_walk;
    bool fmu = s->fluid_mass_uniform, anyf = s->any_fluid;
    float wf = s->wf;
    for (uint32_t k = 0; k < n; k++) {
        for (int a = 0; a < 3; a++) {
            if (!is_finite(pos[3 * k + a]) || (vel && !is_finite(vel[3 * k + a])))
                return fail(s, PS_ERR_NOT_FINITE, "non-finite position or velocity");
        }
        if (!(inv_mass[k] >= 0.0f) || !is_finite(inv_mass[k]))
            return fail(s, PS_ERR_INVALID_ARGUMENT, "inverse mass must be finite and >= 0");
        if (phase[k] < 0) return fail(s, PS_ERR_INVALID_ARGUMENT, "phase must be >= 0");
        if (phase[k] == PS_PHASE_FLUID && inv_mass[k] == 0.0f)
            // cpu/src/simulation.cpp:436-439 rejects fluids of infinite mass
            return fail(s, PS_ERR_INVALID_ARGUMENT, "a fluid particle cannot have infinite mass");
        if (phase[k] == PS_PHASE_FLUID && !(rest_density[k] > 0.0f && is_finite(rest_density[k])))
            return fail(s, PS_ERR_INVALID_ARGUMENT, "fluid rest density must be > 0");
        X[k] = make_float4(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2], 0.0f);
        V[k] = vel ? make_float4(vel[3 * k], vel[3 * k + 1], vel[3 * k + 2], 0.0f)
                   : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        A[k].w = inv_mass[k];
        A[k].phase = phase[k];
        A[k].rho0 = rest_density[k];
        A[k].orig = s->n + k;
        slot[k] = s->n + k;
        if (phase[k] != PS_PHASE_FLUID || inv_mass[k] != wu) uni = false;
        if (phase[k] == PS_PHASE_FLUID) {
            if (!anyf) wf = inv_mass[k];
            else if (inv_mass[k] != wf) fmu = false;
            anyf = true;
        }
        if (phase[k] >= PS_PHASE_CLOTH) contact = true;
        if (phase[k] >= PS_PHASE_CLOTH && inv_mass[k] != 0.0f) nwalk++;
        if (phase[k] >= PS_PHASE_SOLID) solid = true;
    }
    const size_t o = s->n;
    DevBuffers &b = s->b;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    HIP_TRY(s, hipMemcpy(b.X + o, X.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    // the iterate starts equal to the positions (lambda 0)
    HIP_TRY(s, hipMemcpy(b.P + o, X.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.V + o, V.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.A + o, A.data(), sizeof(Attr) * n, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.slot_of + o, slot.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    if (b.force) HIP_TRY(s, hipMemset(b.force + o, 0, sizeof(float4) * n));
    if (b.ttimer) {
        const std::vector<float> t4(n, 4.0f);  // Particle::init, particle.h:37-38
        HIP_TRY(s, hipMemcpy(b.ttimer + o, t4.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    }
    if (s->scene_count[scene] == 0) {
        s->scenes[scene].first = s->n;
        HIP_TRY(s, hipMemcpy(b.scenes + scene, &s->scenes[scene], sizeof(ps::SceneDev),
                             hipMemcpyHostToDevice));
    }
    if (b.sid) {
        std::vector<uint32_t> sid(n, scene);
        HIP_TRY(s, hipMemcpy(b.sid + o, sid.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    }
    s->scene_count[scene] += n;
    s->last_scene = scene;
    drop_graph(s);
    if (first_out) *first_out = s->n;
    s->n += n;
    s->uniform = uni;
    s->wu = wu;
    s->fluid_mass_uniform = fmu;
    s->any_fluid = anyf;
    s->wf = wf;
    s->any_contact_phase = contact;
    s->b.contacts = contact ? 1u : 0u;
    // the solve runs units without contact-walking particles through the
    // variant without contacts when such particles are rare (at most 1 in
    // 32: ropes in fluid); with more of them (cloth, rigid bodies) most
    // units hold one and a second launch over the others would not pay.
    // Only the launch plan depends on this count: the neighbour build
    // classifies the units each substep from the particles' phases and
    // masses as they are then, so a count made stale by a later change of
    // w or phase costs time, never a different result.
    s->n_walk = nwalk;
    s->b.split_solve = (contact && nwalk * 32u <= (uint64_t)s->n) ? 1u : 0u;
    s->any_solid_phase = solid;
    s->nbr_valid = false;
    return PS_OK;
}

int ps_add_distance_constraints(ps_system *s, uint32_t m, const uint32_t *pairs,
                                const float *rest, const float *stiffness)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) return unsupported_in_cpuref(s, "ps_add_distance_constraints");
    if (s->stage) return fail(s, PS_ERR_STATE, "cannot add constraints inside a substep");
    if (m == 0) return PS_OK;
    if (!pairs || !rest) return fail(s, PS_ERR_INVALID_ARGUMENT, "NULL input array");
    for (uint32_t k = 0; k < m; k++) {
        if (pairs[2 * k] >= s->n || pairs[2 * k + 1] >= s->n || pairs[2 * k] == pairs[2 * k + 1])
            return fail(s, PS_ERR_INVALID_ARGUMENT, "constraint index out of range or degenerate");
        if (!(rest[k] >= 0.0f) || !is_finite(rest[k]))
            return fail(s, PS_ERR_INVALID_ARGUMENT, "rest length must be finite and >= 0");
        if (stiffness && !(stiffness[k] > 0.0f && stiffness[k] <= 1.0f))
            return fail(s, PS_ERR_INVALID_ARGUMENT, "stiffness must be in (0, 1]");
    }
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    drop_graph(s);
    if (int rc = enable_slot_tracking(s)) return rc;
    const size_t m0 = s->dist_rest.size();
    for (uint32_t k = 0; k < m; k++) {
        s->dist_pairs.push_back(pairs[2 * k]);
        s->dist_pairs.push_back(pairs[2 * k + 1]);
        s->dist_rest.push_back(rest[k]);
        s->dist_stiff.push_back(stiffness ? stiffness[k] : 1.0f);
        s->occ[pairs[2 * k]]++;
        s->occ[pairs[2 * k + 1]]++;
    }
    const int rc = upload_constraints(s);
    if (rc) {
        // keep the host sets equal to what the device last received: undo
        // this call's constraints and re-upload the earlier ones (best effort;
        // the first error is the one reported)
        for (size_t c = m0; c < s->dist_rest.size(); c++) {
            s->occ[s->dist_pairs[2 * c]]--;
            s->occ[s->dist_pairs[2 * c + 1]]--;
        }
        s->dist_pairs.resize(2 * m0);
        s->dist_rest.resize(m0);
        s->dist_stiff.resize(m0);
        const std::string why = s->err;
        (void)upload_constraints(s);
        s->err = why;
    }
    return rc;
}

uint32_t ps_num_distance_constraints(const ps_system *s)
{
    return s ? (uint32_t)s->dist_rest.size() : 0u;
}

int ps_add_point_constraints(ps_system *s, uint32_t m, const uint32_t *idx, const float *xyz)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) return unsupported_in_cpuref(s, "ps_add_point_constraints");
    if (s->stage) return fail(s, PS_ERR_STATE, "cannot add constraints inside a substep");
    if (m == 0) return PS_OK;
    if (!idx || !xyz) return fail(s, PS_ERR_INVALID_ARGUMENT, "NULL input array");
    for (uint32_t k = 0; k < m; k++) {
        if (idx[k] >= s->n) return fail(s, PS_ERR_INVALID_ARGUMENT, "pin index out of range");
        for (int a = 0; a < 3; a++)
            if (!is_finite(xyz[3 * k + a])) return fail(s, PS_ERR_NOT_FINITE, "non-finite pin");
    }
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    drop_graph(s);
    const size_t np0 = s->pin_xyz.size();
    std::vector<std::pair<uint32_t, uint32_t>> prev;  // (particle, its pin before this call)
    for (uint32_t k = 0; k < m; k++) {
        const uint32_t pid = (uint32_t)(s->pin_xyz.size() / 3);
        s->pin_xyz.push_back(xyz[3 * k]);
        s->pin_xyz.push_back(xyz[3 * k + 1]);
        s->pin_xyz.push_back(xyz[3 * k + 2]);
        prev.emplace_back(idx[k], s->pin_of[idx[k]]);
        s->pin_of[idx[k]] = pid + 1;
        s->occ[idx[k]]++;
    }
    const int rc = upload_pins(s);
    if (rc) {
        // undo this call's pins (in reverse, so a particle pinned twice gets
        // its first value back) and re-upload the earlier ones (best effort)
        for (size_t k = prev.size(); k-- > 0;) {
            s->pin_of[prev[k].first] = prev[k].second;
            s->occ[prev[k].first]--;
        }
        s->pin_xyz.resize(np0);
        const std::string why = s->err;
        (void)upload_pins(s);
        s->err = why;
    }
    return rc;
}

int ps_add_rigid_bodies(ps_system *s, uint32_t nbody, const uint32_t *offsets,
                        const uint32_t *idx, const float *stiffness, uint32_t *first_body_out)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) return unsupported_in_cpuref(s, "ps_add_rigid_bodies");
    if (s->stage) return fail(s, PS_ERR_STATE, "cannot add bodies inside a substep");
    if (nbody == 0) {
        if (first_body_out) *first_body_out = (uint32_t)s->body_off.size() - 1;
        return PS_OK;
    }
    if (!offsets || !idx) return fail(s, PS_ERR_INVALID_ARGUMENT, "NULL input array");
    if (offsets[0] != 0) return fail(s, PS_ERR_INVALID_ARGUMENT, "offsets[0] must be 0");
    for (uint32_t bb = 0; bb < nbody; bb++) {
        if (offsets[bb + 1] < offsets[bb] + 2)
            // cpu/src/simulation.cpp:373-376: rigid bodies need at least 2 points
            return fail(s, PS_ERR_INVALID_ARGUMENT, "a rigid body needs at least 2 particles");
        if (stiffness && (!(stiffness[bb] <= 1.0f) || !is_finite(stiffness[bb])))
            return fail(s, PS_ERR_INVALID_ARGUMENT, "stiffness must be <= 1");
    }
    const uint32_t total = offsets[nbody];
    std::vector<float4> X;
    std::vector<Attr> A;
    if (int rc = read_sorted_positions(s, X, A)) return rc;
    std::vector<uint32_t> slot_of(s->n);
    for (uint32_t t = 0; t < s->n; t++) slot_of[A[t].orig] = t;
    std::vector<uint8_t> seen(s->n, 0);
    for (uint32_t k = 0; k < total; k++) {
        if (idx[k] >= s->n || s->body_of[idx[k]] || seen[idx[k]])
            return fail(s, PS_ERR_INVALID_ARGUMENT,
                        "body member out of range, repeated, or already in a body");
        seen[idx[k]] = 1;
        if (A[slot_of[idx[k]]].w == 0.0f)
            // cpu/src/simulation.cpp:387-390
            return fail(s, PS_ERR_INVALID_ARGUMENT, "a rigid body cannot have a static particle");
    }
    const uint32_t bfirst = (uint32_t)s->body_off.size() - 1;
    for (uint32_t bb = 0; bb < nbody; bb++) {
        const uint32_t base = s->body_off.back();
        const uint32_t k0 = offsets[bb], k1 = offsets[bb + 1];
        // centre of mass and rest offsets, same float operations as the kernels
        float cx = 0.0f, cy = 0.0f, cz = 0.0f, M = 0.0f;
        for (uint32_t k = k0; k < k1; k++) {
            const uint32_t t = slot_of[idx[k]];
            const float m = 1.0f / A[t].w;
            cx = cx + (m * X[t].x);
            cy = cy + (m * X[t].y);
            cz = cz + (m * X[t].z);
            M = M + m;
        }
        cx = cx / M;
        cy = cy / M;
        cz = cz / M;
        for (uint32_t k = k0; k < k1; k++) {
            const uint32_t t = slot_of[idx[k]];
            s->body_idx.push_back(idx[k]);
            s->body_rest.push_back(X[t].x - cx);
            s->body_rest.push_back(X[t].y - cy);
            s->body_rest.push_back(X[t].z - cz);
            s->body_of[idx[k]] = bfirst + bb + 1;
            s->member_of[idx[k]] = base + (k - k0);
        }
        s->body_off.push_back(base + (k1 - k0));
        const float st = stiffness ? stiffness[bb] : 0.0f;
        s->body_stiff.push_back(st > 0.0f ? st : s->prm.shape_stiffness);
        s->body_mass.push_back(M);
    }
    // upload every body table (bodies are added at scene-build time)
    DevBuffers &b = s->b;
    const size_t nb = s->body_stiff.size(), nm = s->body_idx.size();
    std::vector<float4> q(nb, make_float4(1.f, 0.f, 0.f, 0.f));
    if (bfirst > 0)
        HIP_TRY(s, hipMemcpy(q.data(), b.body_q, sizeof(float4) * bfirst, hipMemcpyDeviceToHost));
    if (nb > s->body_cap || nm > s->member_cap) {
        const size_t bc = nb * 2, mc = nm * 2;
        dev_free(b.body_off); dev_free(b.body_idx); dev_free(b.body_rest); dev_free(b.body_stiff);
        dev_free(b.body_mass); dev_free(b.body_q); dev_free(b.body_goal);
        int rc;
        if ((rc = dev_alloc(s, &b.body_off, bc + 1))) return rc;
        if ((rc = dev_alloc(s, &b.body_idx, mc))) return rc;
        if ((rc = dev_alloc(s, &b.body_rest, mc))) return rc;
        if ((rc = dev_alloc(s, &b.body_stiff, bc))) return rc;
        if ((rc = dev_alloc(s, &b.body_mass, bc))) return rc;
        if ((rc = dev_alloc(s, &b.body_q, bc))) return rc;
        if ((rc = dev_alloc(s, &b.body_goal, mc))) return rc;
        s->body_cap = bc;
        s->member_cap = mc;
    }
    std::vector<float4> rest(nm);
    for (size_t k = 0; k < nm; k++)
        rest[k] = make_float4(s->body_rest[3 * k], s->body_rest[3 * k + 1], s->body_rest[3 * k + 2], 0.f);
    if (int rc = enable_slot_tracking(s)) return rc;
    HIP_TRY(s, hipMemcpy(b.body_q, q.data(), sizeof(float4) * nb, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.body_off, s->body_off.data(), sizeof(uint32_t) * s->body_off.size(), hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.body_idx, s->body_idx.data(), sizeof(uint32_t) * nm, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.body_rest, rest.data(), sizeof(float4) * nm, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.body_stiff, s->body_stiff.data(), sizeof(float) * nb, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.body_mass, s->body_mass.data(), sizeof(float) * nb, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.body_of, s->body_of.data(), sizeof(uint32_t) * s->cap, hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(b.member_of, s->member_of.data(), sizeof(uint32_t) * s->cap, hipMemcpyHostToDevice));
    if (first_body_out) *first_body_out = bfirst;
    drop_graph(s);
    return PS_OK;
}

int ps_add_rigid_body(ps_system *s, uint32_t n, const uint32_t *idx, float stiffness,
                      uint32_t *body_out)
{
    const uint32_t off[2] = {0, n};
    return ps_add_rigid_bodies(s, 1, off, idx, &stiffness, body_out);
}

int ps_set_velocities(ps_system *s, uint32_t first, uint32_t n, const float *vel)
{
    if (int rc = check_sys(s)) return rc;
    if (s->stage) return fail(s, PS_ERR_STATE, "cannot set velocities inside a substep");
    if ((uint64_t)first + n > s->n || (n && !vel))
        return fail(s, PS_ERR_INVALID_ARGUMENT, "range out of bounds");
    for (uint32_t k = 0; k < 3 * n; k++)
        if (!is_finite(vel[k])) return fail(s, PS_ERR_NOT_FINITE, "non-finite velocity");
    std::vector<Attr> A(s->n);
    std::vector<float4> V(s->n);
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    HIP_TRY(s, hipMemcpy(A.data(), s->b.A, sizeof(Attr) * s->n, hipMemcpyDeviceToHost));
    HIP_TRY(s, hipMemcpy(V.data(), s->b.V, sizeof(float4) * s->n, hipMemcpyDeviceToHost));
    for (uint32_t t = 0; t < s->n; t++) {
        const uint32_t o = A[t].orig;
        if (o >= first && o < first + n) {
            const uint32_t k = o - first;
            V[t] = make_float4(vel[3 * k], vel[3 * k + 1], vel[3 * k + 2], 0.0f);
        }
    }
    HIP_TRY(s, hipMemcpy(s->b.V, V.data(), sizeof(float4) * s->n, hipMemcpyHostToDevice));
    return PS_OK;
}

// ---- substep --------------------------------------------------------------

int ps_predict(ps_system *s, float dt)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) return unsupported_in_cpuref(s, "ps_predict");
    if (!is_finite(dt) || dt < 0.0f) return fail(s, PS_ERR_INVALID_ARGUMENT, "dt must be finite and >= 0");
    LAUNCH(s, PS_K_PREDICT_HASH, ps::launch_predict_hash(s->stream, s->b, s->kc, s->n, dt, true,
                                                         false, s->block[PS_K_PREDICT_HASH]));
    s->dt = dt;
    s->stage = 1;
    s->it_sub = 0;
    s->hashed = false;
    s->nbr_valid = false;
    return PS_OK;
}

int ps_build_grid(ps_system *s)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) return unsupported_in_cpuref(s, "ps_build_grid");
    if (s->stage != 1) return fail(s, PS_ERR_STATE, "ps_build_grid before ps_predict");
    int rc;
    if (!s->hashed && (rc = stage_hash(s, false, 0.0f))) return rc;
    return stage_sort(s);
}

int ps_find_neighbors(ps_system *s)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) return unsupported_in_cpuref(s, "ps_find_neighbors");
    if (s->stage != 1) return fail(s, PS_ERR_STATE, "ps_find_neighbors before ps_predict");
    if (s->hashed) return fail(s, PS_ERR_STATE, "ps_find_neighbors before ps_build_grid");
    return stage_neighbors(s);
}

int ps_substep_begin(ps_system *s, float dt)
{
    if (int rc = check_sys(s)) return rc;
    if (!(dt > 0.0f) || !is_finite(dt)) return fail(s, PS_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (s->cpuref) return cr_begin(s, dt);
    int rc;
    s->dt = dt;
    s->it_sub = 0;
    if ((rc = stage_hash(s, true, dt))) return rc;  // fused predict + hash
    s->stage = 1;
    if ((rc = stage_sort(s))) return rc;
    return stage_neighbors(s);
}

int ps_substep_iterate(ps_system *s)
{
    if (int rc = check_sys(s)) return rc;
    if (s->stage != 1) return fail(s, PS_ERR_STATE, "ps_substep_iterate before ps_substep_begin");
    return any_iterate(s);
}

int ps_substep_end(ps_system *s)
{
    if (int rc = check_sys(s)) return rc;
    if (s->stage != 1) return fail(s, PS_ERR_STATE, "ps_substep_end before ps_substep_begin");
    if (!(s->dt > 0.0f)) return fail(s, PS_ERR_STATE, "substep has dt = 0");
    return any_end(s);
}

int ps_step(ps_system *s, float dt, uint32_t substeps)
{
    if (int rc = check_sys(s)) return rc;
    if (s->stage) return fail(s, PS_ERR_STATE, "ps_step inside an open substep");
    if (!(dt > 0.0f) || !is_finite(dt)) return fail(s, PS_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (s->n == 0) return PS_OK;
    // emitters overflowing within the call fail it before the first substep
    if (!s->emitters.empty() || !s->femitters.empty())
        if (int rc = emission_check(s, dt, substeps)) return rc;
    for (uint32_t k = 0; k < substeps; k++) {
        int rc;
        if ((rc = ps_substep_begin(s, dt))) return rc;
        for (uint32_t it = 0; it < s->prm.iterations; it++)
            if ((rc = any_iterate(s))) return rc;
        if ((rc = any_end(s))) return rc;
    }
    return PS_OK;
}

int ps_step_graph(ps_system *s, float dt, uint32_t substeps)
{
    if (int rc = check_sys(s)) return rc;
    if (s->stage) return fail(s, PS_ERR_STATE, "ps_step_graph inside an open substep");
    if (!(dt > 0.0f) || !is_finite(dt)) return fail(s, PS_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (s->n == 0 || substeps == 0) return PS_OK;
    if (!s->emitters.empty() || !s->femitters.empty())
        return fail(s, PS_ERR_UNSUPPORTED, "ps_step_graph with emitters (the particle count grows)");
    if (!s->gexec || s->g_dt != dt || s->g_sub != substeps) {
        drop_graph(s);
        const DevBuffers start = s->b;
        const uint64_t sub0 = s->substeps, it0 = s->iters;
        const uint32_t timing = s->timing;
        s->timing = 0;  // no event records inside the graph
        HIP_TRY(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int rc = PS_OK;
        for (uint32_t k = 0; k < substeps && rc == PS_OK; k++) {
            rc = ps_substep_begin(s, dt);
            for (uint32_t it = 0; it < s->prm.iterations && rc == PS_OK; it++) rc = any_iterate(s);
            if (rc == PS_OK) rc = any_end(s);
        }
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s->stream, &g);
        s->timing = timing;
        s->substeps = sub0;
        s->iters = it0;
        s->stage = 0;
        s->it_sub = 0;
        s->cr_swapped = false;
        // the buffer pointers must come back to where the capture began; the
        // staging mode is whatever the captured neighbour builds chose
        DevBuffers end = s->b, begin = start;
        end.staging = begin.staging = 0;
        const bool rotated_back = memcmp(&begin, &end, sizeof(DevBuffers)) == 0;
        const uint32_t staging = s->b.staging;
        s->b = start;
        s->b.staging = staging;
        if (rc != PS_OK || e != hipSuccess || !rotated_back) {
            if (g) (void)hipGraphDestroy(g);
            if (rc != PS_OK) return rc;
            if (e != hipSuccess)
                return fail(s, PS_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
            return fail(s, PS_ERR_INVALID_ARGUMENT,
                        "substeps must be a multiple of the buffer rotation period "
                        "(4 for an even iteration count, 6 for odd)");
        }
        e = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            s->gexec = nullptr;
            return fail(s, PS_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
        s->g_dt = dt;
        s->g_sub = substeps;
    }
    HIP_TRY(s, hipGraphLaunch(s->gexec, s->stream));
    s->substeps += substeps;
    s->iters += (uint64_t)substeps * s->prm.iterations;
    s->nbr_valid = true;
    return PS_OK;
}

// ---- reverse mode ---------------------------------------------------------

namespace {

int adj_reserve(ps_system *s, size_t cks, size_t tslots, size_t iters)
{
    ps_system::AdjWork &w = s->adj;
    const size_t ndist = s->dist_rest.size();
    const size_t c = s->cap, tab = ((c + 63) / 64) * 64 * (size_t)s->kmax;
    if (w.ckX && w.cap >= c && w.cks >= cks && w.tslots >= tslots && w.iters >= iters && w.ndist >= ndist &&
        w.tab >= tab)
        return PS_OK;
    adj_free(s);
    const size_t ti = tslots * iters;
    int rc;
    if ((rc = dev_alloc(s, &w.ckX, cks * c)) || (rc = dev_alloc(s, &w.ckV, cks * c)) ||
        (rc = dev_alloc(s, &w.ckA, cks * c)) || (rc = dev_alloc(s, &w.ckS, cks * c)) ||
        (rc = dev_alloc(s, &w.ckSt, cks)) || (rc = dev_alloc(s, &w.tN, tslots * tab)) ||
        (rc = dev_alloc(s, &w.tC, tslots * c)) || (rc = dev_alloc(s, &w.tS, tslots * c)) ||
        (rc = dev_alloc(s, &w.tX, tslots * c)) || (rc = dev_alloc(s, &w.tP, tslots * c)) ||
        (rc = dev_alloc(s, &w.tA, tslots * c)) || (rc = dev_alloc(s, &w.tPit, ti * c)) ||
        (rc = dev_alloc(s, &w.tW, ti * c)) || (rc = dev_alloc(s, &w.tClamp, ti * c)) ||
        (rc = dev_alloc(s, &w.tLQ, ti * c)) || (rc = dev_alloc(s, &w.tF, ti * c)) ||
        (rc = dev_alloc(s, &w.tRho, ti * c)) ||
        (rc = dev_alloc(s, &w.pb, c)) || (rc = dev_alloc(s, &w.Y, c)) || (rc = dev_alloc(s, &w.Y2, c)) ||
        (rc = dev_alloc(s, &w.M, c)) || (rc = dev_alloc(s, &w.AB, c)) || (rc = dev_alloc(s, &w.AB2, c)) ||
        (rc = dev_alloc(s, &w.xsb, c)) || (rc = dev_alloc(s, &w.gx, c)) || (rc = dev_alloc(s, &w.gv, c)) ||
        (rc = dev_alloc(s, &w.MC, c)) || (rc = dev_alloc(s, &w.KV, c)) || (rc = dev_alloc(s, &w.nc, c)) || (rc = dev_alloc(s, &w.r0b, c)) ||
        (rc = dev_alloc(s, &w.wkb, c)) || (rc = dev_alloc(s, &w.drho, c)) || (rc = dev_alloc(s, &w.dwk, c)) ||
        (rc = dev_alloc(s, &w.sb, ndist))) {
        adj_free(s);
        return rc == PS_ERR_DEVICE ? fail(s, PS_ERR_OUT_OF_MEMORY, "reverse-pass workspace") : rc;
    }
    w.cap = c;
    w.cks = cks;
    w.tslots = tslots;
    w.iters = iters;
    w.ndist = ndist;
    w.tab = tab;
    return PS_OK;
}

// copy the per-particle state (and the device status word) to / from
// checkpoint k
int adj_copy(ps_system *s, size_t k, bool save)
{
    ps_system::AdjWork &w = s->adj;
    DevBuffers &b = s->b;
    const size_t n = s->n, off = k * w.cap;
    hipError_t e = hipSuccess;
    auto cp = [&](void *keep, void *live, size_t bytes) {
        if (e == hipSuccess)
            e = hipMemcpyAsync(save ? keep : live, save ? live : keep, bytes, hipMemcpyDeviceToDevice,
                               s->stream);
    };
    cp(w.ckX + off, b.X, sizeof(float4) * n);
    cp(w.ckV + off, b.V, sizeof(float4) * n);
    cp(w.ckA + off, b.A, sizeof(Attr) * n);
    cp(w.ckS + off, b.slot_of, sizeof(uint32_t) * n);
    cp(w.ckSt + k, b.status, sizeof(ps::DevStatus));
    if (e != hipSuccess) return fail(s, PS_ERR_DEVICE, std::string("checkpoint copy: ") + hipGetErrorString(e));
    return PS_OK;
}

int adj_substep(ps_system *s, float dt)
{
    int rc;
    if ((rc = ps_substep_begin(s, dt))) return rc;
    for (uint32_t it = 0; it < s->prm.iterations; it++)
        if ((rc = stage_iterate(s))) return rc;
    return stage_end(s);
}

// One substep that also writes tape slot t: the kernels write straight into
// the slot its neighbour table, the sorted start state, attributes and slot
// map (reorder), every iterate with its lambda (the taping lambda pass), the
// clamp bits, wall-friction records and fluid sums (the taping solve) and the
// final iterate (the last solve).
int adj_taped_substep(ps_system *s, float dt, size_t t)
{
    ps_system::AdjWork &w = s->adj;
    DevBuffers &b = s->b;
    uint32_t *own_nbr = b.nbr, *own_cnt = b.ncnt;
    // a staged substep keeps its LDS-index table and writes the global-slot
    // copy the reverse kernels read into the slot
    if (staging_ok(s))
        b.nbr_glob = w.tN + t * w.tab;
    else
        b.nbr = w.tN + t * w.tab;
    b.ncnt = w.tC + t * w.cap;
    // the reorder writes the sorted start state, attributes and slot map
    // into the tape slot as well
    b.tape_x = w.tX + t * w.cap;
    b.tape_a = w.tA + t * w.cap;
    b.tape_s = w.tS + t * w.cap;
    int rc = ps_substep_begin(s, dt);
    b.tape_x = nullptr;
    b.tape_a = nullptr;
    b.tape_s = nullptr;
    const size_t c = w.cap, I = s->prm.iterations;
    hipError_t e = hipSuccess;
    const bool uni = uniform_fluid(s);
    for (uint32_t it = 0; !rc && e == hipSuccess && it < I; it++) {
        const size_t q = (t * I + it) * c;
        // the taping lambda pass also writes the iterate after it to tape
        b.tape_lq = w.tLQ + q;
        b.tape_rho = w.tRho + q;
        b.tape_pit = w.tPit + q;
        e = ps::launch_lambda(s->stream, b, s->kc, s->n, uni, s->wu, s->block[PS_K_LAMBDA]);
        b.tape_lq = nullptr;
        b.tape_rho = nullptr;
        b.tape_pit = nullptr;
        b.clamp_rec = w.tClamp + q;
        b.wall_rec = w.tW + q;
        b.tape_f = w.tF + q;
        // the last solve also writes the substep's final iterate to tape
        b.tape_q = it + 1 == I ? w.tP + t * c : nullptr;
        if (e == hipSuccess)
            e = ps::launch_solve(s->stream, b, s->kc, s->n, features(s), uni, s->wu, s->it_sub,
                                 s->block[PS_K_SOLVE]);
        b.clamp_rec = nullptr;
        b.wall_rec = nullptr;
        b.tape_f = nullptr;
        b.tape_q = nullptr;
        std::swap(b.P, b.Q);
        s->iters++;
        s->it_sub++;
    }
    if (!rc && e == hipSuccess) rc = stage_end(s);
    b.nbr = own_nbr;
    b.ncnt = own_cnt;
    b.nbr_glob = nullptr;
    s->nbr_valid = false;
    if (!rc && e != hipSuccess) rc = fail(s, PS_ERR_DEVICE, std::string("taped substep: ") + hipGetErrorString(e));
    return rc;
}

// Reverse of the substep on tape slot t, enqueued on st: insertion-order
// gx/gv of its end state in, of its start state out; rho0, stiffness and
// wall_k terms accumulate.  It reads the tape slot, the constraint tables
// and its own work arrays, never the live particle state, so a recompute
// may run beside it on another stream.
int adj_reverse(ps_system *s, float dt, size_t t, hipStream_t st)
{
    ps_system::AdjWork &w = s->adj;
    const size_t c = w.cap, I = s->prm.iterations;
    DevBuffers bt = s->b;  // the substep as taped
    bt.nbr = w.tN + t * w.tab;
    bt.ncnt = w.tC + t * c;
    bt.X = w.tX + t * c;
    bt.A = w.tA + t * c;
    bt.slot_of = w.tS + t * c;
    bt.P = w.tP + t * c;
    ps::AdjBuffers a{};
    // dL/dy and the attribute records alternate between two buffers by
    // iteration: iteration it uses buffer it % 2, and the kernel that forms
    // its input (adj_position_k of iteration it + 1, or adj_finalize_k, whose
    // dL/dS is in buffer I % 2) writes them while its own are still read
    float4 *Yb[2] = {w.Y, w.Y2};
    ps::AdjAttr *ABb[2] = {w.AB, w.AB2};
    a.M = w.M;
    a.MC = w.MC;
    a.KV = w.KV;
    a.nc = w.nc;
    a.r0b = w.r0b;
    a.wkb = w.wkb;
    a.xsb = w.xsb;
    a.sb = w.sb;
    a.gx = w.gx;
    a.gv = w.gv;
    a.drho = w.drho;
    a.dwk = w.dwk;
    a.contacts = s->any_contact_phase ? 1u : 0u;
    a.dist = s->dist_rest.empty() ? 0u : 1u;
    a.pb = w.pb;
    const uint32_t n = s->n;
    const bool uni = s->uniform;
    s->tstream = st;  // LAUNCH times these launches on st
    struct Restore {
        ps_system *s;
        ~Restore() { s->tstream = nullptr; }
    } restore{s};
    // the pre-step of iteration it, run at the end of the kernel forming
    // its input (none past the first iteration: that kernel writes pb)
    auto next_pre = [&](size_t it, bool any) {
        if (!any) {
            a.clampn = nullptr;
            return;
        }
        const size_t q = (t * I + it) * c;
        a.clampn = w.tClamp + q;
        a.wrecn = w.tW + q;
        a.tfn = w.tF + q;
        a.Yn = Yb[it % 2];
        a.ABn = ABb[it % 2];
    };
    a.Y = Yb[I % 2];  // dL/dS
    LAUNCH(s, PS_K_ADJ_XSPH, ps::launch_adj(st, ps::ADJ_XSPH, bt, s->kc, a, n, dt, uni));
    next_pre(I ? I - 1 : 0, I > 0);
    LAUNCH(s, PS_K_ADJ_FINALIZE, ps::launch_adj(st, ps::ADJ_FINALIZE, bt, s->kc, a, n, dt, uni));
    for (size_t it = I; it-- > 0;) {
        const size_t q = (t * I + it) * c;
        a.Pit = w.tPit + q;
        a.SD = w.tLQ + q;
        a.rho = w.tRho + q;
        a.Y = Yb[it % 2];
        a.AB = ABb[it % 2];
        LAUNCH(s, PS_K_ADJ_LAMBDA, ps::launch_adj(st, ps::ADJ_LAMBDA, bt, s->kc, a, n, dt, uni));
        next_pre(it ? it - 1 : 0, it > 0);
        LAUNCH(s, PS_K_ADJ_POSITION, ps::launch_adj(st, ps::ADJ_POSITION, bt, s->kc, a, n, dt, uni));
    }
    LAUNCH(s, PS_K_ADJ_PREDICT, ps::launch_adj(st, ps::ADJ_PREDICT, bt, s->kc, a, n, dt, uni));
    return PS_OK;
}

int step_backward_impl(ps_system *s, float dt, uint32_t S, uint32_t K, const float *dl_dx,
                       const float *dl_dv, float *dl_dx0, float *dl_dv0, float *dl_drho0,
                       float *dl_dstiffness, float *dl_dmu_k)
{
    const uint32_t n = s->n;
    // segments of K substeps (K = 0: one segment, the whole horizon).  With
    // more than one segment the tape holds two: segment g is taped into
    // buffer g % 2.  The forward pass tapes the last two segments (one
    // buffer each); the others are recomputed from their checkpoints when
    // the reverse pass reaches them, segment g - 2 (on `stream`) into the
    // buffer the reverse of segment g (on `rstream`) has just finished
    // with, while the reverse of segment g - 1 reads the other
    const uint32_t seg = K ? K : S;
    const uint32_t nseg = (S + seg - 1) / seg;
    const size_t fin = nseg;  // checkpoint of the final state
    const size_t nbuf = nseg > 1 ? 2 : 1;
    int rc;
    if ((rc = adj_reserve(s, nseg + 1, nbuf * seg, s->prm.iterations ? s->prm.iterations : 1))) return rc;
    if (!s->rstream) HIP_TRY(s, hipStreamCreateWithFlags(&s->rstream, hipStreamNonBlocking));
    ps_system::AdjWork &w = s->adj;
    std::vector<uint64_t> sub(nseg + 1), its(nseg + 1);
    auto save = [&](size_t k) {
        sub[k] = s->substeps;
        its[k] = s->iters;
        return adj_copy(s, k, true);
    };
    auto load = [&](size_t k) {
        s->substeps = sub[k];
        s->iters = its[k];
        return adj_copy(s, k, false);
    };
    auto tape_segment = [&](uint32_t g) {
        const uint32_t s0 = g * seg, s1 = std::min(S, s0 + seg);
        for (uint32_t t = s0; t < s1; t++)
            if (int r = adj_taped_substep(s, dt, (g % nbuf) * seg + (t - s0))) return r;
        return (int)PS_OK;
    };
    for (uint32_t g = 0; g < nseg; g++) {
        if ((rc = save(g))) return rc;
        if (g + nbuf < nseg) {
            for (uint32_t t = g * seg; t < std::min(S, (g + 1) * seg); t++)
                if ((rc = adj_substep(s, dt))) return rc;
        } else if ((rc = tape_segment(g))) {
            return rc;
        }
    }
    if ((rc = save(fin))) return rc;
    ps::DevStatus ds;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    HIP_TRY(s, hipMemcpy(&ds, s->b.status, sizeof(ds), hipMemcpyDeviceToHost));
    if (ds.flags & PS_STATUS_NEIGHBOR_OVERFLOW)
        return fail(s, PS_ERR_CAPACITY,
                    "a neighbour table overflowed (raise max_neighbors, or clear the status of an earlier "
                    "overflow): the reverse pass needs every pair");
    // seed: dL/dx_T, dL/dv_T in insertion order
    std::vector<float4> g4(n);
    for (int which = 0; which < 2; which++) {
        const float *src = which ? dl_dv : dl_dx;
        for (uint32_t o = 0; o < n; o++)
            g4[o] = src ? make_float4(src[3 * o], src[3 * o + 1], src[3 * o + 2], 0.0f)
                        : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        HIP_TRY(s, hipMemcpy(which ? w.gv : w.gx, g4.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    }
    HIP_TRY(s, hipMemsetAsync(w.drho, 0, sizeof(float) * n, s->stream));
    HIP_TRY(s, hipMemsetAsync(w.dwk, 0, sizeof(float) * n, s->stream));
    if (!s->dist_rest.empty())
        HIP_TRY(s, hipMemsetAsync(w.sb, 0, sizeof(float) * s->dist_rest.size(), s->stream));
    // tape_done[g]: segment g is on tape (stream); rev_done[g]: reversed (rstream)
    struct Events {
        std::vector<hipEvent_t> e;
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } tape_done, rev_done;
    tape_done.e.assign(nseg, nullptr);
    rev_done.e.assign(nseg, nullptr);
    for (uint32_t g = 0; g < nseg; g++) {
        HIP_TRY(s, hipEventCreateWithFlags(&tape_done.e[g], hipEventDisableTiming));
        HIP_TRY(s, hipEventCreateWithFlags(&rev_done.e[g], hipEventDisableTiming));
    }
    auto recompute = [&](uint32_t g) {
        if (int r = load(g)) return r;
        if (int r = tape_segment(g)) return r;
        HIP_TRY(s, hipEventRecord(tape_done.e[g], s->stream));
        return (int)PS_OK;
    };
    // the forward pass has taped the last nbuf segments
    for (uint32_t g = nseg - (uint32_t)nbuf; g < nseg; g++) HIP_TRY(s, hipEventRecord(tape_done.e[g], s->stream));
    for (uint32_t g = nseg; g-- > 0;) {
        const uint32_t s0 = g * seg, s1 = std::min(S, s0 + seg);
        HIP_TRY(s, hipStreamWaitEvent(s->rstream, tape_done.e[g], 0));
        for (uint32_t t = s1; t-- > s0;)
            if ((rc = adj_reverse(s, dt, (g % nbuf) * seg + (t - s0), s->rstream))) return rc;
        HIP_TRY(s, hipEventRecord(rev_done.e[g], s->rstream));
        if (g >= 2) {
            // segment g - 2 goes into the buffer segment g has just been read from
            HIP_TRY(s, hipStreamWaitEvent(s->stream, rev_done.e[g], 0));
            if ((rc = recompute(g - 2))) return rc;
        }
    }
    HIP_TRY(s, hipStreamWaitEvent(s->stream, rev_done.e[0], 0));
    if ((rc = load(fin))) return rc;
    s->stage = 0;
    s->it_sub = 0;
    s->hashed = false;
    s->nbr_valid = false;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    // results
    for (int which = 0; which < 2; which++) {
        float *dst = which ? dl_dv0 : dl_dx0;
        if (!dst) continue;
        HIP_TRY(s, hipMemcpy(g4.data(), which ? w.gv : w.gx, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (uint32_t o = 0; o < n; o++) {
            dst[3 * o] = g4[o].x;
            dst[3 * o + 1] = g4[o].y;
            dst[3 * o + 2] = g4[o].z;
        }
    }
    if (dl_drho0) HIP_TRY(s, hipMemcpy(dl_drho0, w.drho, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (dl_dstiffness && !s->dist_rest.empty())
        HIP_TRY(s, hipMemcpy(dl_dstiffness, w.sb, sizeof(float) * s->dist_rest.size(), hipMemcpyDeviceToHost));
    if (dl_dmu_k) {
        // wall_k = sqrt(mu_k) radius: dwall_k/dmu_k = radius / (2 sqrt(mu_k)); the
        // per-particle terms are summed in insertion order
        std::vector<float> wk(n);
        HIP_TRY(s, hipMemcpy(wk.data(), w.dwk, sizeof(float) * n, hipMemcpyDeviceToHost));
        float sum = 0.0f;
        for (uint32_t o = 0; o < n; o++) sum = sum + wk[o];
        *dl_dmu_k = sum * (s->prm.radius / (2.0f * sqrtf(s->prm.wall_mu_k)));
    }
    return PS_OK;
}

}  // namespace

int ps_step_backward(ps_system *s, float dt, uint32_t substeps, uint32_t checkpoint_every,
                     const float *dl_dx, const float *dl_dv, float *dl_dx0, float *dl_dv0,
                     float *dl_drho0, float *dl_dstiffness, float *dl_dmu_k)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) return unsupported_in_cpuref(s, "ps_step_backward");
    if (s->stage) return fail(s, PS_ERR_STATE, "ps_step_backward inside an open substep");
    if (!(dt > 0.0f) || !is_finite(dt)) return fail(s, PS_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (s->kc.nscenes > 1) return fail(s, PS_ERR_UNSUPPORTED, "ps_step_backward runs on single-scene handles");
    if (s->body_off.size() > 1 || s->any_solid_phase)
        return fail(s, PS_ERR_UNSUPPORTED, "ps_step_backward covers no solid or rigid particles");
    if (s->prm.flags & (PS_FLAG_SLEEPING | PS_FLAG_REBUILD_EVERY_ITERATION))
        return fail(s, PS_ERR_UNSUPPORTED,
                    "ps_step_backward needs one neighbour table per substep and no sleeping");
    if (dl_dmu_k && !(s->prm.wall_mu_k > 0.0f))
        return fail(s, PS_ERR_INVALID_ARGUMENT,
                    "dl_dmu_k needs wall_mu_k > 0 (d sqrt(mu_k)/d mu_k is unbounded at 0)");
    const uint32_t n = s->n;
    const float *in[2] = {dl_dx, dl_dv};
    for (const float *p : in)
        if (p)
            for (size_t q = 0; q < 3 * (size_t)n; q++)
                if (!is_finite(p[q])) return fail(s, PS_ERR_NOT_FINITE, "non-finite loss gradient");
    if (substeps == 0 || n == 0) {
        for (size_t q = 0; q < 3 * (size_t)n; q++) {
            if (dl_dx0) dl_dx0[q] = dl_dx ? dl_dx[q] : 0.0f;
            if (dl_dv0) dl_dv0[q] = dl_dv ? dl_dv[q] : 0.0f;
        }
        if (dl_drho0) std::fill(dl_drho0, dl_drho0 + n, 0.0f);
        if (dl_dstiffness) std::fill(dl_dstiffness, dl_dstiffness + s->dist_rest.size(), 0.0f);
        if (dl_dmu_k) *dl_dmu_k = 0.0f;
        return PS_OK;
    }
    const uint32_t K = checkpoint_every >= substeps ? 0u : checkpoint_every;
    drop_graph(s);
    s->adj_active = true;
    const int rc = step_backward_impl(s, dt, substeps, K, dl_dx, dl_dv, dl_dx0, dl_dv0, dl_drho0,
                                      dl_dstiffness, dl_dmu_k);
    s->adj_active = false;
    // a failure can leave reverse work queued: let it drain before the
    // workspace is reused
    if (rc && s->rstream) (void)hipStreamSynchronize(s->rstream);
    s->b.clamp_rec = nullptr;
    s->b.wall_rec = nullptr;
    s->b.tape_lq = nullptr;
    s->b.tape_rho = nullptr;
    s->b.tape_f = nullptr;
    if (rc) {
        s->stage = 0;
        s->it_sub = 0;
        s->hashed = false;
        s->nbr_valid = false;
    }
    return rc;
}

int ps_export_device(ps_system *s, uint32_t field, uint32_t first, uint32_t n, void *dst)
{
    if (int rc = check_sys(s)) return rc;
    if (field != PS_FIELD_POSITIONS && field != PS_FIELD_VELOCITIES && field != PS_FIELD_PHASES)
        return fail(s, PS_ERR_INVALID_ARGUMENT, "field must be one PS_FIELD_* value");
    if ((uint64_t)first + n > s->n || (n && !dst))
        return fail(s, PS_ERR_INVALID_ARGUMENT, "range out of bounds");
    LAUNCH(s, PS_K_EXPORT, ps::launch_export(s->stream, s->b, s->n, field, first, n, dst));
    return PS_OK;
}

int ps_synchronize(ps_system *s)
{
    if (int rc = check_sys(s)) return rc;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return PS_OK;
}

// ---- readback -------------------------------------------------------------

uint32_t ps_num_particles(const ps_system *s) { return s ? s->n : 0; }

int ps_read_positions(ps_system *s, uint32_t first, uint32_t n, float *xyz)
{
    if (int rc = check_sys(s)) return rc;
    if ((uint64_t)first + n > s->n || (n && !xyz))
        return fail(s, PS_ERR_INVALID_ARGUMENT, "range out of bounds");
    std::vector<float4> X;
    std::vector<Attr> A;
    if (int rc = read_sorted_positions(s, X, A)) return rc;
    for (uint32_t t = 0; t < s->n; t++) {
        const uint32_t o = A[t].orig;
        if (o >= first && o < first + n) {
            xyz[3 * (o - first)] = X[t].x;
            xyz[3 * (o - first) + 1] = X[t].y;
            xyz[3 * (o - first) + 2] = X[t].z;
        }
    }
    return PS_OK;
}

int ps_read_velocities(ps_system *s, uint32_t first, uint32_t n, float *xyz)
{
    if (int rc = check_sys(s)) return rc;
    if ((uint64_t)first + n > s->n || (n && !xyz))
        return fail(s, PS_ERR_INVALID_ARGUMENT, "range out of bounds");
    if (s->n == 0) return PS_OK;
    std::vector<float4> V(s->n);
    std::vector<Attr> A(s->n);
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    HIP_TRY(s, hipMemcpy(V.data(), s->b.V, sizeof(float4) * s->n, hipMemcpyDeviceToHost));
    HIP_TRY(s, hipMemcpy(A.data(), s->b.A, sizeof(Attr) * s->n, hipMemcpyDeviceToHost));
    for (uint32_t t = 0; t < s->n; t++) {
        const uint32_t o = A[t].orig;
        if (o >= first && o < first + n) {
            xyz[3 * (o - first)] = V[t].x;
            xyz[3 * (o - first) + 1] = V[t].y;
            xyz[3 * (o - first) + 2] = V[t].z;
        }
    }
    return PS_OK;
}

int ps_read_grid(ps_system *s, uint32_t *sorted_index, uint32_t *cell, uint32_t *cell_offsets)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) return unsupported_in_cpuref(s, "ps_read_grid");
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    if (sorted_index && s->n) {
        std::vector<Attr> A(s->n);
        HIP_TRY(s, hipMemcpy(A.data(), s->b.A, sizeof(Attr) * s->n, hipMemcpyDeviceToHost));
        for (uint32_t t = 0; t < s->n; t++) sorted_index[t] = A[t].orig;
    }
    if (cell && s->n)
        HIP_TRY(s, hipMemcpy(cell, s->b.key, sizeof(uint32_t) * s->n, hipMemcpyDeviceToHost));
    if (cell_offsets)
        HIP_TRY(s, hipMemcpy(cell_offsets, s->b.cell_off, sizeof(uint32_t) * ((size_t)s->ncell + 1),
                             hipMemcpyDeviceToHost));
    return PS_OK;
}

int ps_read_neighbors(ps_system *s, uint32_t *counts, uint32_t *table)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) return unsupported_in_cpuref(s, "ps_read_neighbors");
    if (s->n == 0) return PS_OK;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    std::vector<uint32_t> cnt(s->n);
    HIP_TRY(s, hipMemcpy(cnt.data(), s->b.ncnt, sizeof(uint32_t) * s->n, hipMemcpyDeviceToHost));
    if (counts) memcpy(counts, cnt.data(), sizeof(uint32_t) * s->n);
    if (table) {
        const size_t groups = (s->n + 63) / 64;
        const uint32_t kq = s->b.kq;
        std::vector<uint32_t> raw(groups * 64 * s->kmax);
        HIP_TRY(s, hipMemcpy(raw.data(), s->b.nbr, sizeof(uint32_t) * raw.size(), hipMemcpyDeviceToHost));
        // staged units hold LDS indices: map them back through their intervals
        std::vector<ps::StageDesc> desc;
        if (s->b.staging) {
            desc.resize((s->n + ps::kStageUnit - 1) / ps::kStageUnit);
            HIP_TRY(s, hipMemcpy(desc.data(), s->b.stage, sizeof(ps::StageDesc) * desc.size(),
                                 hipMemcpyDeviceToHost));
        }
        for (uint32_t i = 0; i < s->n; i++) {
            const ps::StageDesc *d = desc.empty() ? nullptr : &desc[i / ps::kStageUnit];
            if (d && !d->staged) d = nullptr;
            for (uint32_t k = 0; k < s->kmax; k++) {
                const size_t at = (((size_t)(i >> 6) * kq + (k >> 2)) * 64 + (i & 63)) * 4 + (k & 3);
                uint32_t e = k < cnt[i] ? raw[at] : 0u;
                if (d && k < cnt[i]) {
                    const uint32_t c = e / ps::kStageChunk;
                    e = c < d->nchunks && c < ps::kStageChunks ? d->src[c] + e % ps::kStageChunk
                                                               : 0xffffffffu;
                }
                table[(size_t)i * s->kmax + k] = e;
            }
        }
    }
    return PS_OK;
}

int ps_read_sorted_state(ps_system *s, float *iterate_xyzl, float *start_xyz)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) return unsupported_in_cpuref(s, "ps_read_sorted_state");
    if (s->n == 0) return PS_OK;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    if (iterate_xyzl)
        HIP_TRY(s, hipMemcpy(iterate_xyzl, s->b.P, sizeof(float4) * s->n, hipMemcpyDeviceToHost));
    if (start_xyz) {
        std::vector<float4> X(s->n);
        HIP_TRY(s, hipMemcpy(X.data(), s->b.X, sizeof(float4) * s->n, hipMemcpyDeviceToHost));
        for (uint32_t t = 0; t < s->n; t++) {
            start_xyz[3 * t] = X[t].x;
            start_xyz[3 * t + 1] = X[t].y;
            start_xyz[3 * t + 2] = X[t].z;
        }
    }
    return PS_OK;
}

int ps_get_status(ps_system *s, ps_status *out)
{
    if (int rc = check_sys(s)) return rc;
    if (!out) return fail(s, PS_ERR_INVALID_ARGUMENT, "out is NULL");
    ps::DevStatus ds;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    HIP_TRY(s, hipMemcpy(&ds, s->b.status, sizeof(ds), hipMemcpyDeviceToHost));
    memset(out, 0, sizeof(*out));
    out->flags = ds.flags;
    out->max_neighbors = ds.max_neighbors;
    out->num_particles = s->n;
    out->num_cells = s->ncell;
    out->substeps = s->substeps;
    out->iterations = ds.iter + s->it_sub;
    if (s->b.staging && s->n && !s->cpuref) {
        std::vector<ps::StageDesc> desc((s->n + ps::kStageUnit - 1) / ps::kStageUnit);
        HIP_TRY(s, hipMemcpy(desc.data(), s->b.stage, sizeof(ps::StageDesc) * desc.size(),
                             hipMemcpyDeviceToHost));
        out->units = (uint32_t)desc.size();
        for (const ps::StageDesc &d : desc) out->staged_units += d.staged ? 1u : 0u;
    }
    return PS_OK;
}

int ps_clear_status(ps_system *s)
{
    if (int rc = check_sys(s)) return rc;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    HIP_TRY(s, hipMemset(s->b.status, 0, sizeof(ps::DevStatus)));
    return PS_OK;
}

// ---- timing ---------------------------------------------------------------

int ps_timing_enable(ps_system *s, uint32_t kernel_mask)
{
    if (int rc = check_sys(s)) return rc;
    if (int rc = timing_collect(s)) return rc;
    s->timing = kernel_mask;
    return PS_OK;
}

int ps_timing_reset(ps_system *s)
{
    if (int rc = check_sys(s)) return rc;
    if (int rc = timing_collect(s)) return rc;
    for (int k = 0; k < PS_K_COUNT; k++) {
        s->total_ms[k] = 0.0;
        s->launches[k] = 0;
        s->seen[k] = 0;
    }
    return PS_OK;
}

int ps_timing_sample(ps_system *s, uint32_t period)
{
    if (int rc = check_sys(s)) return rc;
    if (period == 0) return fail(s, PS_ERR_INVALID_ARGUMENT, "period must be >= 1");
    if (int rc = timing_collect(s)) return rc;
    s->timing_period = period;
    for (int k = 0; k < PS_K_COUNT; k++) s->seen[k] = 0;
    return PS_OK;
}

int ps_timing_read(ps_system *s, int kernel_id, double *total_ms, uint64_t *launches)
{
    if (int rc = check_sys(s)) return rc;
    if (kernel_id < 0 || kernel_id >= PS_K_COUNT)
        return fail(s, PS_ERR_INVALID_ARGUMENT, "kernel id out of range");
    if (int rc = timing_collect(s)) return rc;
    if (total_ms) *total_ms = s->total_ms[kernel_id];
    if (launches) *launches = s->launches[kernel_id];
    return PS_OK;
}

// ---- autotuning ------------------------------------------------------------

namespace {

const int kTunable[] = {PS_K_PREDICT_HASH, PS_K_REORDER, PS_K_NEIGHBORS, PS_K_LAMBDA, PS_K_SOLVE,
                        PS_K_FINALIZE};
const uint32_t kCandidates[] = {64, 128, 256, 512};
// the tuning-table entry of the staging choice (value 0 off, 1 on)
constexpr int kTuneStaging = 1000;

// (kernel, log2 N bucket, scene features) -> block size, shared by the
// handles of this process and, with PS_AUTOTUNE_CACHE=<file>, across runs
struct TuneKey {
    int kernel;
    uint32_t bucket, features;
    bool operator<(const TuneKey &o) const
    {
        return kernel != o.kernel ? kernel < o.kernel
                                  : (bucket != o.bucket ? bucket < o.bucket : features < o.features);
    }
};
std::map<TuneKey, uint32_t> g_tune;
bool g_tune_loaded = false;

void tune_cache_load()
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *path = getenv("PS_AUTOTUNE_CACHE");
    if (!path) return;
    FILE *f = fopen(path, "r");
    if (!f) return;
    int k;
    unsigned b, ft, bs;
    while (fscanf(f, "%d %u %u %u", &k, &b, &ft, &bs) == 4)
        if ((bs >= 64 && bs <= 512 && bs % 64 == 0 && k >= 0 && k < PS_K_COUNT) ||
            (k == kTuneStaging && bs <= 1u))
            g_tune[{k, b, ft}] = bs;
    fclose(f);
}

void tune_cache_store()
{
    const char *path = getenv("PS_AUTOTUNE_CACHE");
    if (!path) return;
    FILE *f = fopen(path, "w");
    if (!f) return;
    for (const auto &kv : g_tune)
        fprintf(f, "%d %u %u %u\n", kv.first.kernel, kv.first.bucket, kv.first.features, kv.second);
    fclose(f);
}

uint32_t log2_bucket(uint32_t n)
{
    uint32_t b = 0;
    while ((1u << (b + 1)) <= n && b < 31) b++;
    return b;
}

}  // namespace

int ps_autotune(ps_system *s, float dt, uint32_t substeps, uint32_t *blocks_out)
{
    if (int rc = check_sys(s)) return rc;
    if (s->cpuref) return unsupported_in_cpuref(s, "ps_autotune");
    if (s->stage) return fail(s, PS_ERR_STATE, "ps_autotune inside an open substep");
    if (!(dt > 0.0f) || !is_finite(dt)) return fail(s, PS_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (substeps == 0) substeps = 2;
    if (s->n == 0) return fail(s, PS_ERR_STATE, "autotuning needs particles");
    tune_cache_load();
    const uint32_t bucket = log2_bucket(s->n), feat = features(s) | (s->uniform ? 0x100u : 0u);
    bool need = false;
    for (int k : kTunable) {
        auto it = g_tune.find({k, bucket, feat});
        if (it != g_tune.end())
            s->block[k] = it->second;
        else
            need = true;
    }
    // the second knob: LDS staging of each unit's neighbour cells (unless
    // the caller turned it off by flag)
    const bool stageable = can_stage(s);
    if (stageable) {
        auto it = g_tune.find({kTuneStaging, bucket, feat});
        if (it != g_tune.end())
            s->stage_choice = (int)it->second;
        else
            need = true;
    }
    if (need) {
        // snapshot the state the trials will advance
        const uint32_t n = s->n, nbody = (uint32_t)s->body_off.size() - 1;
        float4 *sx = nullptr, *sv = nullptr, *sq = nullptr;
        Attr *sa = nullptr;
        uint32_t *ss = nullptr, *ssid = nullptr;
        ps::DevStatus *sst = nullptr;
        int rc = PS_OK;
        auto cleanup = [&]() {
            dev_free(sx); dev_free(sv); dev_free(sq); dev_free(sa); dev_free(ss); dev_free(ssid);
            dev_free(sst);
        };
        if ((rc = dev_alloc(s, &sx, n)) || (rc = dev_alloc(s, &sv, n)) || (rc = dev_alloc(s, &sa, n)) ||
            (rc = dev_alloc(s, &ss, n)) || (rc = dev_alloc(s, &sst, 1)) ||
            (nbody && (rc = dev_alloc(s, &sq, nbody))) || (s->b.sid && (rc = dev_alloc(s, &ssid, n)))) {
            cleanup();
            return rc;
        }
        DevBuffers &b = s->b;
        hipError_t e = hipStreamSynchronize(s->stream);
        auto cp = [&](void *dst, const void *src, size_t bytes) {
            if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice);
        };
        cp(sx, b.X, sizeof(float4) * n);
        cp(sv, b.V, sizeof(float4) * n);
        cp(sa, b.A, sizeof(Attr) * n);
        cp(ss, b.slot_of, sizeof(uint32_t) * n);
        cp(sst, b.status, sizeof(ps::DevStatus));
        if (nbody) cp(sq, b.body_q, sizeof(float4) * nbody);
        if (b.sid) cp(ssid, b.sid, sizeof(uint32_t) * n);
        if (e != hipSuccess) {
            cleanup();
            return fail(s, PS_ERR_DEVICE, std::string("autotune snapshot: ") + hipGetErrorString(e));
        }
        const uint64_t sub0 = s->substeps, it0 = s->iters;
        const uint32_t timing0 = s->timing;
        double ms0[PS_K_COUNT];
        uint64_t l0[PS_K_COUNT];
        if ((rc = timing_collect(s))) { cleanup(); return rc; }
        memcpy(ms0, s->total_ms, sizeof(ms0));
        memcpy(l0, s->launches, sizeof(l0));
        for (int k : kTunable) {
            if (g_tune.count({k, bucket, feat})) continue;
            uint32_t best = s->block[k];
            double best_ms = 1e30;
            for (uint32_t bs : kCandidates) {
                s->block[k] = bs;
                s->timing = 0;
                if ((rc = ps_step(s, dt, 1))) break;  // warm-up of this shape
                s->timing = 1u << k;
                s->total_ms[k] = 0.0;
                s->launches[k] = 0;
                if ((rc = ps_step(s, dt, substeps))) break;
                if ((rc = timing_collect(s))) break;
                const double ms = s->launches[k] ? s->total_ms[k] / (double)s->launches[k] : 1e30;
                if (ms < best_ms) {
                    best_ms = ms;
                    best = bs;
                }
            }
            s->block[k] = best;
            if (rc) break;
            g_tune[{k, bucket, feat}] = best;
        }
        if (!rc && stageable && !g_tune.count({kTuneStaging, bucket, feat})) {
            // whole substeps with staging off and on, alternating twice (the
            // scene moves on between trials), timed with events on the stream
            hipEvent_t e0 = nullptr, e1 = nullptr;
            double tot[2] = {0.0, 0.0};
            e = hipEventCreate(&e0);
            if (e == hipSuccess) e = hipEventCreate(&e1);
            for (int round = 0; round < 4 && !rc && e == hipSuccess; round++) {
                const int choice = round & 1;
                s->stage_choice = choice;
                s->timing = 0;
                if ((rc = ps_step(s, dt, 1))) break;  // warm-up of this choice
                e = hipEventRecord(e0, s->stream);
                if (e == hipSuccess && !(rc = ps_step(s, dt, substeps))) {
                    e = hipEventRecord(e1, s->stream);
                    if (e == hipSuccess) e = hipEventSynchronize(e1);
                    float ms = 0.0f;
                    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                    tot[choice] += ms;
                }
            }
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (!rc && e == hipSuccess) {
                s->stage_choice = tot[1] <= tot[0] ? 1 : 0;
                g_tune[{kTuneStaging, bucket, feat}] = (uint32_t)s->stage_choice;
            } else if (!rc) {
                rc = fail(s, PS_ERR_DEVICE, std::string("autotune staging trial: ") + hipGetErrorString(e));
            }
        }
        s->timing = timing0;
        memcpy(s->total_ms, ms0, sizeof(ms0));
        memcpy(s->launches, l0, sizeof(l0));
        // restore the snapshot
        e = hipStreamSynchronize(s->stream);
        cp(b.X, sx, sizeof(float4) * n);
        cp(b.V, sv, sizeof(float4) * n);
        cp(b.A, sa, sizeof(Attr) * n);
        cp(b.slot_of, ss, sizeof(uint32_t) * n);
        cp(b.status, sst, sizeof(ps::DevStatus));
        if (nbody) cp(b.body_q, sq, sizeof(float4) * nbody);
        if (b.sid) cp(b.sid, ssid, sizeof(uint32_t) * n);
        cleanup();
        s->substeps = sub0;
        s->iters = it0;
        s->stage = 0;
        s->it_sub = 0;
        s->hashed = false;
        s->nbr_valid = false;
        if (rc) return rc;
        if (e != hipSuccess)
            return fail(s, PS_ERR_DEVICE, std::string("autotune restore: ") + hipGetErrorString(e));
        tune_cache_store();
    }
    drop_graph(s);
    if (blocks_out) memcpy(blocks_out, s->block, sizeof(s->block));
    return PS_OK;
}

int ps_get_tuning(const ps_system *s, uint32_t *blocks_out)
{
    if (!s || !blocks_out) return fail(nullptr, PS_ERR_INVALID_ARGUMENT, "NULL argument");
    memcpy(blocks_out, s->block, sizeof(s->block));
    return PS_OK;
}

int ps_set_tuning(ps_system *s, const uint32_t *blocks)
{
    if (int rc = check_sys(s)) return rc;
    if (!blocks) return fail(s, PS_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int k : kTunable)
        if (blocks[k] < 64 || blocks[k] > 512 || blocks[k] % 64)
            return fail(s, PS_ERR_INVALID_ARGUMENT, "block sizes must be 64..512, multiples of 64");
    for (int k : kTunable) s->block[k] = blocks[k];
    drop_graph(s);
    return PS_OK;
}

const char *ps_kernel_name(int kernel_id)
{
    static const char *names[PS_K_COUNT] = {"predict_hash", "scan", "scatter", "reorder", "neighbors",
                                            "lambda", "solve", "shape", "finalize", "adj_xsph",
                                            "adj_finalize", "adj_pre", "adj_lambda", "adj_position",
                                            "adj_stiffness", "adj_predict", "export"};
    if (kernel_id < 0 || kernel_id >= PS_K_COUNT) return "unknown";
    return names[kernel_id];
}

}  // extern "C"
