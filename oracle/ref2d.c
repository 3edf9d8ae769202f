/*
 * ref2d.c — CPU oracle of the reference's own CPU solver (TEST
 * INFRASTRUCTURE ONLY; loaded by tests/ and bench.py's CPU baseline, never
 * by the product library).
 *
 * Restates Simulation::tick (cpu/src/simulation.cpp:115-369) for scenes of
 * fluid objects inside a box: prediction (:139-161), boundary constraints
 * created from the predicted positions (:201-224) and counted (:232-245),
 * SOLVER_ITERATIONS passes over the CONTACT group (boundaries,
 * BoundaryConstraint::project, cpu/src/constraint/boundaryconstraint.cpp:
 * 14-93) then the STANDARD group (one TotalFluidConstraint::project per
 * fluid object in creation order, cpu/src/constraint/totalfluidconstraint.cpp:
 * 41-116, and one GasConstraint::project per gas object,
 * cpu/src/constraint/gasconstraint.cpp:31-114), velocity v = (ep - p)/dt and
 * sleeping (Particle::confirmGuess, cpu/src/particle.h:60-65).  Gas objects
 * (2D, as the reference's dvec2 forms) also accumulate a per-particle force
 * (open-gas drag and the vorticity term) that the next prediction applies,
 * with gas gravity scaled by ALPHA (simulation.cpp:143-154, simulation.h:21).
 * Immovable solids (new Particle(pos, 0)) and their ContactConstraints with
 * every other particle (simulation.cpp:164-199, contactconstraint.cpp:14-40,
 * with Particle::scaleMass, particle.h:67-73), smoke emitters
 * (opensmokeemitter.cpp:17-46, simulation.cpp:348-362), fluid emitters
 * (fluidemitter.cpp:13-79) and, with PS_FLAG_STABILIZATION, the
 * USE_STABILIZATION pre-pass over the boundaries (simulation.cpp:247-270).
 * Parity unpinned: the height-scaled mass tmass = m e^{-y} of the contacts,
 * which cancels from every contact with an immovable partner (the only
 * contacts here), so no closed form sees more than its rounding.  The reference's compile-time constants
 * (PARTICLE_RAD, H, RELAXATION, K_P, E_P, DQ_P, S_SOLID) are runtime
 * parameters; frand() jitter becomes the counter RNG of ps_oracle.c; the
 * vector operations follow glm's order (normalize(r) = r * (1/sqrt(r.r))).
 * Works in 2D (dims = 2) and 3D.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ps.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#if ORC_DOUBLE
typedef double real;
#define SQRT sqrt
#define SFX(name) name##_f64
#else
typedef float real;
#define SQRT sqrtf
#define SFX(name) name##_f32
#endif
#define R(x) ((real)(x))
#define EXPORT __attribute__((visibility("default")))

#include "orc_exp.h"

#define BX_LOW 1u
#define BX_HIGH 2u
#define BY_LOW 4u
#define BY_HIGH 8u
#define BZ_LOW 16u
#define BZ_HIGH 32u

static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* boundary jitter in [0,1): SplitMix64 finaliser of seed + golden * counter
 * + 1, counter = (iteration, insertion index, axis), top 24 bits (replaces
 * frand(), boundaryconstraint.cpp:19) */
static real uniform(uint64_t seed, uint64_t iter, uint32_t i, uint32_t axis)
{
    uint64_t ctr = (iter << 35) | ((uint64_t)i << 3) | (uint64_t)axis;
    uint64_t z = mix64(seed + ctr * 0x9E3779B97F4A7C15ull + 1u);
    return R((double)(uint32_t)(z >> 40) * (1.0 / 16777216.0));
}

typedef struct r2d_sys r2d_sys;
static real poly6(const r2d_sys *s, real r2);
static real length3(const real *r);

/* GasConstraint's compile-time constants (gasconstraint.h:10-19) and the
 * gas gravity scale (simulation.h:21) */
#define GAS_RELAXATION 0.01
#define GAS_K_P 0.2
#define GAS_E_P 4u
#define GAS_DQ_P 0.25
#define GAS_S_SOLID 0.5
#define GAS_ALPHA (-0.2)
#define GAS_OPEN_DRAG (-50.0)

/* FluidEmitter (cpu/src/fluidemitter.cpp:3-79): double timers, the fluid
 * object it feeds and solidifies, and the count of its emissions (the
 * counter of the frand() replacement) */
typedef struct {
    double timer, total, rate;
    real pos[3];
    int32_t group;
    uint64_t emitted;
} r2d_femitter;

/* OpenSmokeEmitter (cpu/src/opensmokeemitter.cpp:3-51): a double timer,
 * markers (diffuse particles, not simulated) and an optional gas object
 * that receives one particle per period */
typedef struct {
    double timer, rate;
    real pos[3];
    int32_t group; /* gas object, -1 none */
    uint32_t nmark, cap;
    real *mark; /* 3 per marker */
} r2d_emitter;

struct r2d_sys {
    r2d_emitter em[16];
    uint32_t nem;
    r2d_femitter fem[4];
    uint32_t nfem;
    real *tm;          /* tmass, Particle::scaleMass (particle.h:67-73) */
    real *tt;          /* Particle::t, the FluidEmitter's solidification timer */
    uint32_t *cstart;  /* contacts created while visiting particle i: [cstart[i], cstart[i+1]) */
    uint32_t *cpair;   /* 2 per contact, in creation order */
    uint32_t ncontact, ccap;
    uint32_t *members; /* scratch: a fluid object's particles */
    uint8_t *snap;     /* scratch: membership before a FluidEmitter tick */
    ps_params prm;
    uint32_t n, cap, ngroup;
    real *p, *ep, *v; /* 3 per particle */
    real *force;      /* 3 per particle: applied at the next prediction */
    real *w;
    int32_t *phase;
    int32_t *group; /* fluid or gas object, -1 none */
    real *group_rho0;
    int32_t *group_gas; /* 0 fluid, 1 closed gas, 2 open gas */
    real g_wdq;         /* poly6(DQ_P^2 H^2) of the gas constraint */
    uint32_t *bflags, *counts;
    real *lambda, *delta;
    uint32_t *nnb;
    real poly6, spiky, h, h2, w_dq, radius, relax, k, jit;
    uint64_t iters;
};

EXPORT int SFX(r2d_create)(const ps_params *prm, r2d_sys **out)
{
    r2d_sys *s = (r2d_sys *)calloc(1, sizeof(r2d_sys));
    s->prm = *prm;
    s->cap = prm->max_particles;
    size_t c = s->cap ? s->cap : 1;
    s->p = (real *)calloc(3 * c, sizeof(real));
    s->ep = (real *)calloc(3 * c, sizeof(real));
    s->v = (real *)calloc(3 * c, sizeof(real));
    s->force = (real *)calloc(3 * c, sizeof(real));
    s->group_gas = (int32_t *)calloc(c, sizeof(int32_t));
    s->w = (real *)calloc(c, sizeof(real));
    s->phase = (int32_t *)calloc(c, sizeof(int32_t));
    s->group = (int32_t *)calloc(c, sizeof(int32_t));
    s->group_rho0 = (real *)calloc(c, sizeof(real));
    s->bflags = (uint32_t *)calloc(c, sizeof(uint32_t));
    s->counts = (uint32_t *)calloc(c, sizeof(uint32_t));
    s->lambda = (real *)calloc(c, sizeof(real));
    s->delta = (real *)calloc(3 * c, sizeof(real));
    s->nnb = (uint32_t *)calloc(c, sizeof(uint32_t));
    s->tm = (real *)calloc(c, sizeof(real));
    s->tt = (real *)calloc(c, sizeof(real));
    s->cstart = (uint32_t *)calloc(c + 1, sizeof(uint32_t));
    s->members = (uint32_t *)calloc(c, sizeof(uint32_t));
    s->snap = (uint8_t *)calloc(c, 1);
    /* constants as totalfluidconstraint.cpp:123-137 with a runtime H */
    const double pi = 3.14159265358979323846, h = prm->h;
    const double h2 = h * h, h6 = h2 * h2 * h2, h9 = h6 * h2 * h;
    const double poly6 = 315.0 / (64.0 * pi * h9);
    /* poly6(DQ_P * DQ_P * H * H), totalfluidconstraint.cpp:107 */
    const double dq = prm->dq_corr, t = h2 - (((dq * dq) * h) * h);
    s->poly6 = R(poly6);
    s->spiky = R(45.0 / (pi * h6));
    s->h = R(prm->h);
    s->h2 = R(h2);
    s->w_dq = R(poly6 * ((t * t) * t));
    /* poly6(DQ_P * DQ_P * H * H), gasconstraint.cpp:94 */
    const double tg = h2 - (((GAS_DQ_P * GAS_DQ_P) * h) * h);
    s->g_wdq = R(poly6 * ((tg * tg) * tg));
    s->radius = R(prm->radius);
    s->relax = R(prm->relaxation);
    s->k = R(prm->k_corr);
    s->jit = R((double)prm->wall_jitter * (double)prm->radius);
    *out = s;
    return PS_OK;
}

/* derived constants: poly6 and spiky coefficients, W(0), W(dq h), jitter */
EXPORT void SFX(r2d_constants)(const r2d_sys *s, double *out)
{
    out[0] = s->poly6;
    out[1] = s->spiky;
    out[2] = poly6(s, R(0));
    out[3] = s->w_dq;
    out[4] = s->jit;
}

/* OpenMP threads of the per-particle loops (1 = the reference's single
 * thread, 0 = every processor) */
EXPORT void SFX(r2d_set_threads)(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
#else
    (void)n;
#endif
}

EXPORT void SFX(r2d_destroy)(r2d_sys *s)
{
    if (!s) return;
    void *ptrs[] = {s->p, s->ep, s->v, s->force, s->w, s->phase, s->group, s->group_rho0,
                    s->group_gas, s->bflags, s->counts, s->lambda, s->delta, s->nnb,
                    s->tm, s->tt, s->cstart, s->cpair, s->members, s->snap};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    for (uint32_t e = 0; e < s->nem; e++) free(s->em[e].mark);
    free(s);
}

EXPORT int SFX(r2d_add_particles)(r2d_sys *s, uint32_t n, const float *pos, const float *vel,
                                  const float *w, const int32_t *phase)
{
    if ((uint64_t)s->n + n > s->cap) return PS_ERR_CAPACITY;
    /* solids only as immovable particles (new Particle(pos, 0), mass 0 =>
     * imass 0): their contacts are ContactConstraints; movable solids would
     * need RigidContactConstraint, which this restatement leaves out */
    for (uint32_t k = 0; k < n; k++)
        if (phase[k] == PS_PHASE_SOLID && w[k] != 0.0f) return PS_ERR_UNSUPPORTED;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t i = s->n + k;
        for (int a = 0; a < 3; a++) {
            s->p[3 * i + a] = pos[3 * k + a];
            s->v[3 * i + a] = vel ? R(vel[3 * k + a]) : R(0);
        }
        s->w[i] = w[k];
        s->phase[i] = phase[k];
        s->group[i] = -1;
        s->tt[i] = R(4); /* Particle::init, particle.h:37-38 */
    }
    s->n += n;
    return PS_OK;
}

/* createFluid (cpu/src/simulation.cpp:431-449): particles [first, first+n)
 * form one TotalFluidConstraint of rest density rho0 */
EXPORT int SFX(r2d_add_fluid_group)(r2d_sys *s, uint32_t first, uint32_t n, float rho0)
{
    if ((uint64_t)first + n > s->n) return PS_ERR_INVALID_ARGUMENT;
    for (uint32_t i = first; i < first + n; i++) s->group[i] = (int32_t)s->ngroup;
    s->group_rho0[s->ngroup] = rho0;
    s->group_gas[s->ngroup] = 0;
    s->ngroup++;
    return PS_OK;
}

/* createGas (cpu/src/simulation.cpp:408-428): particles [first, first+n)
 * form one GasConstraint of rest density rho0, open (drag towards rest
 * density) or closed; 2D only (the constraint's forms are the reference's
 * dvec2 ones) */
EXPORT int SFX(r2d_add_gas_group)(r2d_sys *s, uint32_t first, uint32_t n, float rho0, int open)
{
    if ((uint64_t)first + n > s->n || s->prm.dims != 2) return PS_ERR_INVALID_ARGUMENT;
    for (uint32_t i = first; i < first + n; i++) {
        s->group[i] = (int32_t)s->ngroup;
        s->phase[i] = PS_PHASE_GAS;
    }
    s->group_rho0[s->ngroup] = rho0;
    s->group_gas[s->ngroup] = open ? 2 : 1;
    s->ngroup++;
    return PS_OK;
}

/* createSmokeEmitter (simulation.cpp:454-457): emits particles_per_sec
 * markers, and as many gas particles into gas object `group` (-1: none) */
EXPORT int SFX(r2d_add_smoke_emitter)(r2d_sys *s, const float *pos, double rate, int32_t group,
                                      uint32_t max_markers)
{
    if (s->nem >= 16 || !(rate > 0.0)) return PS_ERR_INVALID_ARGUMENT;
    if (group >= (int32_t)s->ngroup || (group >= 0 && !s->group_gas[group])) return PS_ERR_INVALID_ARGUMENT;
    r2d_emitter *e = &s->em[s->nem];
    memset(e, 0, sizeof(*e));
    e->rate = rate;
    for (int a = 0; a < 3; a++) e->pos[a] = pos[a];
    e->group = group;
    e->cap = max_markers;
    e->mark = (real *)calloc(3 * (size_t)(max_markers ? max_markers : 1), sizeof(real));
    s->nem++;
    return PS_OK;
}

EXPORT uint32_t SFX(r2d_read_markers)(const r2d_sys *s, uint32_t emitter, real *xyz)
{
    const r2d_emitter *e = &s->em[emitter];
    if (xyz) memcpy(xyz, e->mark, sizeof(real) * 3 * e->nmark);
    return e->nmark;
}

/* OpenSmokeEmitter::tick (opensmokeemitter.cpp:17-46) at the end of the
 * tick, then the markers' box clamp (simulation.cpp:348-362): emit on the
 * timer (a marker, and a gas particle at rest into the object), then move
 * every marker by the poly6-weighted mean velocity of all particles,
 * p += (v dt) / sum */
static int emitter_tick(r2d_sys *s, r2d_emitter *e, real dt)
{
    e->timer += (double)dt;
    while (e->timer >= 1.0 / e->rate) {
        e->timer -= 1.0 / e->rate;
        if (e->nmark >= e->cap) return PS_ERR_CAPACITY;
        for (int a = 0; a < 3; a++) e->mark[3 * e->nmark + a] = e->pos[a];
        e->nmark++;
        if (e->group >= 0) {
            if (s->n >= s->cap) return PS_ERR_CAPACITY;
            uint32_t i = s->n++;
            for (int a = 0; a < 3; a++) {
                s->p[3 * i + a] = e->pos[a];
                s->ep[3 * i + a] = e->pos[a];
                s->v[3 * i + a] = R(0);
                s->force[3 * i + a] = R(0);
            }
            s->w[i] = R(1);
            s->phase[i] = PS_PHASE_GAS;
            s->group[i] = e->group;
        }
    }
    for (uint32_t m = 0; m < e->nmark; m++) {
        real *q = &e->mark[3 * m];
        real v[3] = {R(0), R(0), R(0)}, sum = R(0);
        for (uint32_t j = 0; j < s->n; j++) {
            real r[3] = {q[0] - s->p[3 * j], q[1] - s->p[3 * j + 1], q[2] - s->p[3 * j + 2]};
            real p6 = poly6(s, ((r[0] * r[0]) + (r[1] * r[1])) + (r[2] * r[2]));
            for (int a = 0; a < 3; a++) v[a] = v[a] + (s->v[3 * j + a] * p6);
            sum = sum + p6;
        }
        if (sum > R(0))
            for (int a = 0; a < 3; a++) q[a] = q[a] + ((v[a] * dt) / sum);
        for (int a = 0; a < 2; a++) {
            if (q[a] < R(s->prm.box_min[a]))
                q[a] = R(s->prm.box_min[a]);
            else if (q[a] > R(s->prm.box_max[a]))
                q[a] = R(s->prm.box_max[a]);
        }
    }
    return PS_OK;
}

/* createFluidEmitter (simulation.cpp:459-461): particles_per_sec fluid
 * particles into fluid object `group` for its first 5 seconds */
EXPORT int SFX(r2d_add_fluid_emitter)(r2d_sys *s, const float *pos, double rate, int32_t group)
{
    if (s->nfem >= 4 || !(rate > 0.0)) return PS_ERR_INVALID_ARGUMENT;
    if (group < 0 || group >= (int32_t)s->ngroup || s->group_gas[group]) return PS_ERR_INVALID_ARGUMENT;
    r2d_femitter *e = &s->fem[s->nfem];
    memset(e, 0, sizeof(*e));
    e->rate = rate;
    for (int a = 0; a < 3; a++) e->pos[a] = pos[a];
    e->group = group;
    s->nfem++;
    return PS_OK;
}

/* phases, inverse masses and solidification timers (t) of every particle */
EXPORT int SFX(r2d_read_attrs)(const r2d_sys *s, int32_t *phase, real *w, real *t)
{
    if (phase) memcpy(phase, s->phase, sizeof(int32_t) * s->n);
    if (w) memcpy(w, s->w, sizeof(real) * s->n);
    if (t) memcpy(t, s->tt, sizeof(real) * s->n);
    return PS_OK;
}

/* frand() of an emitted particle's velocity (fluidemitter.cpp:74), in
 * [0, 1): SplitMix64 finaliser of seed + golden * counter + 1 with counter
 * 2^62 | emitter << 40 | emission, top 24 bits (a stream apart from the
 * boundary jitter's counters, which stay below 2^62) */
static real emit_uniform(uint64_t seed, uint32_t emitter, uint64_t k)
{
    uint64_t ctr = (1ull << 62) | ((uint64_t)emitter << 40) | k;
    uint64_t z = mix64(seed + ctr * 0x9E3779B97F4A7C15ull + 1u);
    return R((double)(uint32_t)(z >> 40) * (1.0 / 16777216.0));
}

/* ContactConstraint::project (cpu/src/constraint/contactconstraint.cpp:
 * 14-40) with the height-scaled masses tmass and the particles' counts:
 * diff = ep1 - ep2, mag = |diff| - diameter; when the pair overlaps,
 * dp = (mag / (t1 + t2) / |diff|) diff, ep1 -= t1 dp / n1, ep2 += t2 dp / n2 */
static void contact(r2d_sys *s, uint32_t i1, uint32_t i2)
{
    real *e1 = &s->ep[3 * i1], *e2 = &s->ep[3 * i2];
    const real t1 = s->tm[i1], t2 = s->tm[i2];
    if (t1 == R(0) && t2 == R(0)) return;
    real d[3] = {e1[0] - e2[0], e1[1] - e2[1], e1[2] - e2[2]};
    const real wsum = t1 + t2, dist = length3(d), mag = dist - (R(2) * s->radius);
    if (mag > R(0)) return;
    const real scale = mag / wsum;
    const real c1 = (real)s->counts[i1], c2 = (real)s->counts[i2];
    real dp[3];
    for (int a = 0; a < 3; a++) dp[a] = (scale / dist) * d[a];
    for (int a = 0; a < 3; a++) {
        e1[a] = e1[a] + (((-t1) * dp[a]) / c1);
        e2[a] = e2[a] + ((t2 * dp[a]) / c2);
    }
}

/* FluidEmitter::tick (cpu/src/fluidemitter.cpp:13-79) at the end of the
 * tick, after the smoke emitters (simulation.cpp:364-366).  Over the fluid
 * object's particles (its ps list, in index order; the reference walks it
 * from the back, but no decision depends on another): a particle slower
 * than .06 at height <= 5 loses 1 from t when lambdas[k] <= 0 and turns
 * into an immovable solid when t reaches 0 (imass 0, SOLID, ep = p, v = f =
 * 0, out of the object); otherwise it gains dt, up to 3.  `lambdas` is a
 * QHash keyed by particle index and the lookup passes the position k in ps:
 * it reads the lambda of particle k if particle k was in the object at its
 * last projection, else 0 (the hash's default).  Then emissions: a fluid
 * particle of mass 1 at pos with v = (frand(), 1) per 1/rate of the timer
 * while the total time is below 5 s. */
static int femitter_tick(r2d_sys *s, r2d_femitter *e, uint32_t eidx, real dt)
{
    const int32_t g = e->group;
    uint32_t m = 0;
    for (uint32_t i = 0; i < s->n; i++) {
        s->snap[i] = s->group[i] == g;
        if (s->snap[i]) s->members[m++] = i;
    }
    for (uint32_t k = m; k-- > 0;) {
        const uint32_t i = s->members[k];
        const real *v = &s->v[3 * i], *p = &s->p[3 * i];
        if (!(length3(v) < R(.06) && p[1] <= R(5))) continue;
        const real lam = (k < s->n && s->snap[k]) ? s->lambda[k] : R(0);
        if (lam <= R(0)) {
            s->tt[i] = s->tt[i] - R(1);
            if (s->tt[i] <= R(0)) {
                s->tt[i] = R(0);
                s->w[i] = R(0);
                s->phase[i] = PS_PHASE_SOLID;
                s->group[i] = -1;
                for (int a = 0; a < 3; a++) {
                    s->ep[3 * i + a] = s->p[3 * i + a];
                    s->v[3 * i + a] = R(0);
                    s->force[3 * i + a] = R(0);
                }
            }
        } else {
            s->tt[i] = s->tt[i] + dt;
            if (s->tt[i] > R(3)) s->tt[i] = R(3);
        }
    }
    e->timer += (double)dt;
    e->total += (double)dt;
    while (e->total < 5.0 && e->timer >= 1.0 / e->rate) {
        e->timer -= 1.0 / e->rate;
        if (s->n >= s->cap) return PS_ERR_CAPACITY;
        const uint32_t i = s->n++;
        for (int a = 0; a < 3; a++) {
            s->p[3 * i + a] = e->pos[a];
            s->ep[3 * i + a] = e->pos[a];
            s->force[3 * i + a] = R(0);
        }
        s->v[3 * i] = emit_uniform(s->prm.seed, eidx, e->emitted++);
        s->v[3 * i + 1] = R(1);
        s->v[3 * i + 2] = R(0);
        s->w[i] = R(1);
        s->phase[i] = PS_PHASE_FLUID;
        s->group[i] = g;
        s->tt[i] = R(4);
    }
    return PS_OK;
}

/* the force each particle carries into the next prediction */
EXPORT int SFX(r2d_read_force)(const r2d_sys *s, real *force)
{
    memcpy(force, s->force, sizeof(real) * 3 * s->n);
    return PS_OK;
}

/* spikyGrad (totalfluidconstraint.cpp:131-137): -normalize(r) * c * (H-l) * (H-l) */
static void spiky_grad(const r2d_sys *s, const real *r, real l, real *g)
{
    if (l >= s->h || l == R(0)) {
        g[0] = g[1] = g[2] = R(0);
        return;
    }
    real dot = ((r[0] * r[0]) + (r[1] * r[1])) + (r[2] * r[2]);
    real inv = R(1) / SQRT(dot);
    real q = s->h - l;
    for (int a = 0; a < 3; a++) g[a] = (((-(r[a] * inv)) * s->spiky) * q) * q;
}

/* pow(x, E_P) (totalfluidconstraint.cpp:107): the FP64 build calls pow as the
 * reference does; the FP32 build squares (x^4 = (x^2)^2), which the kernels
 * reproduce bit for bit */
static real ipow_corr(real x, uint32_t n)
{
#if ORC_DOUBLE
    return pow(x, (double)n);
#else
    real r = R(1), b = x;
    int first = 1;
    while (n) {
        if (n & 1u) {
            r = first ? b : (r * b);
            first = 0;
        }
        n >>= 1;
        if (n) b = b * b;
    }
    return r;
#endif
}

static real poly6(const r2d_sys *s, real r2)
{
    if (r2 >= s->h2) return R(0);
    real t = s->h2 - r2;
    return s->poly6 * ((t * t) * t);
}

static real length3(const real *r)
{
    return SQRT(((r[0] * r[0]) + (r[1] * r[1])) + (r[2] * r[2]));
}

/* TotalFluidConstraint::project for fluid object g: lambdas, then deltas,
 * then ep += delta / (neighbours incl. self + counts) (Jacobi inside the
 * object; objects run one after another) */
static void fluid_group(r2d_sys *s, int32_t g)
{
    const uint32_t n = s->n;
    const real p0 = s->group_rho0[g];
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; i++) {
        if (s->group[i] != g) continue;
        real pi = R(0), den = R(0);
        real gi[3] = {R(0), R(0), R(0)};
        for (uint32_t j = 0; j < n; j++) {
            if (j != i) {
                if (s->w[j] == R(0)) continue;
                real r[3] = {s->ep[3 * i] - s->ep[3 * j], s->ep[3 * i + 1] - s->ep[3 * j + 1],
                             s->ep[3 * i + 2] - s->ep[3 * j + 2]};
                real r2 = ((r[0] * r[0]) + (r[1] * r[1])) + (r[2] * r[2]);
                if (r2 < s->h2) {
                    real incr = poly6(s, r2) / s->w[j];
                    if (s->phase[j] == PS_PHASE_SOLID) incr = incr * R(s->prm.s_solid);
                    pi = pi + incr;
                    real l = length3(r), sg[3];
                    spiky_grad(s, r, l, sg);
                    real gr[3];
                    for (int a = 0; a < 3; a++) gr[a] = (-sg[a]) / p0;
                    den = den + (((gr[0] * gr[0]) + (gr[1] * gr[1])) + (gr[2] * gr[2]));
                    real f = s->phase[j] == PS_PHASE_SOLID ? R(s->prm.s_solid) : R(1);
                    for (int a = 0; a < 3; a++) gi[a] = gi[a] + (f * sg[a]);
                }
            } else {
                pi = pi + (poly6(s, R(0)) / s->w[i]);
            }
        }
        real gr[3] = {gi[0] / p0, gi[1] / p0, gi[2] / p0};
        den = den + (((gr[0] * gr[0]) + (gr[1] * gr[1])) + (gr[2] * gr[2]));
        s->lambda[i] = -((pi / p0) - R(1)) / (den + s->relax);
    }
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; i++) {
        if (s->group[i] != g) continue;
        real d[3] = {R(0), R(0), R(0)};
        uint32_t cnt = 0;
        for (uint32_t j = 0; j < n; j++) {
            if (j == i) {
                cnt++;
                continue;
            }
            if (s->w[j] == R(0)) continue;
            real r[3] = {s->ep[3 * i] - s->ep[3 * j], s->ep[3 * i + 1] - s->ep[3 * j + 1],
                         s->ep[3 * i + 2] - s->ep[3 * j + 2]};
            real r2 = ((r[0] * r[0]) + (r[1] * r[1])) + (r[2] * r[2]);
            if (!(r2 < s->h2)) continue;
            cnt++;
            real l = length3(r), sg[3];
            spiky_grad(s, r, l, sg);
            real x = poly6(s, l * l) / s->w_dq;
            real corr = -(s->k * ipow_corr(x, s->prm.n_corr));
            real lj = s->group[j] == g ? s->lambda[j] : R(0);
            real coef = (s->lambda[i] + lj) + corr;
            for (int a = 0; a < 3; a++) d[a] = d[a] + (coef * sg[a]);
        }
        for (int a = 0; a < 3; a++) s->delta[3 * i + a] = d[a] / p0;
        s->nnb[i] = cnt;
    }
    for (uint32_t i = 0; i < n; i++) {
        if (s->group[i] != g) continue;
        real c = (real)(s->nnb[i] + s->counts[i]);
        for (int a = 0; a < 3; a++) s->ep[3 * i + a] = s->ep[3 * i + a] + (s->delta[3 * i + a] / c);
    }
}

/* GasConstraint::project for gas object g (gasconstraint.cpp:31-114): the
 * fluid constraint's lambda with the gas constants (self gradient summed
 * over every neighbour without the solid weight, :220-232), the open-gas
 * drag f += v (1 - rho/rho0) (-50) (:80), deltas with the gas tensile term
 * and the vorticity force f += sum_j cross((0,0,|G_j * v_j|), r) W(r^2)
 * (:96-103; G_j = spikyGrad(r, r.r), the squared length where the length
 * is expected, and G_j * v_j component-wise, as written there), then
 * ep += delta / (neighbours incl. self + counts). */
static void gas_group(r2d_sys *s, int32_t g)
{
    const uint32_t n = s->n;
    const real p0 = s->group_rho0[g];
    const int open = s->group_gas[g] == 2;
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; i++) {
        if (s->group[i] != g) continue;
        real pi = R(0), den = R(0);
        real gi[3] = {R(0), R(0), R(0)};
        for (uint32_t j = 0; j < n; j++) {
            if (j != i) {
                if (s->w[j] == R(0)) continue;
                real r[3] = {s->ep[3 * i] - s->ep[3 * j], s->ep[3 * i + 1] - s->ep[3 * j + 1],
                             s->ep[3 * i + 2] - s->ep[3 * j + 2]};
                real r2 = ((r[0] * r[0]) + (r[1] * r[1])) + (r[2] * r[2]);
                if (r2 < s->h2) {
                    real incr = poly6(s, r2) / s->w[j];
                    if (s->phase[j] == PS_PHASE_SOLID) incr = incr * R(GAS_S_SOLID);
                    pi = pi + incr;
                    real l = length3(r), sg[3];
                    spiky_grad(s, r, l, sg);
                    real gr[3];
                    for (int a = 0; a < 3; a++) gr[a] = (-sg[a]) / p0;
                    den = den + (((gr[0] * gr[0]) + (gr[1] * gr[1])) + (gr[2] * gr[2]));
                    for (int a = 0; a < 3; a++) gi[a] = gi[a] + sg[a];
                }
            } else {
                pi = pi + (poly6(s, R(0)) / s->w[i]);
            }
        }
        real gr[3] = {gi[0] / p0, gi[1] / p0, gi[2] / p0};
        den = den + (((gr[0] * gr[0]) + (gr[1] * gr[1])) + (gr[2] * gr[2]));
        real p_rat = pi / p0;
        if (open)
            for (int a = 0; a < 2; a++)
                s->force[3 * i + a] = s->force[3 * i + a] + ((s->v[3 * i + a] * (R(1) - p_rat)) * R(GAS_OPEN_DRAG));
        s->lambda[i] = -(p_rat - R(1)) / (den + R(GAS_RELAXATION));
    }
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; i++) {
        if (s->group[i] != g) continue;
        real d[3] = {R(0), R(0), R(0)};
        real fv[2] = {R(0), R(0)};
        uint32_t cnt = 0;
        for (uint32_t j = 0; j < n; j++) {
            if (j == i) {
                cnt++;
                continue;
            }
            if (s->w[j] == R(0)) continue;
            real r[3] = {s->ep[3 * i] - s->ep[3 * j], s->ep[3 * i + 1] - s->ep[3 * j + 1],
                         s->ep[3 * i + 2] - s->ep[3 * j + 2]};
            real r2 = ((r[0] * r[0]) + (r[1] * r[1])) + (r[2] * r[2]);
            if (!(r2 < s->h2)) continue;
            cnt++;
            real l = length3(r), sg[3];
            spiky_grad(s, r, l, sg);
            real x = poly6(s, l * l) / s->g_wdq;
            real corr = -(R(GAS_K_P) * ipow_corr(x, GAS_E_P));
            real lj = s->group[j] == g ? s->lambda[j] : R(0);
            real coef = (s->lambda[i] + lj) + corr;
            for (int a = 0; a < 3; a++) d[a] = d[a] + (coef * sg[a]);
            /* vorticity */
            real vg[3];
            spiky_grad(s, r, r2, vg);
            real wv[2] = {vg[0] * s->v[3 * j], vg[1] * s->v[3 * j + 1]};
            real W = SQRT((wv[0] * wv[0]) + (wv[1] * wv[1]));
            real cx = (R(0) * R(0)) - (r[1] * W), cy = (W * r[0]) - (R(0) * R(0));
            real p6 = poly6(s, r2);
            fv[0] = fv[0] + (cx * p6);
            fv[1] = fv[1] + (cy * p6);
        }
        for (int a = 0; a < 3; a++) s->delta[3 * i + a] = d[a] / p0;
        s->nnb[i] = cnt;
        s->force[3 * i] = s->force[3 * i] + fv[0];
        s->force[3 * i + 1] = s->force[3 * i + 1] + fv[1];
    }
    for (uint32_t i = 0; i < n; i++) {
        if (s->group[i] != g) continue;
        real c = (real)(s->nnb[i] + s->counts[i]);
        for (int a = 0; a < 3; a++) s->ep[3 * i + a] = s->ep[3 * i + a] + (s->delta[3 * i + a] / c);
    }
}

/* BoundaryConstraint::project (boundaryconstraint.cpp:14-93), one axis */
static void boundary(r2d_sys *s, uint32_t i, int axis, int greater, uint64_t iter)
{
    const real r = s->radius;
    const real value = greater ? R(s->prm.box_min[axis]) : R(s->prm.box_max[axis]);
    const int fl = s->phase[i] == PS_PHASE_FLUID || s->phase[i] == PS_PHASE_GAS;
    const real extra = fl ? uniform(s->prm.seed, iter, i, (uint32_t)axis) * s->jit : R(0);
    const real d = r + extra;
    real *e = &s->ep[3 * i];
    real nv[3] = {R(0), R(0), R(0)};
    if (greater) {
        if (e[axis] >= value + r) return;
        e[axis] = value + d;
        nv[axis] = R(1);
    } else {
        if (e[axis] <= value - r) return;
        e[axis] = value - d;
        nv[axis] = R(-1);
    }
    real c = (real)s->counts[i];
    real dp[3];
    for (int a = 0; a < 3; a++) dp[a] = (e[a] - s->p[3 * i + a]) / c;
    real dn = ((dp[0] * nv[0]) + (dp[1] * nv[1])) + (dp[2] * nv[2]);
    real t[3];
    for (int a = 0; a < 3; a++) t[a] = dp[a] - (dn * nv[a]);
    real l = length3(t);
    if (l < R(1e-4)) return;
    const real sf = SQRT(R(s->prm.wall_mu_s)), kf = SQRT(R(s->prm.wall_mu_k));
    if (l < sf * d) {
        for (int a = 0; a < 3; a++) e[a] = e[a] - t[a];
    } else {
        real f = (kf * d) / l;
        if (f > R(1)) f = R(1);
        for (int a = 0; a < 3; a++) e[a] = e[a] - (t[a] * f);
    }
}

/* BoundaryConstraint::project with stabile = true (boundaryconstraint.cpp:
 * 14-71), the stabilisation pre-pass of USE_STABILIZATION: p moves with ep,
 * no friction; jitter from the counter 2^63 | iteration << 36 | pass << 35
 * | index << 3 | axis */
static void stab_boundary(r2d_sys *s, uint32_t i, int axis, int greater, uint64_t iter, uint32_t pass)
{
    const real r = s->radius;
    const real value = greater ? R(s->prm.box_min[axis]) : R(s->prm.box_max[axis]);
    const int fl = s->phase[i] == PS_PHASE_FLUID || s->phase[i] == PS_PHASE_GAS;
    real extra = R(0);
    if (fl) {
        uint64_t ctr = (1ull << 63) | (iter << 36) | ((uint64_t)pass << 35) | ((uint64_t)i << 3) |
                       (uint64_t)axis;
        uint64_t z = mix64(s->prm.seed + ctr * 0x9E3779B97F4A7C15ull + 1u);
        extra = R((double)(uint32_t)(z >> 40) * (1.0 / 16777216.0)) * s->jit;
    }
    const real d = r + extra;
    real *e = &s->ep[3 * i], *p = &s->p[3 * i];
    if (greater) {
        if (e[axis] >= value + r) return;
        e[axis] = value + d;
        p[axis] = value + d;
    } else {
        if (e[axis] <= value - r) return;
        e[axis] = value - d;
        p[axis] = value - d;
    }
}

EXPORT int SFX(r2d_step)(r2d_sys *s, real dt, uint32_t steps)
{
    const real gx = R(s->prm.gravity[0]), gy = R(s->prm.gravity[1]), gz = R(s->prm.gravity[2]);
    const int twod = s->prm.dims == 2;
    for (uint32_t st = 0; st < steps; st++) {
        /* (1)-(5) predict; (8) boundary contacts from the predicted positions */
        for (uint32_t i = 0; i < s->n; i++) {
            real *v = &s->v[3 * i], *p = &s->p[3 * i], *e = &s->ep[3 * i], *fo = &s->force[3 * i];
            /* v = v + dt g' + dt f, g' = g ALPHA for gas; f = 0 */
            real g3[3] = {gx, gy, gz};
            if (s->phase[i] == PS_PHASE_GAS)
                for (int a = 0; a < 3; a++) g3[a] = g3[a] * R(GAS_ALPHA);
            v[0] = (v[0] + (dt * g3[0])) + (dt * fo[0]);
            v[1] = (v[1] + (dt * g3[1])) + (dt * fo[1]);
            if (!twod) v[2] = (v[2] + (dt * g3[2])) + (dt * fo[2]);
            fo[0] = fo[1] = fo[2] = R(0);
            for (int a = 0; a < 3; a++)
                e[a] = s->w[i] == R(0) ? p[a] : ((twod && a == 2) ? p[a] : (p[a] + (dt * v[a])));
            uint32_t f = 0;
            const real r = s->radius;
            if (e[0] < R(s->prm.box_min[0]) + r) f |= BX_LOW;
            else if (e[0] > R(s->prm.box_max[0]) - r) f |= BX_HIGH;
            if (e[1] < R(s->prm.box_min[1]) + r) f |= BY_LOW;
            else if (e[1] > R(s->prm.box_max[1]) - r) f |= BY_HIGH;
            if (!twod) {
                if (e[2] < R(s->prm.box_min[2]) + r) f |= BZ_LOW;
                else if (e[2] > R(s->prm.box_max[2]) - r) f |= BZ_HIGH;
            }
            s->bflags[i] = f;
            s->counts[i] = (uint32_t)__builtin_popcount(f);
            /* (4) scaleMass from the start height */
            s->tm[i] = s->w[i] != R(0) ? R(1) / ((R(1) / s->w[i]) * orc_exp(-p[1])) : R(0);
        }
        /* (7) contacts of a solid with any other particle, not both
         * immovable, closer than a diameter less EPSILON; counted for both */
        s->ncontact = 0;
        const real touch = R(2.0 * (double)s->prm.radius - 1e-4);
        for (uint32_t i = 0; i < s->n; i++) {
            s->cstart[i] = s->ncontact;
            for (uint32_t j = i + 1; j < s->n; j++) {
                if (s->w[i] == R(0) && s->w[j] == R(0)) continue;
                if (s->phase[i] != PS_PHASE_SOLID && s->phase[j] != PS_PHASE_SOLID) continue;
                const real *a = &s->ep[3 * i], *b = &s->ep[3 * j];
                real d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
                if (!(length3(d) < touch)) continue;
                if (s->ncontact == s->ccap) {
                    s->ccap = s->ccap ? 2 * s->ccap : 1024;
                    s->cpair = (uint32_t *)realloc(s->cpair, 2 * sizeof(uint32_t) * s->ccap);
                }
                s->cpair[2 * s->ncontact] = i;
                s->cpair[2 * s->ncontact + 1] = j;
                s->ncontact++;
                s->counts[i]++;
                s->counts[j]++;
            }
        }
        s->cstart[s->n] = s->ncontact;
        /* (10)-(15) the stabilisation pre-pass (simulation.cpp:247-270) */
        if (s->prm.flags & PS_FLAG_STABILIZATION) {
            const uint64_t it0 = s->iters & ((1ull << 27) - 1ull);
            for (uint32_t q = 0; q < 2; q++)
                for (uint32_t i = 0; i < s->n; i++) {
                    uint32_t f = s->bflags[i];
                    if (f & BX_LOW) stab_boundary(s, i, 0, 1, it0, q);
                    if (f & BX_HIGH) stab_boundary(s, i, 0, 0, it0, q);
                    if (f & BY_LOW) stab_boundary(s, i, 1, 1, it0, q);
                    if (f & BY_HIGH) stab_boundary(s, i, 1, 0, it0, q);
                    if (f & BZ_LOW) stab_boundary(s, i, 2, 1, it0, q);
                    if (f & BZ_HIGH) stab_boundary(s, i, 2, 0, it0, q);
                }
        }
        /* (16)-(22) iterations: CONTACT group (per particle, the contacts
         * made while visiting it, then its boundaries x, y, z), STANDARD
         * group (fluid objects in order) */
        for (uint32_t it = 0; it < s->prm.iterations; it++) {
            for (uint32_t i = 0; i < s->n; i++) {
                for (uint32_t c = s->cstart[i]; c < s->cstart[i + 1]; c++)
                    contact(s, s->cpair[2 * c], s->cpair[2 * c + 1]);
                uint32_t f = s->bflags[i];
                if (f & BX_LOW) boundary(s, i, 0, 1, s->iters);
                if (f & BX_HIGH) boundary(s, i, 0, 0, s->iters);
                if (f & BY_LOW) boundary(s, i, 1, 1, s->iters);
                if (f & BY_HIGH) boundary(s, i, 1, 0, s->iters);
                if (f & BZ_LOW) boundary(s, i, 2, 1, s->iters);
                if (f & BZ_HIGH) boundary(s, i, 2, 0, s->iters);
            }
            for (int32_t g = 0; g < (int32_t)s->ngroup; g++) {
                if (s->group_gas[g])
                    gas_group(s, g);
                else
                    fluid_group(s, g);
            }
            s->iters++;
        }
        /* (23)-(28) velocities, sleeping */
        for (uint32_t i = 0; i < s->n; i++) {
            real *v = &s->v[3 * i], *p = &s->p[3 * i], *e = &s->ep[3 * i];
            real d[3] = {e[0] - p[0], e[1] - p[1], e[2] - p[2]};
            for (int a = 0; a < 3; a++) v[a] = d[a] / dt;
            if ((s->prm.flags & PS_FLAG_SLEEPING) && length3(d) < R(1e-4)) {
                v[0] = v[1] = v[2] = R(0);
                continue;
            }
            for (int a = 0; a < 3; a++) p[a] = e[a];
        }
        for (uint32_t k = 0; k < s->nem; k++) {
            int rc = emitter_tick(s, &s->em[k], dt);
            if (rc) return rc;
        }
        for (uint32_t k = 0; k < s->nfem; k++) {
            int rc = femitter_tick(s, &s->fem[k], k, dt);
            if (rc) return rc;
        }
    }
    return PS_OK;
}

EXPORT uint32_t SFX(r2d_num_particles)(const r2d_sys *s) { return s->n; }

EXPORT int SFX(r2d_read)(const r2d_sys *s, real *pos, real *vel)
{
    if (pos) memcpy(pos, s->p, sizeof(real) * 3 * s->n);
    if (vel) memcpy(vel, s->v, sizeof(real) * 3 * s->n);
    return PS_OK;
}
