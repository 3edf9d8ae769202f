/*
 * ps_oracle.c — CPU oracle for the unified particle solver substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py load it; the product
 * library never links or calls it.
 *
 * It restates, in plain C99, the reference GPU substep
 * (gpu/src/particlesystem.cpp:144-246) with the semantic decisions recorded
 * for this build (grid rebuilt once per substep, canonical (cell, index)
 * sort order, standard PBF normalisation, per-particle counter RNG for the
 * wall jitter).  It is compiled twice:
 *   -DORC_DOUBLE=0  -> liboracle_f32.so, symbols *_f32  (FP32, same
 *                      operation order as the HIP kernels: bit-exact target)
 *   -DORC_DOUBLE=1  -> liboracle_f64.so, symbols *_f64  (FP64 trajectory for
 *                      error statements)
 * Every expression is written with explicit parentheses, and the build uses
 * -ffp-contract=off, so the FP32 build performs exactly the IEEE operations
 * written here.
 *
 * Pinning (tests/test_oracle_*.py): kernel constants against the
 * reference's POLY6_COEFF/SPIKEY_COEFF (integration_kernel.cuh:29-30); cell
 * keys, sort order, cell offsets and neighbour sets against numpy brute
 * force; lambda/delta-p/contacts/distance/walls/XSPH against hand-derived
 * closed forms on 1-3 particle scenes and invariants (rest lattice,
 * momentum, rigid motion); shape matching against an FP64 SVD polar
 * decomposition; FP32 against FP64 over short horizons.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ps.h"

#if ORC_DOUBLE
typedef double real;
#define SQRT sqrt
#define FLOOR floor
#define SFX(name) name##_f64
#else
typedef float real;
#define SQRT sqrtf
#define FLOOR floorf
#define SFX(name) name##_f32
#endif
#define R(x) ((real)(x))

/* Pair arithmetic of the neighbour search, lambda, delta-p and XSPH, shared
 * with the kernels: squared distances and accumulations use fused
 * multiply-add (one rounding, as fmaf and the GPU's v_fma_f32 both give),
 * and 1/r comes from RSQ below instead of a square root and a division. */
#if ORC_DOUBLE
#define FMA fma
static double RSQ(double x) { return 1.0 / sqrt(x); }
#else
#define FMA fmaf
/* 1/sqrt(x), x > 0 normal: the integer seed y0 = 0x5f375a86 - (bits(x) >> 1)
 * of the "fast inverse square root" (constant from C. Lomont, Fast Inverse
 * Square Root, 2003; relative error up to 3.4e-2), then one fourth-order
 * correction.  With e = 1 - x y0^2, 1/sqrt(x) = y0 (1 - e)^(-1/2) =
 * y0 (1 + e p(e)), p the binomial series 1/2 + 3/8 e + 5/16 e^2 +
 * 35/128 e^3 + ..., truncated after e^3, its four coefficients re-fitted
 * minimax over the seed's range of e (|e| < 0.07); y = fma(y0 e, p(e), y0),
 * p by Horner in FMA.  Relative error below 1e-7 over [1e-8, 16].  Only
 * IEEE operations, so the kernels' copy gives the same float. */
static float f32_of_bits(uint32_t u)
{
    float f;
    memcpy(&f, &u, sizeof(f));
    return f;
}
static float RSQ(float x)
{
    uint32_t i;
    memcpy(&i, &x, sizeof(i));
    const float y0 = f32_of_bits(0x5f375a86u - (i >> 1));
    const float xy = x * y0;
    const float e = fmaf(-xy, y0, 1.0f);
    float p = fmaf(f32_of_bits(0x3e8c5420u), e, f32_of_bits(0x3ea0bc22u));
    p = fmaf(p, e, f32_of_bits(0x3ec0000cu));
    p = fmaf(p, e, f32_of_bits(0x3effffc8u));
    return fmaf(y0 * e, p, y0);
}
#endif
/* RSQ exported for its accuracy test */
__attribute__((visibility("default"))) real SFX(orc_rsq)(real x) { return RSQ(x); }

/* |d|^2 = fma(dz, dz, fma(dy, dy, dx dx)) */
static real DIST2(real dx, real dy, real dz) { return FMA(dz, dz, FMA(dy, dy, dx * dx)); }
/* below this squared distance a pair is degenerate (no direction) */
#define DEGENERATE_R2 R(1e-8)
/* the half of a neighbour-table entry k: its chunk of four entries, even or odd */
#define HALF(k) ((int)(((k) >> 2) & 1u))
/* the quarter of table entry k: its chunk of four entries modulo 4 (the
 * XSPH sum and the reverse pass's sums) */
#define QUARTER(k) ((int)(((k) >> 2) & 3u))

#include "orc_exp.h"
#define EXP orc_exp
/* EXP exported for its accuracy test */
__attribute__((visibility("default"))) real SFX(orc_exp)(real x) { return EXP(x); }

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* derived constants: FP64 from the FP32 parameters, rounded once            */

typedef struct orc_derived {
    real h, h2, poly6, spiky, w_zero, w_dq, inv_w_dq, search_r2, collide_r2, collide_dist;
    /* 1/(h^2 - (dq h)^2) and h^2 times it: with s = FMA(-w_scale, r^2,
     * w_scale_h2), W(r)/W(dq h) = s^3 */
    real w_scale, w_scale_h2;
    real wall_s, wall_k, jitter, radius, relax, k_corr, xsph_c, s_solid, mu_s, mu_k;
    uint32_t n_corr, num_cells;
} orc_derived;

/* Kernel normalisations of gpu/src/cuda/integration_kernel.cuh:25-30
 * (POLY6 = 315/(64 pi H^9), SPIKEY = 45/(pi H^6)) for a runtime h. */
static void derive(const ps_params *p, orc_derived *d)
{
    const double pi = 3.14159265358979323846;
    double h = p->h, h2 = h * h, h3 = h2 * h, h6 = h3 * h3, h9 = h6 * h3;
    double poly6 = 315.0 / (64.0 * pi * h9);
    double spiky = 45.0 / (pi * h6);
    double dq = (double)p->dq_corr * h;
    double tq = h2 - dq * dq;
    double sr = p->h > p->collide_dist ? (double)p->h : (double)p->collide_dist;
    d->h = R(p->h);
    d->h2 = R(h2);
    d->poly6 = R(poly6);
    d->spiky = R(spiky);
    d->w_zero = R(poly6 * h6);
    d->w_dq = R(poly6 * tq * tq * tq);
    d->inv_w_dq = R(1.0 / (poly6 * tq * tq * tq));
    d->w_scale = R(1.0 / tq);
    d->w_scale_h2 = R(h2 / tq);
    d->search_r2 = R(sr * sr);
    d->collide_r2 = R((double)p->collide_dist * (double)p->collide_dist);
    d->collide_dist = R(p->collide_dist);
    d->wall_s = R(sqrt((double)p->wall_mu_s) * (double)p->radius);
    d->wall_k = R(sqrt((double)p->wall_mu_k) * (double)p->radius);
    d->jitter = R((double)p->wall_jitter * (double)p->radius);
    d->radius = R(p->radius);
    d->relax = R(p->relaxation);
    /* s_corr is added to lambda (units length^2): the paper's dimensionless
     * k measures lengths in units of h, so the term is k h^2 (...)^n; with
     * PS_FLAG_UNSCALED_TENSILE k as it is, as the reference's
     * lambdaCorr = -K_P pow(W/W(DQ_P H), E_P) (integration_kernel.cuh:628-633) */
    d->k_corr = (p->flags & PS_FLAG_UNSCALED_TENSILE) ? R(p->k_corr) : R((double)p->k_corr * h2);
    d->xsph_c = R(p->xsph_c);
    d->s_solid = R(p->s_solid);
    d->mu_s = R(p->mu_s);
    d->mu_k = R(p->mu_k);
    d->n_corr = p->n_corr;
    d->num_cells = p->grid[0] * p->grid[1] * p->grid[2];
}

/* Derived constants (FP64 values when ORC_DOUBLE) for pinning tests:
 * out[0..13] = h2, poly6, spiky, w_zero, w_dq, inv_w_dq, tensile_k,
 * search_r2, collide_r2, wall_s, wall_k, jitter, w_scale, w_scale_h2. */
ORC_EXPORT void SFX(orc_derive)(const ps_params *p, real *out)
{
    orc_derived d;
    derive(p, &d);
    out[0] = d.h2; out[1] = d.poly6; out[2] = d.spiky; out[3] = d.w_zero; out[4] = d.w_dq;
    out[5] = d.inv_w_dq; out[6] = d.k_corr; out[7] = d.search_r2; out[8] = d.collide_r2;
    out[9] = d.wall_s; out[10] = d.wall_k; out[11] = d.jitter; out[12] = d.w_scale;
    out[13] = d.w_scale_h2;
}

/* ------------------------------------------------------------------------ */
/* counter-based uniform for the wall jitter (replaces the reference's 6
 * cuRAND draws per call, integration.cu:326, and the CPU's per-particle
 * frand(), boundaryconstraint.cpp:19): SplitMix64 finaliser of
 * seed + golden * counter, top 24 bits -> [0, 1).                          */

static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static real wall_uniform(uint64_t seed, uint64_t iter, uint32_t orig, uint32_t wall)
{
    uint64_t ctr = (iter << 35) | ((uint64_t)orig << 3) | (uint64_t)wall;
    uint64_t z = mix64(seed + ctr * 0x9E3779B97F4A7C15ull);
    return R((double)(uint32_t)(z >> 40) * (1.0 / 16777216.0));
}

ORC_EXPORT double SFX(orc_wall_uniform)(uint64_t seed, uint64_t iter, uint32_t orig,
                                        uint32_t wall)
{
    return (double)wall_uniform(seed, iter, orig, wall);
}

/* ------------------------------------------------------------------------ */
/* system state (storage is kept in the last canonical sorted order)          */

typedef struct orc_sys {
    ps_params prm;
    orc_derived d;
    uint32_t n, cap, kmax, ncell;
    real *X;   /* 3*cap: positions at substep start (x*) */
    real *P;   /* 4*cap: iterate (x, y, z, lambda)      */
    real *Q;   /* 4*cap: next iterate                   */
    real *V;   /* 3*cap                                 */
    real *w, *rho0;
    int32_t *phase;
    uint32_t *orig, *key, *slot_of;
    uint32_t *off;    /* ncell + 1 */
    uint32_t *nbr;    /* cap * kmax, row-major */
    uint32_t *ncnt;
    /* scratch for the sort */
    uint64_t *sortkey;
    real *tmpr;
    uint32_t *tmpu;
    int32_t *tmpi;
    /* distance constraints, CSR by particle (insertion index) */
    uint32_t ndist;
    uint32_t *dist_pairs; /* 2*ndist */
    real *dist_rest, *dist_stiff;
    uint32_t *dist_off;   /* cap + 1 */
    uint32_t *dist_adj;   /* 2*ndist: constraint ids                  */
    uint32_t *occ;        /* per particle: distance + point incidences */
    /* point constraints */
    uint32_t npin;
    uint32_t *pin_of;     /* cap: pin index + 1, 0 = none (last wins) */
    real *pin_xyz;
    /* rigid bodies (shape matching) */
    uint32_t nbody, body_members_total;
    uint32_t *body_off;   /* nbody + 1 */
    uint32_t *body_idx;   /* members, insertion indices */
    real *body_rest;      /* 3 per member: rest offsets */
    real *body_stiff;
    real *body_mass;      /* total mass */
    real *body_q;         /* 4 per body: rotation quaternion (w,x,y,z), warm start */
    real *body_goal;      /* 3 per member: goal position of the current iteration */
    uint32_t *body_of;    /* cap: body id + 1 */
    uint32_t *member_of;  /* cap: member index within body_idx */
    real dt;
    uint64_t substeps, iters;
    uint32_t status_flags, max_seen;
    int stage; /* 0 idle, 1 begun */
    /* adjoint recompute: when set, solve_pass records per particle which
     * coordinates the walls or a pin overwrote (bits 0-2 x, y, z) */
    uint8_t *clamp_rec;
    real *wall_rec; /* 4 per particle, with clamp_rec */
} orc_sys;

static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    return p;
}

ORC_EXPORT int SFX(orc_create)(const ps_params *p, orc_sys **out)
{
    orc_sys *s = (orc_sys *)xcalloc(1, sizeof(orc_sys));
    if (!s) return PS_ERR_OUT_OF_MEMORY;
    s->prm = *p;
    derive(p, &s->d);
    s->cap = p->max_particles;
    s->kmax = p->max_neighbors;
    s->ncell = s->d.num_cells;
    size_t c = s->cap;
    s->X = (real *)xcalloc(3 * c, sizeof(real));
    s->P = (real *)xcalloc(4 * c, sizeof(real));
    s->Q = (real *)xcalloc(4 * c, sizeof(real));
    s->V = (real *)xcalloc(3 * c, sizeof(real));
    s->w = (real *)xcalloc(c, sizeof(real));
    s->rho0 = (real *)xcalloc(c, sizeof(real));
    s->phase = (int32_t *)xcalloc(c, sizeof(int32_t));
    s->orig = (uint32_t *)xcalloc(c, sizeof(uint32_t));
    s->key = (uint32_t *)xcalloc(c, sizeof(uint32_t));
    s->slot_of = (uint32_t *)xcalloc(c, sizeof(uint32_t));
    s->off = (uint32_t *)xcalloc((size_t)s->ncell + 1, sizeof(uint32_t));
    s->nbr = (uint32_t *)xcalloc(c * s->kmax, sizeof(uint32_t));
    s->ncnt = (uint32_t *)xcalloc(c, sizeof(uint32_t));
    s->sortkey = (uint64_t *)xcalloc(c, sizeof(uint64_t));
    s->tmpr = (real *)xcalloc(4 * c, sizeof(real));
    s->tmpu = (uint32_t *)xcalloc(c, sizeof(uint32_t));
    s->tmpi = (int32_t *)xcalloc(c, sizeof(int32_t));
    s->dist_off = (uint32_t *)xcalloc(c + 1, sizeof(uint32_t));
    s->occ = (uint32_t *)xcalloc(c, sizeof(uint32_t));
    s->pin_of = (uint32_t *)xcalloc(c, sizeof(uint32_t));
    s->body_of = (uint32_t *)xcalloc(c, sizeof(uint32_t));
    s->member_of = (uint32_t *)xcalloc(c, sizeof(uint32_t));
    s->body_off = (uint32_t *)xcalloc(1, sizeof(uint32_t));
    if (!s->X || !s->P || !s->Q || !s->V || !s->w || !s->rho0 || !s->phase || !s->orig ||
        !s->key || !s->slot_of || !s->off || !s->nbr || !s->ncnt || !s->sortkey || !s->tmpr ||
        !s->tmpu || !s->tmpi || !s->dist_off || !s->occ || !s->pin_of || !s->body_of ||
        !s->member_of || !s->body_off)
        return PS_ERR_OUT_OF_MEMORY;
    *out = s;
    return PS_OK;
}

ORC_EXPORT void SFX(orc_destroy)(orc_sys *s)
{
    if (!s) return;
    void *ptrs[] = {s->X, s->P, s->Q, s->V, s->w, s->rho0, s->phase, s->orig, s->key,
                    s->slot_of, s->off, s->nbr, s->ncnt, s->sortkey, s->tmpr, s->tmpu,
                    s->tmpi, s->dist_pairs, s->dist_rest, s->dist_stiff, s->dist_off,
                    s->dist_adj, s->occ, s->pin_of, s->pin_xyz, s->body_off, s->body_idx,
                    s->body_rest, s->body_stiff, s->body_mass, s->body_q, s->body_goal,
                    s->body_of, s->member_of};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(s);
}

ORC_EXPORT int SFX(orc_set_params)(orc_sys *s, const ps_params *p)
{
    s->prm = *p;
    derive(p, &s->d);
    return PS_OK;
}

/* Append particles in insertion order (addParticleMultiple,
 * gpu/src/particlesystem.cpp:333-348). */
ORC_EXPORT int SFX(orc_add_particles)(orc_sys *s, uint32_t n, const float *pos, const float *vel,
                                      const float *w, const int32_t *phase, const float *rho0)
{
    if ((uint64_t)s->n + n > s->cap) return PS_ERR_CAPACITY;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t i = s->n + k;
        s->X[3 * i + 0] = pos[3 * k + 0];
        s->X[3 * i + 1] = pos[3 * k + 1];
        s->X[3 * i + 2] = pos[3 * k + 2];
        for (int a = 0; a < 3; a++) {
            s->P[4 * i + a] = s->X[3 * i + a];
            s->V[3 * i + a] = vel ? R(vel[3 * k + a]) : R(0);
        }
        s->P[4 * i + 3] = R(0);
        s->w[i] = w[k];
        s->phase[i] = phase[k];
        s->rho0[i] = rho0[k];
        s->orig[i] = i;
        s->slot_of[i] = i;
    }
    s->n += n;
    return PS_OK;
}

/* ------------------------------------------------------------------------ */
/* constraints                                                              */

static void rebuild_incidence(orc_sys *s)
{
    memset(s->dist_off, 0, sizeof(uint32_t) * ((size_t)s->cap + 1));
    for (uint32_t c = 0; c < s->ndist; c++) {
        s->dist_off[s->dist_pairs[2 * c] + 1]++;
        s->dist_off[s->dist_pairs[2 * c + 1] + 1]++;
    }
    for (uint32_t i = 0; i < s->cap; i++) s->dist_off[i + 1] += s->dist_off[i];
    free(s->dist_adj);
    s->dist_adj = (uint32_t *)xcalloc(2 * (size_t)s->ndist, sizeof(uint32_t));
    uint32_t *fill = (uint32_t *)xcalloc(s->cap, sizeof(uint32_t));
    /* constraint ids in ascending order within each particle's list */
    for (uint32_t c = 0; c < s->ndist; c++) {
        uint32_t a = s->dist_pairs[2 * c], b = s->dist_pairs[2 * c + 1];
        s->dist_adj[s->dist_off[a] + fill[a]++] = c;
        s->dist_adj[s->dist_off[b] + fill[b]++] = c;
    }
    free(fill);
}

ORC_EXPORT int SFX(orc_add_distance_constraints)(orc_sys *s, uint32_t m, const uint32_t *pairs,
                                                 const float *rest, const float *stiff)
{
    for (uint32_t k = 0; k < 2 * m; k++)
        if (pairs[k] >= s->n) return PS_ERR_INVALID_ARGUMENT;
    uint32_t nd = s->ndist + m;
    s->dist_pairs = (uint32_t *)realloc(s->dist_pairs, sizeof(uint32_t) * 2 * (size_t)nd);
    s->dist_rest = (real *)realloc(s->dist_rest, sizeof(real) * (size_t)nd);
    s->dist_stiff = (real *)realloc(s->dist_stiff, sizeof(real) * (size_t)nd);
    for (uint32_t k = 0; k < m; k++) {
        s->dist_pairs[2 * (s->ndist + k)] = pairs[2 * k];
        s->dist_pairs[2 * (s->ndist + k) + 1] = pairs[2 * k + 1];
        s->dist_rest[s->ndist + k] = rest[k];
        s->dist_stiff[s->ndist + k] = stiff ? R(stiff[k]) : R(1);
        s->occ[pairs[2 * k]]++;
        s->occ[pairs[2 * k + 1]]++;
    }
    s->ndist = nd;
    rebuild_incidence(s);
    return PS_OK;
}

ORC_EXPORT int SFX(orc_add_point_constraints)(orc_sys *s, uint32_t m, const uint32_t *idx,
                                              const float *xyz)
{
    for (uint32_t k = 0; k < m; k++)
        if (idx[k] >= s->n) return PS_ERR_INVALID_ARGUMENT;
    uint32_t np = s->npin + m;
    s->pin_xyz = (real *)realloc(s->pin_xyz, sizeof(real) * 3 * (size_t)np);
    for (uint32_t k = 0; k < m; k++) {
        uint32_t pid = s->npin + k;
        s->pin_xyz[3 * pid + 0] = xyz[3 * k + 0];
        s->pin_xyz[3 * pid + 1] = xyz[3 * k + 1];
        s->pin_xyz[3 * pid + 2] = xyz[3 * k + 2];
        s->pin_of[idx[k]] = pid + 1;
        s->occ[idx[k]]++;
    }
    s->npin = np;
    return PS_OK;
}

/* forward declaration: defined with the shape-matching code below */
static void body_init(orc_sys *s, uint32_t b);

ORC_EXPORT int SFX(orc_add_rigid_body)(orc_sys *s, uint32_t n, const uint32_t *idx, float stiffness,
                                       uint32_t *body_out)
{
    if (n < 2) return PS_ERR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < n; k++) {
        if (idx[k] >= s->n || s->body_of[idx[k]]) return PS_ERR_INVALID_ARGUMENT;
        if (s->w[s->slot_of[idx[k]]] == R(0)) return PS_ERR_INVALID_ARGUMENT;
    }
    uint32_t b = s->nbody, tot = s->body_members_total + n;
    s->body_off = (uint32_t *)realloc(s->body_off, sizeof(uint32_t) * (b + 2));
    s->body_idx = (uint32_t *)realloc(s->body_idx, sizeof(uint32_t) * tot);
    s->body_rest = (real *)realloc(s->body_rest, sizeof(real) * 3 * (size_t)tot);
    s->body_goal = (real *)realloc(s->body_goal, sizeof(real) * 3 * (size_t)tot);
    s->body_stiff = (real *)realloc(s->body_stiff, sizeof(real) * (b + 1));
    s->body_mass = (real *)realloc(s->body_mass, sizeof(real) * (b + 1));
    s->body_q = (real *)realloc(s->body_q, sizeof(real) * 4 * (b + 1));
    s->body_off[b] = s->body_members_total;
    s->body_off[b + 1] = tot;
    for (uint32_t k = 0; k < n; k++) {
        s->body_idx[s->body_members_total + k] = idx[k];
        s->body_of[idx[k]] = b + 1;
        s->member_of[idx[k]] = s->body_members_total + k;
    }
    s->body_stiff[b] = stiffness > 0.f ? R(stiffness) : R(s->prm.shape_stiffness);
    s->nbody = b + 1;
    s->body_members_total = tot;
    body_init(s, b);
    if (body_out) *body_out = b;
    return PS_OK;
}

/* ------------------------------------------------------------------------ */
/* a1: predict (integrate_functor, integration_kernel.cuh:159-184; static
 * particles keep their position, cpu/src/particle.h:56-58)                 */

ORC_EXPORT int SFX(orc_predict)(orc_sys *s, real dt)
{
    const real gx = R(s->prm.gravity[0]), gy = R(s->prm.gravity[1]), gz = R(s->prm.gravity[2]);
    const int twod = s->prm.dims == 2;
    s->dt = dt;
    s->stage = 1; /* the substep is open from the prediction on (as ps_predict) */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)s->n; i++) {
        real x = s->X[3 * i], y = s->X[3 * i + 1], z = s->X[3 * i + 2];
        real px = x, py = y, pz = z;
        if (s->w[i] != R(0)) {
            real vx = s->V[3 * i] + (gx * dt);
            real vy = s->V[3 * i + 1] + (gy * dt);
            real vz = s->V[3 * i + 2] + (gz * dt);
            px = x + (vx * dt);
            py = y + (vy * dt);
            pz = twod ? z : (z + (vz * dt));
        }
        s->P[4 * i] = px;
        s->P[4 * i + 1] = py;
        s->P[4 * i + 2] = pz;
        s->P[4 * i + 3] = R(0);
    }
    return PS_OK;
}

/* ------------------------------------------------------------------------ */
/* a2: cell of a position (calcGridPos, integration_kernel.cuh:187-194):
 * floor((p - origin) / cell) by IEEE division, clamped into the dense grid
 * (the reference wraps with & (g-1), :199-201; a dense clamp keeps every
 * pair within one cell per axis adjacent, so no neighbour is lost).        */

static uint32_t cell_axis(real p, real o, real cell, uint32_t g, uint32_t *flags)
{
    real f = FLOOR((p - o) / cell);
    if (!(f >= R(0))) {
        if (f != f) *flags |= PS_STATUS_NONFINITE;
        else *flags |= PS_STATUS_OUTSIDE_GRID;
        return 0;
    }
    if (f >= (real)g) {
        *flags |= PS_STATUS_OUTSIDE_GRID;
        return g - 1;
    }
    return (uint32_t)f;
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* a2-a4: hash (calcHashD :206-225), sort by (cell, insertion index)
 * (sortParticles, integration.cu:270-275, with ties ordered by insertion
 * index), cell offsets (reorderDataAndFindCellStartD :229-299: cell c holds
 * sorted slots [off[c], off[c+1])) and reorder of every per-particle array. */
ORC_EXPORT int SFX(orc_build_grid)(orc_sys *s)
{
    const uint32_t n = s->n, gx = s->prm.grid[0], gy = s->prm.grid[1], gz = s->prm.grid[2];
    const real ox = R(s->prm.origin[0]), oy = R(s->prm.origin[1]), oz = R(s->prm.origin[2]);
    const real cell = R(s->prm.cell);
    uint32_t flags = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t cx = cell_axis(s->P[4 * i], ox, cell, gx, &flags);
        uint32_t cy = cell_axis(s->P[4 * i + 1], oy, cell, gy, &flags);
        uint32_t cz = cell_axis(s->P[4 * i + 2], oz, cell, gz, &flags);
        uint32_t key = (cz * gy + cy) * gx + cx;
        s->sortkey[i] = ((uint64_t)key << 32) | (uint64_t)s->orig[i];
    }
    s->status_flags |= flags;
    qsort(s->sortkey, n, sizeof(uint64_t), cmp_u64);
    /* permutation: sorted slot t <- storage slot slot_of[orig] */
    uint32_t *src = (uint32_t *)xcalloc(n, sizeof(uint32_t));
    for (uint32_t t = 0; t < n; t++) src[t] = s->slot_of[(uint32_t)(s->sortkey[t] & 0xffffffffu)];
    /* reorder reals: X(3), P(4), V(3), w, rho0 */
    real *tr = s->tmpr;
    for (uint32_t t = 0; t < n; t++)
        for (int a = 0; a < 3; a++) tr[3 * t + a] = s->X[3 * src[t] + a];
    memcpy(s->X, tr, sizeof(real) * 3 * n);
    for (uint32_t t = 0; t < n; t++) {
        for (int a = 0; a < 3; a++) tr[4 * t + a] = s->P[4 * src[t] + a];
        tr[4 * t + 3] = R(0);
    }
    memcpy(s->P, tr, sizeof(real) * 4 * n);
    for (uint32_t t = 0; t < n; t++)
        for (int a = 0; a < 3; a++) tr[3 * t + a] = s->V[3 * src[t] + a];
    memcpy(s->V, tr, sizeof(real) * 3 * n);
    for (uint32_t t = 0; t < n; t++) tr[t] = s->w[src[t]];
    memcpy(s->w, tr, sizeof(real) * n);
    for (uint32_t t = 0; t < n; t++) tr[t] = s->rho0[src[t]];
    memcpy(s->rho0, tr, sizeof(real) * n);
    for (uint32_t t = 0; t < n; t++) s->tmpi[t] = s->phase[src[t]];
    memcpy(s->phase, s->tmpi, sizeof(int32_t) * n);
    for (uint32_t t = 0; t < n; t++) {
        s->orig[t] = (uint32_t)(s->sortkey[t] & 0xffffffffu);
        s->key[t] = (uint32_t)(s->sortkey[t] >> 32);
        s->slot_of[s->orig[t]] = t;
    }
    free(src);
    /* cell offsets = exclusive scan of the per-cell counts */
    memset(s->off, 0, sizeof(uint32_t) * ((size_t)s->ncell + 1));
    for (uint32_t t = 0; t < n; t++) s->off[s->key[t] + 1]++;
    for (uint32_t c = 0; c < s->ncell; c++) s->off[c + 1] += s->off[c];
    (void)gz;
    return PS_OK;
}

/* Neighbour table (collideCell / collideCellRadius,
 * integration_kernel.cuh:303-356,481-518): every j != i within the search
 * radius, scanned over the 3x3 rows of x-adjacent cells (each row is one
 * contiguous range of sorted slots), so the list is in ascending slot order.
 * r^2 = DIST2.  Overflow keeps the first kmax entries and
 * raises PS_STATUS_NEIGHBOR_OVERFLOW (the reference truncates silently at
 * 500, :344,508). */
ORC_EXPORT int SFX(orc_find_neighbors)(orc_sys *s)
{
    const uint32_t n = s->n, gx = s->prm.grid[0], gy = s->prm.grid[1], gz = s->prm.grid[2];
    const uint32_t kmax = s->kmax;
    const real r2max = s->d.search_r2;
    uint32_t flags = 0, maxc = s->max_seen;
#pragma omp parallel for schedule(static) reduction(| : flags) reduction(max : maxc)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        uint32_t i = (uint32_t)ii;
        uint32_t key = s->key[i];
        uint32_t cx = key % gx, cy = (key / gx) % gy, cz = key / (gx * gy);
        uint32_t x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < gx ? cx + 1 : gx - 1;
        real px = s->P[4 * i], py = s->P[4 * i + 1], pz = s->P[4 * i + 2];
        uint32_t cnt = 0;
        for (int dz = -1; dz <= 1; dz++) {
            int64_t z = (int64_t)cz + dz;
            if (z < 0 || z >= (int64_t)gz) continue;
            for (int dy = -1; dy <= 1; dy++) {
                int64_t y = (int64_t)cy + dy;
                if (y < 0 || y >= (int64_t)gy) continue;
                uint32_t row = ((uint32_t)z * gy + (uint32_t)y) * gx;
                uint32_t j0 = s->off[row + x0], j1 = s->off[row + x1 + 1];
                for (uint32_t j = j0; j < j1; j++) {
                    if (j == i) continue;
                    real dx_ = px - s->P[4 * j], dy_ = py - s->P[4 * j + 1],
                         dz_ = pz - s->P[4 * j + 2];
                    real r2 = DIST2(dx_, dy_, dz_);
                    if (r2 < r2max) {
                        if (cnt < kmax) s->nbr[(size_t)i * kmax + cnt] = j;
                        cnt++;
                    }
                }
            }
        }
        if (cnt > kmax) flags |= PS_STATUS_NEIGHBOR_OVERFLOW;
        if (cnt > maxc) maxc = cnt;
        s->ncnt[i] = cnt < kmax ? cnt : kmax;
    }
    s->status_flags |= flags;
    s->max_seen = maxc;
    return PS_OK;
}

ORC_EXPORT int SFX(orc_begin)(orc_sys *s, real dt)
{
    SFX(orc_predict)(s, dt);
    SFX(orc_build_grid)(s);
    SFX(orc_find_neighbors)(s);
    s->stage = 1;
    return PS_OK;
}

/* ------------------------------------------------------------------------ */
/* a5: lambda (findLambdasD, integration_kernel.cuh:521-593; CPU
 * TotalFluidConstraint::project first loop, totalfluidconstraint.cpp:41-90).
 *   rho_i = W(0)/w_i + sum_j s_j W(r_ij)/w_j   (s_j = 1 fluid, s_solid else)
 *   grad_j C_i = a d,  a = (spiky / rho0_i) (h-r)^2 / r,   d = p_i - p_j
 *   lambda_i = -(rho_i/rho0_i - 1) / (sum_j |grad_j C_i|^2 + |sum_j grad_j C_i|^2 + eps)
 * Per pair, with t = h^2 - r^2, 1/r = RSQ(r^2) and q = h - r =
 * FMA(-r^2, 1/r, h): the density term t^3 m_j, the gradient term q^2 (1/r) d
 * and |grad_j C_i|^2 as q^4 (times sk^2); poly6 and sk = spiky/rho0_i
 * multiply the finished sums.  Degenerate pairs (r^2 < 1e-8) add density
 * only.  Static neighbours (w_j = 0) are skipped (cpu
 * totalfluidconstraint.cpp:122). */

/* The sums of particle i (fluid) at iterate P: density rho, S = sum_j
 * grad_j C_i and D = sum_j |grad_j C_i|^2 + |S|^2.  Two halves of the
 * table, [0] over its chunks of four entries 0, 2, 4, ... and [1] over
 * chunks 1, 3, 5, ..., each in table order, added as [0] + [1] (HALF(k)). */
static void lambda_sums(const orc_sys *s, const real *P, uint32_t i, real *rho_out, real *S, real *D)
{
    const orc_derived *d = &s->d;
    const uint32_t kmax = s->kmax;
    real px = P[4 * i], py = P[4 * i + 1], pz = P[4 * i + 2];
    real sk = d->spiky * (R(1) / s->rho0[i]);
    real rho[2] = {R(0), R(0)};
    real gix[2] = {R(0), R(0)}, giy[2] = {R(0), R(0)}, giz[2] = {R(0), R(0)}, den[2] = {R(0), R(0)};
    for (uint32_t k = 0; k < s->ncnt[i]; k++) {
        uint32_t j = s->nbr[(size_t)i * kmax + k];
        const int hf = HALF(k);
        real wj = s->w[j];
        if (wj == R(0)) continue;
        real dx = px - P[4 * j], dy = py - P[4 * j + 1], dz = pz - P[4 * j + 2];
        real r2 = DIST2(dx, dy, dz);
        if (!(r2 < d->h2)) continue;
        real t = d->h2 - r2;
        real m = R(1) / wj;
        if (s->phase[j] != PS_PHASE_FLUID) m = m * d->s_solid;
        rho[hf] = FMA((t * t) * t, m, rho[hf]);
        if (r2 < DEGENERATE_R2) continue;
        real ri = RSQ(r2);
        real q = FMA(-r2, ri, d->h);
        real q2 = q * q;
        real a = q2 * ri;
        den[hf] = FMA(q2, q2, den[hf]);
        gix[hf] = FMA(a, dx, gix[hf]);
        giy[hf] = FMA(a, dy, giy[hf]);
        giz[hf] = FMA(a, dz, giz[hf]);
    }
    real gx = sk * (gix[0] + gix[1]), gy = sk * (giy[0] + giy[1]), gz = sk * (giz[0] + giz[1]);
    *rho_out = FMA(d->poly6, rho[0] + rho[1], d->w_zero * (R(1) / s->w[i]));
    S[0] = gx;
    S[1] = gy;
    S[2] = gz;
    *D = FMA(sk * sk, den[0] + den[1], DIST2(gx, gy, gz));
}

static void lambda_pass(orc_sys *s)
{
    const orc_derived *d = &s->d;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ii++) {
        uint32_t i = (uint32_t)ii;
        if (s->phase[i] != PS_PHASE_FLUID) {
            s->P[4 * i + 3] = R(0);
            continue;
        }
        real rho, S[3], D;
        lambda_sums(s, s->P, i, &rho, S, &D);
        real C = (rho / s->rho0[i]) - R(1);
        s->P[4 * i + 3] = -C / (D + d->relax);
    }
}

/* x^n by square-and-multiply over the bits of n, low bit first. */
static real ipow(real x, uint32_t n)
{
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
}

/* a6: the fluid sum of particle i at iterate P (lambda in P.w)
 * (solveFluidsD, integration_kernel.cuh:596-642; cpu
 * totalfluidconstraint.cpp:92-115):
 *   f_i = sum_j (l_i + l_j + s_corr) gradW_ij,   (delta-p_i = f_i / rho0_i)
 *   gradW_ij = -spiky (h-r)^2 d/r,  s_corr = -k (W(r)/W(dq h))^n  (k: k h^2,
 *   or k with PS_FLAG_UNSCALED_TENSILE);
 *   r^2 < 1e-8: gradW = (0, -0.001 spiky h^2, 0)   (:625-626).
 * Per pair, with t = h^2 - r^2, 1/r = RSQ(r^2) and q = h - r =
 * FMA(-r^2, 1/r, h): W(r)/W(dq h) = s^3 with s = (h^2 - r^2)/(h^2 -
 * (dq h)^2) = FMA(-w_scale, r^2, w_scale_h2), coef = FMA(-k,
 * (W/W(dq h))^n, l_i + l_j), the term -coef q^2 (1/r) d (a degenerate
 * pair: -coef 0.001 h^2 in y); spiky multiplies the finished sum.  In two
 * halves of the table, as lambda_sums.  Static neighbours are skipped. */
static void fluid_sum(const orc_sys *s, const real *P, uint32_t i, real *f)
{
    const orc_derived *d = &s->d;
    const uint32_t kmax = s->kmax;
    real px = P[4 * i], py = P[4 * i + 1], pz = P[4 * i + 2];
    real li = P[4 * i + 3];
    real gdeg = (d->h * d->h) * R(0.001);
    real hx[2] = {R(0), R(0)}, hy[2] = {R(0), R(0)}, hz[2] = {R(0), R(0)};
    for (uint32_t k = 0; k < s->ncnt[i]; k++) {
        uint32_t j = s->nbr[(size_t)i * kmax + k];
        const int hf = HALF(k);
        if (s->w[j] == R(0)) continue;
        real dx = px - P[4 * j], dy = py - P[4 * j + 1], dz = pz - P[4 * j + 2];
        real r2 = DIST2(dx, dy, dz);
        if (!(r2 < d->h2)) continue;
        real sw = FMA(-d->w_scale, r2, d->w_scale_h2);
        real x = (sw * sw) * sw;
        real coef = FMA(-d->k_corr, ipow(x, d->n_corr), li + P[4 * j + 3]);
        if (r2 < DEGENERATE_R2) {
            hy[hf] = hy[hf] - (coef * gdeg);
        } else {
            real ri = RSQ(r2);
            real q = FMA(-r2, ri, d->h);
            real g = coef * ((q * q) * ri);
            hx[hf] = FMA(-g, dx, hx[hf]);
            hy[hf] = FMA(-g, dy, hy[hf]);
            hz[hf] = FMA(-g, dz, hz[hf]);
        }
    }
    f[0] = d->spiky * (hx[0] + hx[1]);
    f[1] = d->spiky * (hy[0] + hy[1]);
    f[2] = d->spiky * (hz[0] + hz[1]);
}

/* wall push-out + wall friction for one particle (collide_world_functor,
 * integration_kernel.cuh:57-157).  Order of the walls as there: floor,
 * +x, -x, ceiling, +z, -z. */
/* returns the coordinates it set (bits 0-2), bit 3 if friction acted, bit 4
 * for the kinetic branch with f < 1; when friction acts and rec is given,
 * rec = (position before friction, (n_x+1)*9 + (n_y+1)*3 + (n_z+1)) */
static uint32_t walls(const orc_sys *s, uint32_t i, real *p, uint64_t iter, real *rec)
{
    const orc_derived *d = &s->d;
    const real r = d->radius;
    const real bx0 = R(s->prm.box_min[0]), by0 = R(s->prm.box_min[1]), bz0 = R(s->prm.box_min[2]);
    const real bx1 = R(s->prm.box_max[0]), by1 = R(s->prm.box_max[1]), bz1 = R(s->prm.box_max[2]);
    const int32_t ph = s->phase[i];
    const uint32_t o = s->orig[i];
    const uint64_t seed = s->prm.seed;
    real nx = R(0), ny = R(0), nz = R(0);
    real efl = ph < PS_PHASE_SOLID ? d->jitter : R(0);
    uint32_t set = 0;
    if (p[1] < by0 + r) {
        p[1] = (by0 + r) + (wall_uniform(seed, iter, o, 5) * efl);
        ny = ny + R(1);
        set |= 2u;
    }
    if (p[0] > bx1 - r) {
        p[0] = bx1 - (r + (wall_uniform(seed, iter, o, 0) * d->jitter));
        nx = nx - R(1);
        set |= 1u;
    }
    if (p[0] < bx0 + r) {
        p[0] = bx0 + (r + (wall_uniform(seed, iter, o, 1) * d->jitter));
        nx = nx + R(1);
        set |= 1u;
    }
    if (p[1] > by1 - r) {
        p[1] = by1 - (r + (wall_uniform(seed, iter, o, 2) * d->jitter));
        ny = ny - R(1);
        set |= 2u;
    }
    if (s->prm.dims != 2) {
        if (p[2] > bz1 - r) {
            p[2] = bz1 - (r + (wall_uniform(seed, iter, o, 3) * d->jitter));
            nz = nz - R(1);
            set |= 4u;
        }
        if (p[2] < bz0 + r) {
            p[2] = bz0 + (r + (wall_uniform(seed, iter, o, 4) * d->jitter));
            nz = nz + R(1);
            set |= 4u;
        }
    }
    if (ph < PS_PHASE_CLOTH) return set;
    if (nx == R(0) && ny == R(0) && nz == R(0)) return set;
    real ex = p[0] - s->X[3 * i], ey = p[1] - s->X[3 * i + 1], ez = p[2] - s->X[3 * i + 2];
    real dn = ((ex * nx) + (ey * ny)) + (ez * nz);
    real tx = ex - (dn * nx), ty = ey - (dn * ny), tz = ez - (dn * nz);
    real l = SQRT(((tx * tx) + (ty * ty)) + (tz * tz));
    if (l < R(0.001)) return set;
    set |= 8u;
    if (rec) {
        rec[0] = p[0];
        rec[1] = p[1];
        rec[2] = p[2];
        rec[3] = (real)((((int)nx + 1) * 9 + ((int)ny + 1) * 3) + ((int)nz + 1));
    }
    if (l < d->wall_s) {
        p[0] = p[0] - tx;
        p[1] = p[1] - ty;
        p[2] = p[2] - tz;
    } else {
        real f = d->wall_k / l;
        if (f > R(1))
            f = R(1);
        else
            set |= 16u;
        p[0] = p[0] - (tx * f);
        p[1] = p[1] - (ty * f);
        p[2] = p[2] - (tz * f);
    }
    return set;
}

/* ------------------------------------------------------------------------ */
/* a12 shape matching (3D Mueller et al. 2005 meshless shape matching; the
 * reference's CPU form is 2D, cpu/src/solver/particle.cpp:15-57 and
 * totalshapeconstraint.cpp:14-24,80-87): c = sum m p / sum m,
 * A = sum m (p - c) q^T, R = rotation of A by the iterative quaternion
 * method (10 sweeps, warm-started from the previous R), goal g = R q + c. */

static void quat_to_mat(const real *q, real *m)
{
    real w = q[0], x = q[1], y = q[2], z = q[3];
    m[0] = R(1) - (R(2) * ((y * y) + (z * z)));
    m[1] = R(2) * ((x * y) - (w * z));
    m[2] = R(2) * ((x * z) + (w * y));
    m[3] = R(2) * ((x * y) + (w * z));
    m[4] = R(1) - (R(2) * ((x * x) + (z * z)));
    m[5] = R(2) * ((y * z) - (w * x));
    m[6] = R(2) * ((x * z) - (w * y));
    m[7] = R(2) * ((y * z) + (w * x));
    m[8] = R(1) - (R(2) * ((x * x) + (y * y)));
}

/* Rotation part of A (row-major 3x3) by the fixed-point iteration of
 * Mueller, Bender, Chentanez and Macklin, "A Robust Method to Extract the
 * Rotational Part of Deformations" (MIG 2016):
 * omega = sum_k r_k x a_k / (|sum_k r_k . a_k| + 1e-9), r_k/a_k the columns
 * of R(q)/A, and q <- normalize((1, omega/2) q).  The first-order
 * quaternion step replaces exp(omega) (no trigonometry); the fixed point,
 * omega = 0, is the same: the rotation factor of the polar decomposition. */
static void extract_rotation(const real *A, real *q, int sweeps)
{
    for (int it = 0; it < sweeps; it++) {
        real m[9];
        quat_to_mat(q, m);
        real ox = R(0), oy = R(0), oz = R(0), dot = R(0);
        for (int c = 0; c < 3; c++) {
            real rx = m[c], ry = m[3 + c], rz = m[6 + c];
            real ax = A[c], ay = A[3 + c], az = A[6 + c];
            ox = ox + ((ry * az) - (rz * ay));
            oy = oy + ((rz * ax) - (rx * az));
            oz = oz + ((rx * ay) - (ry * ax));
            dot = dot + (((rx * ax) + (ry * ay)) + (rz * az));
        }
        real den = (dot < R(0) ? -dot : dot) + R(1e-9);
        ox = ox / den;
        oy = oy / den;
        oz = oz / den;
        real wlen2 = ((ox * ox) + (oy * oy)) + (oz * oz);
        if (wlen2 < R(1e-18)) break;
        real dw = R(1), dx = ox * R(0.5), dy = oy * R(0.5), dz = oz * R(0.5);
        real w = q[0], x = q[1], y = q[2], z = q[3];
        real nw = (((dw * w) - (dx * x)) - (dy * y)) - (dz * z);
        real nx = (((dw * x) + (dx * w)) + (dy * z)) - (dz * y);
        real ny = (((dw * y) - (dx * z)) + (dy * w)) + (dz * x);
        real nz = (((dw * z) + (dx * y)) - (dy * x)) + (dz * w);
        real nl = SQRT((((nw * nw) + (nx * nx)) + (ny * ny)) + (nz * nz));
        q[0] = nw / nl;
        q[1] = nx / nl;
        q[2] = ny / nl;
        q[3] = nz / nl;
    }
}

static void body_init(orc_sys *s, uint32_t b)
{
    uint32_t a = s->body_off[b], e = s->body_off[b + 1];
    real cx = R(0), cy = R(0), cz = R(0), M = R(0);
    for (uint32_t k = a; k < e; k++) {
        uint32_t t = s->slot_of[s->body_idx[k]];
        real m = R(1) / s->w[t];
        cx = cx + (m * s->X[3 * t]);
        cy = cy + (m * s->X[3 * t + 1]);
        cz = cz + (m * s->X[3 * t + 2]);
        M = M + m;
    }
    cx = cx / M;
    cy = cy / M;
    cz = cz / M;
    for (uint32_t k = a; k < e; k++) {
        uint32_t t = s->slot_of[s->body_idx[k]];
        s->body_rest[3 * k] = s->X[3 * t] - cx;
        s->body_rest[3 * k + 1] = s->X[3 * t + 1] - cy;
        s->body_rest[3 * k + 2] = s->X[3 * t + 2] - cz;
    }
    s->body_mass[b] = M;
    s->body_q[4 * b] = R(1);
    s->body_q[4 * b + 1] = R(0);
    s->body_q[4 * b + 2] = R(0);
    s->body_q[4 * b + 3] = R(0);
}

/* goals of every body from the current iterate */
/* A body's sums over its members in the order of a 64-lane wavefront:
 * member k (0-based within the body) goes to lane k % 64, each lane sums
 * its members in order, then the lanes are added pairwise, v[l] + v[l + s]
 * for s = 32, 16, ..., 1.  lane[64][w]: the per-lane partial sums of w
 * quantities; the totals go to out[w]. */
static void lane_tree_sum(real (*lane)[9], int w, real *out)
{
    for (uint32_t st = 32; st > 0; st >>= 1)
        for (uint32_t l = 0; l < st; l++)
            for (int e = 0; e < w; e++) lane[l][e] = lane[l][e] + lane[l + st][e];
    for (int e = 0; e < w; e++) out[e] = lane[0][e];
}

static void shape_pass(orc_sys *s)
{
#pragma omp parallel for schedule(static)
    for (int64_t bb = 0; bb < (int64_t)s->nbody; bb++) {
        uint32_t b = (uint32_t)bb, a = s->body_off[b], e = s->body_off[b + 1];
        real lane[64][9];
        memset(lane, 0, sizeof(lane));
        for (uint32_t k = a; k < e; k++) {
            uint32_t t = s->slot_of[s->body_idx[k]];
            real m = R(1) / s->w[t];
            real *v = lane[(k - a) % 64];
            v[0] = v[0] + (m * s->P[4 * t]);
            v[1] = v[1] + (m * s->P[4 * t + 1]);
            v[2] = v[2] + (m * s->P[4 * t + 2]);
        }
        real c[3];
        lane_tree_sum(lane, 3, c);
        real cx = c[0] / s->body_mass[b], cy = c[1] / s->body_mass[b], cz = c[2] / s->body_mass[b];
        memset(lane, 0, sizeof(lane));
        for (uint32_t k = a; k < e; k++) {
            uint32_t t = s->slot_of[s->body_idx[k]];
            real m = R(1) / s->w[t];
            real px = m * (s->P[4 * t] - cx), py = m * (s->P[4 * t + 1] - cy),
                 pz = m * (s->P[4 * t + 2] - cz);
            real qx = s->body_rest[3 * k], qy = s->body_rest[3 * k + 1], qz = s->body_rest[3 * k + 2];
            real *A = lane[(k - a) % 64];
            A[0] = A[0] + (px * qx); A[1] = A[1] + (px * qy); A[2] = A[2] + (px * qz);
            A[3] = A[3] + (py * qx); A[4] = A[4] + (py * qy); A[5] = A[5] + (py * qz);
            A[6] = A[6] + (pz * qx); A[7] = A[7] + (pz * qy); A[8] = A[8] + (pz * qz);
        }
        real A[9];
        lane_tree_sum(lane, 9, A);
        extract_rotation(A, &s->body_q[4 * b], 10);
        real m[9];
        quat_to_mat(&s->body_q[4 * b], m);
        for (uint32_t k = a; k < e; k++) {
            real qx = s->body_rest[3 * k], qy = s->body_rest[3 * k + 1], qz = s->body_rest[3 * k + 2];
            s->body_goal[3 * k] = (((m[0] * qx) + (m[1] * qy)) + (m[2] * qz)) + cx;
            s->body_goal[3 * k + 1] = (((m[3] * qx) + (m[4] * qy)) + (m[5] * qz)) + cy;
            s->body_goal[3 * k + 2] = (((m[6] * qx) + (m[7] * qy)) + (m[8] * qz)) + cz;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* One Jacobi iteration.  Every correction reads the same snapshot P; the
 * per-particle sum is formed in a fixed order (contacts, fluid, distance,
 * shape), then the walls (collideWorld) and pins (solvePointConstraints)
 * are applied to the result, as in particlesystem.cpp:184-234.            */

static void solve_pass(orc_sys *s)
{
    const orc_derived *d = &s->d;
    const uint32_t kmax = s->kmax;
    const uint64_t iter = s->iters;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ii++) {
        uint32_t i = (uint32_t)ii;
        real px = s->P[4 * i], py = s->P[4 * i + 1], pz = s->P[4 * i + 2];
        int32_t ph = s->phase[i];
        real wi = s->w[i];
        real ax = R(0), ay = R(0), az = R(0);

        /* a7: particle contacts + friction (collideD,
         * integration_kernel.cuh:358-462), for phases >= CLOTH, Jacobi-
         * averaged over the contact count; same-body RIGID pairs skip
         * (:336); mass scaling w* = w e^{y} for solid pairs, each particle
         * with its own height (:403,425); friction against the relative
         * displacement since the substep start (:439-457, using p_j for
         * the second term). */
        if (ph >= PS_PHASE_CLOTH && wi != R(0)) {
            uint32_t nc = 0;
            for (uint32_t k = 0; k < s->ncnt[i]; k++) {
                uint32_t j = s->nbr[(size_t)i * kmax + k];
                int32_t pj = s->phase[j];
                if (ph > PS_PHASE_SOLID && ph == pj) continue;
                real dx = px - s->P[4 * j], dy = py - s->P[4 * j + 1], dz = pz - s->P[4 * j + 2];
                real r2 = ((dx * dx) + (dy * dy)) + (dz * dz);
                if (r2 < d->collide_r2) nc++;
            }
            if (nc > 0) {
                real fnc = (real)nc;
                real ewi = wi * EXP(py);
                /* the corrections are summed in the two halves of the table
                 * (HALF), each in table order, then added half 0 + half 1,
                 * as the fluid sums are */
                real ch[2][3] = {{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
                for (uint32_t k = 0; k < s->ncnt[i]; k++) {
                    real *c3 = ch[HALF(k)];
                    uint32_t j = s->nbr[(size_t)i * kmax + k];
                    int32_t pj = s->phase[j];
                    if (ph > PS_PHASE_SOLID && ph == pj) continue;
                    real qx = s->P[4 * j], qy = s->P[4 * j + 1], qz = s->P[4 * j + 2];
                    real dx = px - qx, dy = py - qy, dz = pz - qz;
                    real r2 = ((dx * dx) + (dy * dy)) + (dz * dz);
                    if (!(r2 < d->collide_r2)) continue;
                    real dist = SQRT(r2);
                    if (dist < R(1e-6)) continue;
                    real wj = s->w[j];
                    real cw = wi, cw2 = wj;
                    int solid = ph >= PS_PHASE_SOLID && pj >= PS_PHASE_SOLID;
                    if (solid) {
                        cw = ewi;
                        cw2 = wj * EXP(qy);
                    }
                    real wsum = cw + cw2;
                    if (wsum == R(0)) continue;
                    real mag = dist - d->collide_dist;
                    /* dp1 = -cw * (mag / wsum) * d / dist / nc */
                    real sc = (((mag / wsum) / dist) * cw) / fnc;
                    real ex = -(sc * dx), ey = -(sc * dy), ez = -(sc * dz);
                    c3[0] = c3[0] + ex;
                    c3[1] = c3[1] + ey;
                    c3[2] = c3[2] + ez;
                    if (!solid) continue;
                    real sc2 = (((mag / wsum) / dist) * cw2) / fnc;
                    real fx = sc2 * dx, fy = sc2 * dy, fz = sc2 * dz;
                    /* relative displacement over the substep, after the
                     * normal corrections: (p_i + dp1 - x*_i) - (p_j + dp2 - x*_j) */
                    real rx = ((px + ex) - s->X[3 * i]) - ((qx + fx) - s->X[3 * j]);
                    real ry = ((py + ey) - s->X[3 * i + 1]) - ((qy + fy) - s->X[3 * j + 1]);
                    real rz = ((pz + ez) - s->X[3 * i + 2]) - ((qz + fz) - s->X[3 * j + 2]);
                    real nx = dx / dist, ny = dy / dist, nz = dz / dist;
                    real dn = ((rx * nx) + (ry * ny)) + (rz * nz);
                    real tx = rx - (dn * nx), ty = ry - (dn * ny), tz = rz - (dn * nz);
                    real lt = SQRT(((tx * tx) + (ty * ty)) + (tz * tz));
                    if (lt < R(0.001)) continue;
                    real share = cw / wsum;
                    if (lt < (d->mu_s * dist)) {
                        c3[0] = c3[0] - (tx * share);
                        c3[1] = c3[1] - (ty * share);
                        c3[2] = c3[2] - (tz * share);
                    } else {
                        real f = (d->mu_k * dist) / lt;
                        if (f > R(1)) f = R(1);
                        c3[0] = c3[0] - ((tx * f) * share);
                        c3[1] = c3[1] - ((ty * f) * share);
                        c3[2] = c3[2] - ((tz * f) * share);
                    }
                }
                real cx = ch[0][0] + ch[1][0], cy = ch[0][1] + ch[1][1], cz = ch[0][2] + ch[1][2];
                ax = ax + cx;
                ay = ay + cy;
                az = az + cz;
            }
        }

        /* a6: delta-p with tensile correction: f_i / rho0_i (fluid_sum) */
        if (ph == PS_PHASE_FLUID) {
            real f[3];
            fluid_sum(s, s->P, i, f);
            real ir0 = R(1) / s->rho0[i];
            ax = ax + (f[0] * ir0);
            ay = ay + (f[1] * ir0);
            az = az + (f[2] * ir0);
        }

        /* a9: distance constraints, mass-weighted with stiffness and
         * Jacobi-averaged by the incidence count (CPU form,
         * cpu/src/constraint/distanceconstraint.cpp:20-40; the GPU form
         * solver.cu:196-231 sums equal splits then divides by occurrences),
         * gathered per particle over its incident constraints. */
        {
            uint32_t o = s->orig[i];
            uint32_t a0 = s->dist_off[o], a1 = s->dist_off[o + 1];
            if (a1 > a0 && wi != R(0)) {
                real cx = R(0), cy = R(0), cz = R(0);
                for (uint32_t k = a0; k < a1; k++) {
                    uint32_t c = s->dist_adj[k];
                    uint32_t oa = s->dist_pairs[2 * c], ob = s->dist_pairs[2 * c + 1];
                    uint32_t other = oa == o ? ob : oa;
                    uint32_t j = s->slot_of[other];
                    real dx = px - s->P[4 * j], dy = py - s->P[4 * j + 1], dz = pz - s->P[4 * j + 2];
                    real dist = SQRT(((dx * dx) + (dy * dy)) + (dz * dz));
                    real wsum = wi + s->w[j];
                    if (dist < R(1e-6) || wsum == R(0)) continue;
                    real sc = ((((dist - s->dist_rest[c]) / wsum) / dist) * wi) * s->dist_stiff[c];
                    cx = cx - (sc * dx);
                    cy = cy - (sc * dy);
                    cz = cz - (sc * dz);
                }
                real occ = (real)s->occ[o];
                ax = ax + (cx / occ);
                ay = ay + (cy / occ);
                az = az + (cz / occ);
            }
        }

        /* a12: shape matching goal pull (totalshapeconstraint.cpp:19-23) */
        {
            uint32_t o = s->orig[i];
            uint32_t b = s->body_of[o];
            if (b) {
                uint32_t k = s->member_of[o];
                real st = s->body_stiff[b - 1];
                ax = ax + ((s->body_goal[3 * k] - px) * st);
                ay = ay + ((s->body_goal[3 * k + 1] - py) * st);
                az = az + ((s->body_goal[3 * k + 2] - pz) * st);
            }
        }

        real p[3] = {px + ax, py + ay, pz + az};
        if (s->prm.dims == 2) p[2] = pz;
        /* a8: walls */
        uint32_t set = walls(s, i, p, iter, s->clamp_rec ? s->wall_rec + 4 * (size_t)i : NULL);
        /* a10: pins (point_constraint_functor, solver_kernel.cuh:10-25) */
        {
            uint32_t pin = s->pin_of[s->orig[i]];
            if (pin) {
                p[0] = s->pin_xyz[3 * (pin - 1)];
                p[1] = s->pin_xyz[3 * (pin - 1) + 1];
                p[2] = s->pin_xyz[3 * (pin - 1) + 2];
                set |= 39u; /* all three coordinates, and bit 5: pinned */
            }
        }
        if (s->clamp_rec) s->clamp_rec[i] = (uint8_t)set;
        s->Q[4 * i] = p[0];
        s->Q[4 * i + 1] = p[1];
        s->Q[4 * i + 2] = p[2];
        s->Q[4 * i + 3] = R(0);
    }
    real *t = s->P;
    s->P = s->Q;
    s->Q = t;
}

/* lambda pass alone (pinning tests read it from the iterate's 4th lane) */
ORC_EXPORT int SFX(orc_compute_lambda)(orc_sys *s)
{
    if (s->stage != 1) return PS_ERR_STATE;
    lambda_pass(s);
    return PS_OK;
}

ORC_EXPORT int SFX(orc_iterate)(orc_sys *s)
{
    if (s->stage != 1) return PS_ERR_STATE;
    if (s->prm.flags & PS_FLAG_REBUILD_EVERY_ITERATION) {
        SFX(orc_build_grid)(s);
        SFX(orc_find_neighbors)(s);
    }
    lambda_pass(s);
    if (s->nbody) shape_pass(s);
    solve_pass(s);
    s->iters++;
    return PS_OK;
}

/* a11: v = (x - x*)/dt (calcVelocity / subtract_functor,
 * integration.cu:416-426, integration_kernel.cuh:465-475) plus XSPH for
 * fluids, v_i += (c / rho0_i) sum_{fluid j} (v_j - v_i) m_j W(r_ij), formed
 * from displacements u = x - x*:  v_i = u_i/dt + ((c/rho0_i) S)/dt with
 * S = sum_j (u_j - u_i) m_j W_ij; then x <- x_final (optionally sleeping,
 * cpu/src/particle.h:60-65).                                               */

/* S of fluid particle i at iterate P with start positions X (3 per
 * particle): per pair (u_j - u_i) m_j t^3, t = h^2 - r^2, in four quarters
 * of the table (QUARTER: chunks of four entries in turn, each in table
 * order), added as ([0] + [2]) + ([1] + [3]); poly6 multiplies the
 * finished sum. */
static void xsph_sum(const orc_sys *s, const real *P, const real *X, uint32_t i, real *S)
{
    const orc_derived *d = &s->d;
    const uint32_t kmax = s->kmax;
    real px = P[4 * i], py = P[4 * i + 1], pz = P[4 * i + 2];
    real ux = px - X[3 * i], uy = py - X[3 * i + 1], uz = pz - X[3 * i + 2];
    real hx[4] = {R(0), R(0), R(0), R(0)}, hy[4] = {R(0), R(0), R(0), R(0)}, hz[4] = {R(0), R(0), R(0), R(0)};
    for (uint32_t k = 0; k < s->ncnt[i]; k++) {
        uint32_t j = s->nbr[(size_t)i * kmax + k];
        const int hf = QUARTER(k);
        if (s->phase[j] != PS_PHASE_FLUID) continue;
        real qx = P[4 * j], qy = P[4 * j + 1], qz = P[4 * j + 2];
        real dx = px - qx, dy = py - qy, dz = pz - qz;
        real r2 = DIST2(dx, dy, dz);
        if (!(r2 < d->h2)) continue;
        real t = d->h2 - r2;
        real m = ((t * t) * t) * (R(1) / s->w[j]);
        real ex = qx - X[3 * j], ey = qy - X[3 * j + 1], ez = qz - X[3 * j + 2];
        hx[hf] = FMA(ex - ux, m, hx[hf]);
        hy[hf] = FMA(ey - uy, m, hy[hf]);
        hz[hf] = FMA(ez - uz, m, hz[hf]);
    }
    S[0] = d->poly6 * ((hx[0] + hx[2]) + (hx[1] + hx[3]));
    S[1] = d->poly6 * ((hy[0] + hy[2]) + (hy[1] + hy[3]));
    S[2] = d->poly6 * ((hz[0] + hz[2]) + (hz[1] + hz[3]));
}

ORC_EXPORT int SFX(orc_end)(orc_sys *s)
{
    if (s->stage != 1) return PS_ERR_STATE;
    const orc_derived *d = &s->d;
    const real dt = s->dt;
    real *Vn = s->tmpr; /* 3*n */
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->n; ii++) {
        uint32_t i = (uint32_t)ii;
        real px = s->P[4 * i], py = s->P[4 * i + 1], pz = s->P[4 * i + 2];
        real ux = px - s->X[3 * i], uy = py - s->X[3 * i + 1], uz = pz - s->X[3 * i + 2];
        real vx = ux / dt, vy = uy / dt, vz = uz / dt;
        if (s->phase[i] == PS_PHASE_FLUID && d->xsph_c != R(0)) {
            real S[3];
            xsph_sum(s, s->P, s->X, i, S);
            real cr = d->xsph_c / s->rho0[i];
            vx = vx + ((cr * S[0]) / dt);
            vy = vy + ((cr * S[1]) / dt);
            vz = vz + ((cr * S[2]) / dt);
        }
        Vn[3 * i] = vx;
        Vn[3 * i + 1] = vy;
        Vn[3 * i + 2] = vz;
    }
    const int sleeping = (s->prm.flags & PS_FLAG_SLEEPING) != 0;
    for (uint32_t i = 0; i < s->n; i++) {
        real px = s->P[4 * i], py = s->P[4 * i + 1], pz = s->P[4 * i + 2];
        if (sleeping) {
            real ex = px - s->X[3 * i], ey = py - s->X[3 * i + 1], ez = pz - s->X[3 * i + 2];
            real l = SQRT(((ex * ex) + (ey * ey)) + (ez * ez));
            if (l < R(1e-4)) {
                s->V[3 * i] = R(0);
                s->V[3 * i + 1] = R(0);
                s->V[3 * i + 2] = R(0);
                continue;
            }
        }
        s->V[3 * i] = Vn[3 * i];
        s->V[3 * i + 1] = Vn[3 * i + 1];
        s->V[3 * i + 2] = Vn[3 * i + 2];
        s->X[3 * i] = px;
        s->X[3 * i + 1] = py;
        s->X[3 * i + 2] = pz;
    }
    s->substeps++;
    s->stage = 0;
    return PS_OK;
}

ORC_EXPORT int SFX(orc_step)(orc_sys *s, real dt, uint32_t substeps)
{
    for (uint32_t k = 0; k < substeps; k++) {
        SFX(orc_begin)(s, dt);
        for (uint32_t it = 0; it < s->prm.iterations; it++) SFX(orc_iterate)(s);
        SFX(orc_end)(s);
    }
    return PS_OK;
}

/* ------------------------------------------------------------------------ */
/* readback                                                                 */

ORC_EXPORT uint32_t SFX(orc_num_particles)(const orc_sys *s) { return s->n; }

ORC_EXPORT int SFX(orc_read_positions)(const orc_sys *s, real *xyz)
{
    for (uint32_t t = 0; t < s->n; t++) {
        uint32_t o = s->orig[t];
        xyz[3 * o] = s->X[3 * t];
        xyz[3 * o + 1] = s->X[3 * t + 1];
        xyz[3 * o + 2] = s->X[3 * t + 2];
    }
    return PS_OK;
}

ORC_EXPORT int SFX(orc_read_velocities)(const orc_sys *s, real *xyz)
{
    for (uint32_t t = 0; t < s->n; t++) {
        uint32_t o = s->orig[t];
        xyz[3 * o] = s->V[3 * t];
        xyz[3 * o + 1] = s->V[3 * t + 1];
        xyz[3 * o + 2] = s->V[3 * t + 2];
    }
    return PS_OK;
}

ORC_EXPORT int SFX(orc_read_grid)(const orc_sys *s, uint32_t *sorted_index, uint32_t *cell,
                                  uint32_t *off)
{
    if (sorted_index) memcpy(sorted_index, s->orig, sizeof(uint32_t) * s->n);
    if (cell) memcpy(cell, s->key, sizeof(uint32_t) * s->n);
    if (off) memcpy(off, s->off, sizeof(uint32_t) * ((size_t)s->ncell + 1));
    return PS_OK;
}

ORC_EXPORT int SFX(orc_read_neighbors)(const orc_sys *s, uint32_t *counts, uint32_t *table)
{
    if (counts) memcpy(counts, s->ncnt, sizeof(uint32_t) * s->n);
    if (table) memcpy(table, s->nbr, sizeof(uint32_t) * (size_t)s->n * s->kmax);
    return PS_OK;
}

ORC_EXPORT int SFX(orc_read_sorted_state)(const orc_sys *s, real *iterate, real *start)
{
    if (iterate) memcpy(iterate, s->P, sizeof(real) * 4 * s->n);
    if (start) memcpy(start, s->X, sizeof(real) * 3 * s->n);
    return PS_OK;
}

ORC_EXPORT int SFX(orc_get_status)(const orc_sys *s, ps_status *out)
{
    memset(out, 0, sizeof(*out));
    out->flags = s->status_flags;
    out->max_neighbors = s->max_seen;
    out->num_particles = s->n;
    out->num_cells = s->ncell;
    out->substeps = s->substeps;
    out->iterations = s->iters;
    return PS_OK;
}

/* Shape-matching rotation of a 3x3 matrix (row-major), for pinning the
 * polar decomposition against an FP64 SVD: q in/out (w,x,y,z). */
ORC_EXPORT void SFX(orc_extract_rotation)(const real *A, real *q, int sweeps)
{
    extract_rotation(A, q, sweeps);
}

/* ------------------------------------------------------------------------ */
/* Full-precision setters for finite-difference checks (insertion order; a
 * NULL array is left as it is). */
ORC_EXPORT int SFX(orc_set_state)(orc_sys *s, const real *x, const real *v)
{
    for (uint32_t t = 0; t < s->n; t++) {
        uint32_t o = s->orig[t];
        for (int a = 0; a < 3; a++) {
            if (x) s->X[3 * t + a] = x[3 * o + a];
            if (v) s->V[3 * t + a] = v[3 * o + a];
        }
    }
    return PS_OK;
}

/* wall friction coefficients in full precision: wall_k = sqrt(mu_k) radius */
ORC_EXPORT int SFX(orc_set_wall_k)(orc_sys *s, real wall_k)
{
    s->d.wall_k = wall_k;
    return PS_OK;
}

ORC_EXPORT int SFX(orc_set_coefficients)(orc_sys *s, const real *rho0, const real *stiffness)
{
    if (rho0)
        for (uint32_t t = 0; t < s->n; t++) s->rho0[t] = rho0[s->orig[t]];
    if (stiffness)
        for (uint32_t c = 0; c < s->ndist; c++) s->dist_stiff[c] = stiffness[c];
    return PS_OK;
}

#include "ps_adjoint.inc"
