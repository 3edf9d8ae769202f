// ps_cpuref.hip — kernels of the CPU-reference mode (PS_FLAG_CPU_REFERENCE).
//
// This mode evolves a scene with the reference's own CPU solver semantics,
// Simulation::tick (cpu/src/simulation.cpp:115-369) with its ITERATIVE path:
//   - prediction v += dt g, ep = p + dt v (static particles keep p)
//     (:139-161, Particle::guess cpu/src/particle.h:56-58);
//   - box boundaries become constraints at prediction time (:201-224) and
//     count towards the particle's Jacobi divisor (:232-245);
//   - each iteration projects the CONTACT group (BoundaryConstraint::project,
//     cpu/src/constraint/boundaryconstraint.cpp:14-93) then the STANDARD
//     group: one TotalFluidConstraint per fluid object and one GasConstraint
//     per gas object, in creation order
//     (cpu/src/constraint/totalfluidconstraint.cpp:41-116,
//     cpu/src/constraint/gasconstraint.cpp:31-114);
//   - gas objects add the open-gas drag and a vorticity term to a
//     per-particle force, which the next prediction applies together with
//     gravity scaled by ALPHA for gas particles (simulation.cpp:143-154);
//   - v = (ep - p)/dt, sleeping below |ep - p| < 1e-4 (:322-333,
//     particle.h:60-65).
// The fluid object scans every particle for neighbours (O(N^2), as the
// reference does) and accumulates in ascending particle index, so its sums
// are the reference's sums term for term.
//
// Storage is insertion order.  X: positions (.w unused), E/E2: iterate
// ping-pong (x, y, z, inverse mass), V: velocities, lam: lambda per
// particle, bc: boundary flags (bits 0-5) | boundary count << 8.
//
// Work shape: one wavefront per particle for the two fluid passes.  The 64
// lanes test 64 candidates at once; the neighbours found are ballot-masked
// and their terms are added one after another in lane (= index) order with
// readlane, which keeps the sequential summation order of the reference.
// Boundaries and the Jacobi update are particle-local, so the boundary pass
// of iteration k+1 is fused into the last fluid pass of iteration k (and the
// one of iteration 0 into prediction).

#include "ps_device.hpp"
#include "ps_internal.hpp"

namespace ps {

namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kWaves = kBlock / 64;

enum : uint32_t { BX_LOW = 1u, BX_HIGH = 2u, BY_LOW = 4u, BY_HIGH = 8u, BZ_LOW = 16u, BZ_HIGH = 32u };
// bc bit 6: the particle is neither fluid nor gas, so its boundary
// projections add no jitter (boundaryconstraint.cpp:18-19)
constexpr uint32_t kNoJitter = 64u;

// GasConstraint's compile-time constants (gasconstraint.h:10-19) and the gas
// gravity scale ALPHA (simulation.h:21), rounded once to float
constexpr float kGasRelax = 0.01f, kGasKP = 0.2f, kGasAlpha = -0.2f, kGasOpenDrag = -50.0f;
constexpr uint32_t kGasEP = 4;

__device__ __forceinline__ uint64_t cr_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Boundary jitter in [0,1) (replaces frand(), boundaryconstraint.cpp:19):
// SplitMix64 finaliser of seed + golden * counter + 1, counter = (iteration,
// insertion index, axis), top 24 bits.
__device__ __forceinline__ float cr_uniform(uint64_t seed, uint64_t iter, uint32_t i, uint32_t axis)
{
    uint64_t ctr = (iter << 35) | ((uint64_t)i << 3) | (uint64_t)axis;
    uint64_t z = cr_mix64(seed + ctr * 0x9E3779B97F4A7C15ull + 1ull);
    return (float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f);
}

// poly6 (totalfluidconstraint.cpp:123-129): 0 at r2 >= h2
__device__ __forceinline__ float cr_poly6(const CrConst &c, float r2)
{
    if (r2 >= c.h2) return 0.0f;
    const float t = c.h2 - r2;
    return c.poly6 * ((t * t) * t);
}

// spikyGrad (totalfluidconstraint.cpp:131-137): -normalize(r) c (h-l)^2,
// zero at l >= h or l == 0; normalize(r) = r * (1 / sqrt(r.r))
__device__ __forceinline__ void cr_spiky(const CrConst &c, float rx, float ry, float rz, float r2,
                                         float l, float &gx, float &gy, float &gz)
{
    const float inv = 1.0f / sqrtf(r2);
    const float q = c.h - l;
    const bool on = !(l >= c.h || l == 0.0f);
    gx = on ? (((-(rx * inv)) * c.spiky) * q) * q : 0.0f;
    gy = on ? (((-(ry * inv)) * c.spiky) * q) * q : 0.0f;
    gz = on ? (((-(rz * inv)) * c.spiky) * q) * q : 0.0f;
}

// x^n by binary powering (x^4 = (x^2)^2)
__device__ __forceinline__ float cr_ipow(float x, uint32_t n)
{
    float r = 1.0f, b = x;
    bool first = true;
    while (n) {
        if (n & 1u) {
            r = first ? b : (r * b);
            first = false;
        }
        n >>= 1;
        if (n) b = b * b;
    }
    return r;
}

__device__ __forceinline__ float lane_value(float v, uint32_t lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)lane));
}

// BoundaryConstraint::project for one axis (boundaryconstraint.cpp:14-93):
// push to value +- (r + jitter), then friction on the tangential part of
// the displacement since the substep start divided by the boundary count.
__device__ __forceinline__ void cr_boundary_axis(const CrConst &c, float e[3], const float p[3],
                                                 uint32_t axis, bool greater, float cnt,
                                                 float extra)
{
    const float r = c.radius;
    const float value = greater ? c.box_min[axis] : c.box_max[axis];
    const float d = r + extra;
    float nv[3] = {0.0f, 0.0f, 0.0f};
    if (greater) {
        if (e[axis] >= value + r) return;
        e[axis] = value + d;
        nv[axis] = 1.0f;
    } else {
        if (e[axis] <= value - r) return;
        e[axis] = value - d;
        nv[axis] = -1.0f;
    }
    float dp[3], t[3];
    for (int a = 0; a < 3; a++) dp[a] = (e[a] - p[a]) / cnt;
    const float dn = ((dp[0] * nv[0]) + (dp[1] * nv[1])) + (dp[2] * nv[2]);
    for (int a = 0; a < 3; a++) t[a] = dp[a] - (dn * nv[a]);
    const float l = sqrtf(((t[0] * t[0]) + (t[1] * t[1])) + (t[2] * t[2]));
    if (l < 1e-4f) return;
    if (l < c.sf * d) {
        for (int a = 0; a < 3; a++) e[a] = e[a] - t[a];
    } else {
        float f = (c.kf * d) / l;
        f = f > 1.0f ? 1.0f : f;
        for (int a = 0; a < 3; a++) e[a] = e[a] - (t[a] * f);
    }
}

// All boundary constraints of particle i, in the order the reference creates
// them (x then y then z, simulation.cpp:201-224).  Fluids get the jitter.
__device__ __forceinline__ void cr_boundaries(const CrConst &c, float e[3], const float p[3],
                                              uint32_t bc, uint64_t iter, uint32_t i)
{
    if (!(bc & 0x3Fu)) return;
    const float cnt = (float)(bc >> 8);
    for (uint32_t axis = 0; axis < 3; axis++) {
        const uint32_t lo = 1u << (2 * axis), hi = 2u << (2 * axis);
        if (!(bc & (lo | hi))) continue;
        const float extra = (bc & kNoJitter) ? 0.0f : cr_uniform(c.seed, iter, i, axis) * c.jitter;
        cr_boundary_axis(c, e, p, axis, (bc & lo) != 0, cnt, extra);
    }
}

// (1)-(9): velocity and predicted position, boundary flags and counts, then
// the boundary projections of iteration 0.
__global__ __launch_bounds__(kBlock) void cr_predict_k(const float4 *__restrict__ X,
                                                       float4 *__restrict__ V,
                                                       const Attr *__restrict__ A,
                                                       float4 *__restrict__ E,
                                                       uint32_t *__restrict__ bcs,
                                                       float4 *__restrict__ F,
                                                       float *__restrict__ tm,
                                                       const DevStatus *__restrict__ st, CrConst c,
                                                       uint32_t n, float dt, int boundaries)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 x = X[i];
    float4 v = V[i];
    const Attr at = A[i];
    const float w = at.w;
    const bool two = c.dims == 2;
    // v = v + dt g' + dt f (g' = g ALPHA for gas), then f = 0
    float g3[3] = {c.gravity[0], c.gravity[1], c.gravity[2]};
    if (at.phase == PS_PHASE_GAS)
        for (int a = 0; a < 3; a++) g3[a] = g3[a] * kGasAlpha;
    const float4 fo = F[i];
    v.x = (v.x + (dt * g3[0])) + (dt * fo.x);
    v.y = (v.y + (dt * g3[1])) + (dt * fo.y);
    if (!two) v.z = (v.z + (dt * g3[2])) + (dt * fo.z);
    F[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float p[3] = {x.x, x.y, x.z};
    float e[3];
    if (w == 0.0f) {
        e[0] = x.x;
        e[1] = x.y;
        e[2] = x.z;
    } else {
        e[0] = x.x + (dt * v.x);
        e[1] = x.y + (dt * v.y);
        e[2] = two ? x.z : (x.z + (dt * v.z));
    }
    const float r = c.radius;
    uint32_t f = 0;
    if (e[0] < c.box_min[0] + r) f |= BX_LOW;
    else if (e[0] > c.box_max[0] - r) f |= BX_HIGH;
    if (e[1] < c.box_min[1] + r) f |= BY_LOW;
    else if (e[1] > c.box_max[1] - r) f |= BY_HIGH;
    if (!two) {
        if (e[2] < c.box_min[2] + r) f |= BZ_LOW;
        else if (e[2] > c.box_max[2] - r) f |= BZ_HIGH;
    }
    const bool fl = at.phase == PS_PHASE_FLUID || at.phase == PS_PHASE_GAS;
    const uint32_t bc = f | (fl ? 0u : kNoJitter) | ((uint32_t)__popc(f) << 8);
    if (boundaries) cr_boundaries(c, e, p, bc, st->iter, i);
    bcs[i] = bc;
    V[i] = v;
    E[i] = make_float4(e[0], e[1], e[2], w);
    // (4) Particle::scaleMass (particle.h:67-73) from the start height
    if (tm) tm[i] = w != 0.0f ? 1.0f / ((1.0f / w) * ps_exp(-x.y)) : 0.0f;
}

// (10)-(15) the stabilisation pre-pass (simulation.cpp:247-270, compiled
// with USE_STABILIZATION): STABILIZATION_ITERATIONS = 2 passes over the
// particle's boundary constraints, x then y then z, each with stabile =
// true (boundaryconstraint.cpp:14-71): ep and p both move to the push-out
// value, no friction.  Jitter from the stream 2^63 | iteration << 36 |
// pass << 35 | index << 3 | axis.
__global__ __launch_bounds__(kBlock) void cr_stabilize_k(float4 *__restrict__ E,
                                                         float4 *__restrict__ X,
                                                         const uint32_t *__restrict__ bcs,
                                                         const DevStatus *__restrict__ st,
                                                         CrConst c, uint32_t n)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t bc = bcs[i];
    if (!(bc & 0x3Fu)) return;
    float4 e = E[i], x = X[i];
    float ev[3] = {e.x, e.y, e.z}, xv[3] = {x.x, x.y, x.z};
    const uint64_t iter = st->iter & ((1ull << 27) - 1ull);
    for (uint32_t pass = 0; pass < 2; pass++) {
        for (uint32_t axis = 0; axis < 3; axis++) {
            const uint32_t lo = 1u << (2 * axis), hi = 2u << (2 * axis);
            if (!(bc & (lo | hi))) continue;
            const bool greater = (bc & lo) != 0;
            float extra = 0.0f;
            if (!(bc & kNoJitter)) {
                const uint64_t ctr = (1ull << 63) | (iter << 36) | ((uint64_t)pass << 35) |
                                     ((uint64_t)i << 3) | (uint64_t)axis;
                const uint64_t z = cr_mix64(c.seed + ctr * 0x9E3779B97F4A7C15ull + 1ull);
                extra = ((float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f)) * c.jitter;
            }
            const float r = c.radius, d = r + extra;
            const float value = greater ? c.box_min[axis] : c.box_max[axis];
            if (greater) {
                if (ev[axis] >= value + r) continue;
                ev[axis] = value + d;
                xv[axis] = value + d;
            } else {
                if (ev[axis] <= value - r) continue;
                ev[axis] = value - d;
                xv[axis] = value - d;
            }
        }
    }
    E[i] = make_float4(ev[0], ev[1], ev[2], e.w);
    X[i] = make_float4(xv[0], xv[1], xv[2], x.w);
}

// (7) contacts (simulation.cpp:164-199): a pair with exactly one immovable
// particle (every solid of this mode is immovable, and every other particle
// movable) whose predicted centres are closer than c.touch.  One wavefront
// per particle: its partner count joins its constraint count (both ends of
// a ContactConstraint count it, contactconstraint.cpp:94-98), and a movable
// particle lists its partners in index order.
__global__ __launch_bounds__(kBlock) void cr_find_contacts_k(const float4 *__restrict__ E,
                                                             uint32_t *__restrict__ bcs,
                                                             uint32_t *__restrict__ cnum,
                                                             uint32_t *__restrict__ clist,
                                                             DevStatus *__restrict__ st, CrConst c,
                                                             uint32_t n)
{
    const uint32_t i = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (i >= n) return;  // whole wavefront
    const float4 ei = E[i];
    const bool movable = ei.w != 0.0f;
    uint32_t cnt = 0;
    for (uint32_t base = 0; base < n; base += 64) {
        const uint32_t j = base + lane;
        const bool valid = j < n && j != i;
        const float4 ej = E[valid ? j : i];
        const float dx = ei.x - ej.x, dy = ei.y - ej.y, dz = ei.z - ej.z;
        const float dist = sqrtf(((dx * dx) + (dy * dy)) + (dz * dz));
        const bool hit = valid && ((ej.w != 0.0f) != movable) && dist < c.touch;
        const uint64_t mask = __ballot(hit);
        if (movable && hit) {
            const uint32_t k = cnt + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            if (k < kCrMaxContacts) clist[(size_t)i * kCrMaxContacts + k] = j;
        }
        cnt += (uint32_t)__popcll(mask);
    }
    if (lane != 0) return;
    if (movable && cnt > kCrMaxContacts) atomicOr(&st->flags, PS_STATUS_NEIGHBOR_OVERFLOW);
    cnum[i] = movable ? (cnt < kCrMaxContacts ? cnt : kCrMaxContacts) : 0u;
    bcs[i] = bcs[i] + (cnt << 8);
}

// The CONTACT group of one iteration for particle i: ContactConstraint::
// project (contactconstraint.cpp:14-40) with each immovable partner in index
// order (the order the reference creates them), then the boundaries.  The
// partner's position is read from E, before its own boundaries of this
// iteration, as the reference's creation order has it; it does not move.
__global__ __launch_bounds__(kBlock) void cr_contacts_k(const float4 *__restrict__ E,
                                                        float4 *__restrict__ E2,
                                                        const float4 *__restrict__ X,
                                                        const uint32_t *__restrict__ bcs,
                                                        const uint32_t *__restrict__ cnum,
                                                        const uint32_t *__restrict__ clist,
                                                        const float *__restrict__ tm,
                                                        const DevStatus *__restrict__ st, CrConst c,
                                                        uint32_t n, uint32_t it)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 x = X[i];
    const float4 ei = E[i];
    const uint32_t bc = bcs[i];
    float e[3] = {ei.x, ei.y, ei.z};
    const uint32_t m = cnum[i];
    if (m) {
        const float ti = tm[i], cnt = (float)(bc >> 8), diam = 2.0f * c.radius;
        for (uint32_t k = 0; k < m; k++) {
            const uint32_t j = clist[(size_t)i * kCrMaxContacts + k];
            const float4 ej = E[j];
            // diff = ep(lower index) - ep(higher index); w sum = t + 0
            const bool first = i < j;
            float d[3];
            if (first) {
                d[0] = e[0] - ej.x;
                d[1] = e[1] - ej.y;
                d[2] = e[2] - ej.z;
            } else {
                d[0] = ej.x - e[0];
                d[1] = ej.y - e[1];
                d[2] = ej.z - e[2];
            }
            const float wsum = ti;  // t_i + 0 (the partner is immovable)
            const float dist = sqrtf(((d[0] * d[0]) + (d[1] * d[1])) + (d[2] * d[2]));
            const float mag = dist - diam;
            if (mag > 0.0f) continue;
            const float scale = mag / wsum;
            for (int a = 0; a < 3; a++) {
                const float dp = (scale / dist) * d[a];
                e[a] = e[a] + (first ? (((-ti) * dp) / cnt) : ((ti * dp) / cnt));
            }
        }
    }
    const float p[3] = {x.x, x.y, x.z};
    cr_boundaries(c, e, p, bc, st->iter + it, i);
    E2[i] = make_float4(e[0], e[1], e[2], ei.w);
}

// TotalFluidConstraint::project, first loop (totalfluidconstraint.cpp:44-90):
// lambda of every particle of the object [first, first + count).
// gas != 0: GasConstraint::project's first loop (gasconstraint.cpp:34-84):
// its relaxation, and for an open gas (gas == 2) the drag
// f += v (1 - rho/rho0) (-50).  (Neither kind of object has solid
// neighbours in this mode, so no solid weights appear.)
__global__ __launch_bounds__(kBlock) void cr_lambda_k(const float4 *__restrict__ E,
                                                      float *__restrict__ lam,
                                                      const float4 *__restrict__ V,
                                                      float4 *__restrict__ F, CrConst c,
                                                      uint32_t n, uint32_t first, uint32_t count,
                                                      float rho0, int gas)
{
    const uint32_t wv = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (wv >= count) return;  // whole wavefront
    const uint32_t i = first + wv;
    const float4 ei = E[i];
    // a particle turned solid by a FluidEmitter has left the object
    if (ei.w == 0.0f) return;  // whole wavefront
    float acc = 0.0f, den = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
    for (uint32_t base = 0; base < n; base += 64) {
        const uint32_t j = base + lane;
        const bool valid = j < n;
        const float4 ej = E[valid ? j : i];
        const float rx = ei.x - ej.x, ry = ei.y - ej.y, rz = ei.z - ej.z;
        const float r2 = ((rx * rx) + (ry * ry)) + (rz * rz);
        const bool self = valid && j == i;
        const bool nb = valid && !self && ej.w != 0.0f && r2 < c.h2;
        uint64_t mask = __ballot(nb || self);
        if (!mask) continue;
        float incr = 0.0f, dd = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
        if (nb) {
            incr = cr_poly6(c, r2) / ej.w;
            const float l = sqrtf(r2);
            cr_spiky(c, rx, ry, rz, r2, l, sx, sy, sz);
            const float grx = (-sx) / rho0, gry = (-sy) / rho0, grz = (-sz) / rho0;
            dd = ((grx * grx) + (gry * gry)) + (grz * grz);
        } else if (self) {
            incr = cr_poly6(c, 0.0f) / ei.w;
        }
        // in index order: the self lane adds +0 to den and the gradient sum,
        // which leaves them unchanged (neither can be -0)
        while (mask) {
            const uint32_t b = (uint32_t)__builtin_ctzll(mask);
            mask &= mask - 1;
            acc = acc + lane_value(incr, b);
            den = den + lane_value(dd, b);
            gx = gx + lane_value(sx, b);
            gy = gy + lane_value(sy, b);
            gz = gz + lane_value(sz, b);
        }
    }
    const float grx = gx / rho0, gry = gy / rho0, grz = gz / rho0;
    den = den + (((grx * grx) + (gry * gry)) + (grz * grz));
    if (lane != 0) return;
    const float p_rat = acc / rho0;
    if (gas == 2) {
        const float4 vi = V[i];
        float4 f = F[i];
        f.x = f.x + ((vi.x * (1.0f - p_rat)) * kGasOpenDrag);
        f.y = f.y + ((vi.y * (1.0f - p_rat)) * kGasOpenDrag);
        F[i] = f;
    }
    lam[i] = -(p_rat - 1.0f) / (den + (gas ? kGasRelax : c.relax));
}

// TotalFluidConstraint::project, delta and update loops
// (totalfluidconstraint.cpp:92-115) for the object [first, first + count),
// ping-pong E -> E2 over every particle (others are copied).  With
// next_iter >= 0 the boundary projections of the next iteration follow.
// gas != 0: GasConstraint::project's delta loop (gasconstraint.cpp:86-107):
// the gas tensile term and the vorticity force
// f += sum_j cross((0,0,|G_j * v_j|), r) W(r^2), G_j = spikyGrad(r, r.r)
// (the squared length where the length is expected, and G_j * v_j
// component-wise, as written there; 2D).
__global__ __launch_bounds__(kBlock) void cr_delta_k(const float4 *__restrict__ E,
                                                     float4 *__restrict__ E2,
                                                     const float *__restrict__ lam,
                                                     const float4 *__restrict__ X,
                                                     const uint32_t *__restrict__ bcs,
                                                     const float4 *__restrict__ V,
                                                     float4 *__restrict__ F,
                                                     const DevStatus *__restrict__ st, CrConst c,
                                                     uint32_t n, uint32_t first, uint32_t count,
                                                     float rho0, int gas, int next_iter)
{
    const uint32_t i = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (i >= n) return;  // whole wavefront
    float4 ei = E[i];
    const uint32_t bc = bcs[i];
    if (i - first < count && ei.w != 0.0f) {
        const float li = lam[i];
        const float kc = gas ? kGasKP : c.k_corr, wdq = gas ? c.g_wdq : c.w_dq;
        const uint32_t nc = gas ? kGasEP : c.n_corr;
        float dx = 0.0f, dy = 0.0f, dz = 0.0f, fvx = 0.0f, fvy = 0.0f;
        uint32_t cnt = 1;  // the particle itself is in its neighbour list
        for (uint32_t base = 0; base < n; base += 64) {
            const uint32_t j = base + lane;
            const bool valid = j < n;
            const float4 ej = E[valid ? j : i];
            const float rx = ei.x - ej.x, ry = ei.y - ej.y, rz = ei.z - ej.z;
            const float r2 = ((rx * rx) + (ry * ry)) + (rz * rz);
            const bool nb = valid && j != i && ej.w != 0.0f && r2 < c.h2;
            uint64_t mask = __ballot(nb);
            if (!mask) continue;
            cnt += (uint32_t)__popcll(mask);
            float tx = 0.0f, ty = 0.0f, tz = 0.0f, ux = 0.0f, uy = 0.0f;
            if (nb) {
                const float l = sqrtf(r2);
                float sx, sy, sz;
                cr_spiky(c, rx, ry, rz, r2, l, sx, sy, sz);
                const float xr = cr_poly6(c, l * l) / wdq;
                const float corr = -(kc * cr_ipow(xr, nc));
                // lambdas of particles outside this object read as 0 (QHash default)
                const float lj = (j - first < count) ? lam[j] : 0.0f;
                const float coef = (li + lj) + corr;
                tx = coef * sx;
                ty = coef * sy;
                tz = coef * sz;
                if (gas) {
                    float vx, vy, vz;
                    cr_spiky(c, rx, ry, rz, r2, r2, vx, vy, vz);
                    const float4 vj = V[j];
                    const float wx = vx * vj.x, wy = vy * vj.y;
                    const float W = sqrtf((wx * wx) + (wy * wy));
                    const float p6 = cr_poly6(c, r2);
                    ux = ((0.0f * 0.0f) - (ry * W)) * p6;
                    uy = ((W * rx) - (0.0f * 0.0f)) * p6;
                }
            }
            while (mask) {
                const uint32_t b = (uint32_t)__builtin_ctzll(mask);
                mask &= mask - 1;
                dx = dx + lane_value(tx, b);
                dy = dy + lane_value(ty, b);
                dz = dz + lane_value(tz, b);
                if (gas) {
                    fvx = fvx + lane_value(ux, b);
                    fvy = fvy + lane_value(uy, b);
                }
            }
        }
        const float div = (float)(cnt + (bc >> 8));
        ei.x = ei.x + ((dx / rho0) / div);
        ei.y = ei.y + ((dy / rho0) / div);
        ei.z = ei.z + ((dz / rho0) / div);
        if (gas && lane == 0) {
            float4 f = F[i];
            f.x = f.x + fvx;
            f.y = f.y + fvy;
            F[i] = f;
        }
    }
    if (next_iter >= 0) {
        const float4 x = X[i];
        const float p[3] = {x.x, x.y, x.z};
        float e[3] = {ei.x, ei.y, ei.z};
        cr_boundaries(c, e, p, bc, st->iter + (uint64_t)next_iter, i);
        ei.x = e[0];
        ei.y = e[1];
        ei.z = e[2];
    }
    if (lane == 0) E2[i] = ei;
}

// Boundary projections alone (a scene with no fluid object).
__global__ __launch_bounds__(kBlock) void cr_boundary_k(float4 *__restrict__ E,
                                                        const float4 *__restrict__ X,
                                                        const uint32_t *__restrict__ bcs,
                                                        const DevStatus *__restrict__ st,
                                                        CrConst c, uint32_t n, uint32_t it)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 x = X[i];
    float4 ei = E[i];
    const float p[3] = {x.x, x.y, x.z};
    float e[3] = {ei.x, ei.y, ei.z};
    cr_boundaries(c, e, p, bcs[i], st->iter + it, i);
    E[i] = make_float4(e[0], e[1], e[2], ei.w);
}

// (23)-(28): v = (ep - p)/dt; sleeping keeps p and zeroes v.
__global__ __launch_bounds__(kBlock) void cr_end_k(const float4 *__restrict__ E,
                                                   float4 *__restrict__ X, float4 *__restrict__ V,
                                                   DevStatus *__restrict__ st, uint32_t n,
                                                   float dt, int sleeping, uint32_t its)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) st->iter += its;
    if (i >= n) return;
    const float4 e = E[i], x = X[i];
    const float dx = e.x - x.x, dy = e.y - x.y, dz = e.z - x.z;
    float4 v = make_float4(dx / dt, dy / dt, dz / dt, 0.0f);
    if (!(isfinite(e.x) && isfinite(e.y) && isfinite(e.z))) atomicOr(&st->flags, PS_STATUS_NONFINITE);
    if (sleeping && sqrtf(((dx * dx) + (dy * dy)) + (dz * dz)) < 1e-4f) {
        V[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    V[i] = v;
    X[i] = make_float4(e.x, e.y, e.z, 0.0f);
}

// OpenSmokeEmitter (opensmokeemitter.cpp:17-46): `count` emissions at pos.
// Each writes a marker (slot first_mark + k) and, with first_part != ~0, a
// gas particle at rest of inverse mass 1 (slot first_part + k, Particle's
// constructor, cpu/src/particle.h:31-52).
__global__ __launch_bounds__(64) void cr_emit_k(float4 *__restrict__ X, float4 *__restrict__ V,
                                                Attr *__restrict__ A, float4 *__restrict__ F,
                                                uint32_t *__restrict__ slot_of, float4 *__restrict__ M,
                                                uint32_t first_mark, uint32_t first_part, uint32_t count,
                                                float px, float py, float pz, float rho0)
{
    const uint32_t k = threadIdx.x;
    if (k >= count) return;
    M[first_mark + k] = make_float4(px, py, pz, 0.0f);
    if (first_part == ~0u) return;
    const uint32_t i = first_part + k;
    X[i] = make_float4(px, py, pz, 0.0f);
    V[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    F[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    Attr a;
    a.w = 1.0f;
    a.phase = PS_PHASE_GAS;
    a.rho0 = rho0;
    a.orig = i;
    A[i] = a;
    slot_of[i] = i;
}

// Marker advection (opensmokeemitter.cpp:37-45) and box clamp
// (simulation.cpp:348-362): v = sum_j v_j W(|m - x_j|^2), s = sum_j W,
// m += (v dt) / s, then x and y clamped to the box.  One wavefront per
// marker; the terms are added in particle order (zero terms, which leave
// the sums unchanged, are skipped).
__global__ __launch_bounds__(kBlock) void cr_markers_k(float4 *__restrict__ M,
                                                       const float4 *__restrict__ X,
                                                       const float4 *__restrict__ V, CrConst c,
                                                       uint32_t nmark, uint32_t n, float dt)
{
    const uint32_t m = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (m >= nmark) return;  // whole wavefront
    float4 q = M[m];
    float vx = 0.0f, vy = 0.0f, vz = 0.0f, sum = 0.0f;
    for (uint32_t base = 0; base < n; base += 64) {
        const uint32_t j = base + lane;
        const bool valid = j < n;
        const float4 xj = X[valid ? j : 0];
        const float rx = q.x - xj.x, ry = q.y - xj.y, rz = q.z - xj.z;
        const float p6 = valid ? cr_poly6(c, ((rx * rx) + (ry * ry)) + (rz * rz)) : 0.0f;
        uint64_t mask = __ballot(p6 != 0.0f);
        if (!mask) continue;
        float tx = 0.0f, ty = 0.0f, tz = 0.0f;
        if (p6 != 0.0f) {
            const float4 vj = V[j];
            tx = vj.x * p6;
            ty = vj.y * p6;
            tz = vj.z * p6;
        }
        while (mask) {
            const uint32_t b = (uint32_t)__builtin_ctzll(mask);
            mask &= mask - 1;
            vx = vx + lane_value(tx, b);
            vy = vy + lane_value(ty, b);
            vz = vz + lane_value(tz, b);
            sum = sum + lane_value(p6, b);
        }
    }
    if (lane != 0) return;
    if (sum > 0.0f) {
        q.x = q.x + ((vx * dt) / sum);
        q.y = q.y + ((vy * dt) / sum);
        q.z = q.z + ((vz * dt) / sum);
    }
    if (q.x < c.box_min[0]) q.x = c.box_min[0];
    else if (q.x > c.box_max[0]) q.x = c.box_max[0];
    if (q.y < c.box_min[1]) q.y = c.box_min[1];
    else if (q.y > c.box_max[1]) q.y = c.box_max[1];
    M[m] = q;
}

// FluidEmitter::tick's first loop (fluidemitter.cpp:13-62) over fluid
// object [first, first + count), one workgroup.  Members are the particles
// of the range that are not immovable; their position in the object's list
// is their rank among members (block scan), and `lambdas[k]` is keyed by
// particle index with that position k: the lambda of particle k when it was
// a member at the object's last projection, else 0.  Pass 1 updates t and
// marks the particles that turn solid; pass 2 turns them into immovable
// solids (imass 0, SOLID, v = f = 0; ep = p follows at the next prediction).
constexpr uint32_t kSolidifyBlock = 1024;

__global__ __launch_bounds__(kSolidifyBlock) void cr_solidify_k(const float4 *__restrict__ X,
                                                                float4 *__restrict__ V,
                                                                Attr *__restrict__ A,
                                                                float4 *__restrict__ F,
                                                                const float *__restrict__ lam,
                                                                float *__restrict__ tt,
                                                                uint32_t *__restrict__ sflag,
                                                                uint32_t first, uint32_t count,
                                                                float dt)
{
    __shared__ uint32_t wave_sum[kSolidifyBlock / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < count; base += kSolidifyBlock) {
        const uint32_t off = base + tid;
        const uint32_t j = first + off;
        const bool member = off < count && A[j].w != 0.0f;
        const uint64_t mask = __ballot(member);
        if (lane == 0) wave_sum[wave] = (uint32_t)__popcll(mask);
        __syncthreads();
        uint32_t before = carry, total = 0;
        for (uint32_t w = 0; w < kSolidifyBlock / 64; w++) {
            if (w < wave) before += wave_sum[w];
            total += wave_sum[w];
        }
        const uint32_t k = before + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (off < count) sflag[j] = 0u;
        if (member) {
            const float4 v = V[j], x = X[j];
            const float speed = sqrtf(((v.x * v.x) + (v.y * v.y)) + (v.z * v.z));
            if (speed < 0.06f && x.y <= 5.0f) {
                // particle k was a member: inside the range and movable
                // (pass 1 changes no attribute)
                const bool was_member = k - first < count && A[k].w != 0.0f;
                const float lk = was_member ? lam[k] : 0.0f;
                float t = tt[j];
                if (lk <= 0.0f) {
                    t = t - 1.0f;
                    if (t <= 0.0f) {
                        t = 0.0f;
                        sflag[j] = 1u;
                    }
                } else {
                    t = t + dt;
                    if (t > 3.0f) t = 3.0f;
                }
                tt[j] = t;
            }
        }
        carry += total;
        __syncthreads();  // wave_sum is rewritten by the next chunk
    }
    __syncthreads();
    for (uint32_t off = tid; off < count; off += kSolidifyBlock) {
        const uint32_t j = first + off;
        if (!sflag[j]) continue;
        Attr a = A[j];
        a.w = 0.0f;
        a.phase = PS_PHASE_SOLID;
        A[j] = a;
        V[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        F[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// FluidEmitter's emissions (fluidemitter.cpp:69-78): fluid particles of
// mass 1 at pos with v = (u, 1), u the emission's frand() replacement:
// SplitMix64 of seed + golden * (2^62 | emitter << 40 | emission) + 1, top
// 24 bits.
__global__ __launch_bounds__(64) void cr_emit_fluid_k(float4 *__restrict__ X, float4 *__restrict__ V,
                                                      Attr *__restrict__ A, float4 *__restrict__ F,
                                                      float *__restrict__ tt,
                                                      uint32_t *__restrict__ slot_of, uint64_t seed,
                                                      uint32_t first_part, uint32_t count, float px,
                                                      float py, float pz, float rho0, uint32_t eidx,
                                                      uint64_t emitted)
{
    const uint32_t k = threadIdx.x;
    if (k >= count) return;
    const uint32_t i = first_part + k;
    const uint64_t ctr = (1ull << 62) | ((uint64_t)eidx << 40) | (emitted + k);
    const uint64_t z = cr_mix64(seed + ctr * 0x9E3779B97F4A7C15ull + 1ull);
    const float u = (float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f);
    X[i] = make_float4(px, py, pz, 0.0f);
    V[i] = make_float4(u, 1.0f, 0.0f, 0.0f);
    F[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    tt[i] = 4.0f;
    Attr a;
    a.w = 1.0f;
    a.phase = PS_PHASE_FLUID;
    a.rho0 = rho0;
    a.orig = i;
    A[i] = a;
    slot_of[i] = i;
}

inline uint32_t blocks_for(uint32_t threads) { return (threads + kBlock - 1) / kBlock; }

}  // namespace

hipError_t launch_cr_solidify(hipStream_t st, const DevBuffers &b, uint32_t first, uint32_t count,
                              float dt)
{
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(cr_solidify_k, dim3(1), dim3(kSolidifyBlock), 0, st, b.X, b.V, b.A, b.force,
                       b.lam, b.ttimer, b.sflag, first, count, dt);
    return hipGetLastError();
}

hipError_t launch_cr_emit_fluid(hipStream_t st, const DevBuffers &b, const CrConst &c,
                                uint32_t first_part, uint32_t count, const float *pos, float rho0,
                                uint32_t eidx, uint64_t emitted)
{
    if (!count) return hipSuccess;
    if (count > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cr_emit_fluid_k, dim3(1), dim3(64), 0, st, b.X, b.V, b.A, b.force, b.ttimer,
                       b.slot_of, c.seed, first_part, count, pos[0], pos[1], pos[2], rho0, eidx,
                       emitted);
    return hipGetLastError();
}

hipError_t launch_cr_emit(hipStream_t st, const DevBuffers &b, float4 *markers, uint32_t first_mark,
                          uint32_t first_part, uint32_t count, const float *pos, float rho0)
{
    if (!count) return hipSuccess;
    if (count > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cr_emit_k, dim3(1), dim3(64), 0, st, b.X, b.V, b.A, b.force, b.slot_of, markers,
                       first_mark, first_part, count, pos[0], pos[1], pos[2], rho0);
    return hipGetLastError();
}

hipError_t launch_cr_markers(hipStream_t st, const DevBuffers &b, float4 *markers, const CrConst &c,
                             uint32_t nmark, uint32_t n, float dt)
{
    if (!nmark) return hipSuccess;
    hipLaunchKernelGGL(cr_markers_k, dim3((nmark + kWaves - 1) / kWaves), dim3(kBlock), 0, st, markers,
                       b.X, b.V, c, nmark, n, dt);
    return hipGetLastError();
}

hipError_t launch_cr_predict(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n,
                             float dt, bool boundaries)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(cr_predict_k, dim3(blocks_for(n)), dim3(kBlock), 0, st, b.X, b.V, b.A, b.P,
                       b.bc, b.force, b.tmass, b.status, c, n, dt, boundaries ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_cr_stabilize(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(cr_stabilize_k, dim3(blocks_for(n)), dim3(kBlock), 0, st, b.P, b.X, b.bc,
                       b.status, c, n);
    return hipGetLastError();
}

hipError_t launch_cr_find_contacts(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(cr_find_contacts_k, dim3((n + kWaves - 1) / kWaves), dim3(kBlock), 0, st, b.P,
                       b.bc, b.cnum, b.clist, b.status, c, n);
    return hipGetLastError();
}

hipError_t launch_cr_contacts(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n,
                              uint32_t it)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(cr_contacts_k, dim3(blocks_for(n)), dim3(kBlock), 0, st, b.P, b.Q, b.X, b.bc,
                       b.cnum, b.clist, b.tmass, b.status, c, n, it);
    return hipGetLastError();
}

hipError_t launch_cr_lambda(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n,
                            uint32_t first, uint32_t count, float rho0, int gas)
{
    if (!n || !count) return hipSuccess;
    hipLaunchKernelGGL(cr_lambda_k, dim3((count + kWaves - 1) / kWaves), dim3(kBlock), 0, st, b.P,
                       b.lam, b.V, b.force, c, n, first, count, rho0, gas);
    return hipGetLastError();
}

hipError_t launch_cr_delta(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n,
                           uint32_t first, uint32_t count, float rho0, int gas, int next_iter)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(cr_delta_k, dim3((n + kWaves - 1) / kWaves), dim3(kBlock), 0, st, b.P, b.Q,
                       b.lam, b.X, b.bc, b.V, b.force, b.status, c, n, first, count, rho0, gas,
                       next_iter);
    return hipGetLastError();
}

hipError_t launch_cr_boundary(hipStream_t st, const DevBuffers &b, const CrConst &c, uint32_t n,
                              uint32_t it)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(cr_boundary_k, dim3(blocks_for(n)), dim3(kBlock), 0, st, b.P, b.X, b.bc,
                       b.status, c, n, it);
    return hipGetLastError();
}

hipError_t launch_cr_end(hipStream_t st, const DevBuffers &b, uint32_t n, float dt, bool sleeping,
                         uint32_t its)
{
    hipLaunchKernelGGL(cr_end_k, dim3(blocks_for(n ? n : 1)), dim3(kBlock), 0, st, b.P, b.X, b.V,
                       b.status, n, dt, sleeping ? 1 : 0, its);
    return hipGetLastError();
}

}  // namespace ps
