// ps_device.hpp — device helpers shared by the forward kernels
// (ps_kernels.hip) and the adjoint kernels (ps_adjoint.hip): launch mapping,
// neighbour-table access, the pair arithmetic shared with the CPU oracle,
// the wall-jitter RNG and the world-box projection.
#pragma once

#include "ps_internal.hpp"

namespace ps {

__device__ __forceinline__ uint32_t xcd_block()
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t x = b & 7u, q = nb >> 3, r = nb & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// A status word as this kernel has updated it so far, read before an atomic
// that may change nothing.  The words only grow (bits set, maxima raised)
// within a kernel, so a stale value can only cost a redundant atomic.
__device__ __forceinline__ uint32_t status_peek(const uint32_t *w)
{
    return __atomic_load_n(w, __ATOMIC_RELAXED);
}

// Cell coordinate of p on one axis (calcGridPos, integration_kernel.cuh:
// 187-194): floor((p - o) / cell) by IEEE division, clamped into [0, g - 1]
// of the dense grid (positions outside raise PS_STATUS_OUTSIDE_GRID in
// flags, non-finite ones PS_STATUS_NONFINITE and land in cell 0).
__device__ __forceinline__ uint32_t cell_axis(float p, float o, float cell, uint32_t g,
                                              uint32_t &flags)
{
    float f = floorf((p - o) / cell);
    if (!(f >= 0.0f)) {
        flags |= (f != f) ? PS_STATUS_NONFINITE : PS_STATUS_OUTSIDE_GRID;
        return 0u;
    }
    if (f >= (float)g) {
        flags |= PS_STATUS_OUTSIDE_GRID;
        return g - 1u;
    }
    return (uint32_t)f;
}

// chunk m (entries 4m..4m+3) of particle i is nbr_lane(b, i)[m * 64]
__device__ __forceinline__ const uint4 *nbr_lane(const DevBuffers &b, uint32_t i)
{
    return reinterpret_cast<const uint4 *>(b.nbr) + (size_t)(i >> 6) * b.kq * 64 + (i & 63);
}

// Visit the first c neighbours of particle i in table order, calling
// fn(j, gather(j)) for each.  Software-pipelined: while chunk m is being
// processed, the four gathers of chunk m+1 and the index load of chunk m+2
// are already in flight.  The loop is unrolled over two chunks whose record
// sets swap roles, so no records are copied between registers.  Every entry
// of a chunk at or below the last one is a valid slot (the builder pads the
// last chunk with the particle's own slot, or with the sentinel record in a
// staged unit), so all loads are unconditional; only fn is guarded.
template <typename G, typename Fn>
__device__ __forceinline__ void for_neighbors(const DevBuffers &b, uint32_t i, uint32_t c, G gather,
                                              Fn fn)
{
    if (c == 0) return;
    const uint4 *nb = nbr_lane(b, i);
    const uint32_t last = (c - 1) >> 2;
    auto chunk = [&](uint32_t m) { return nb[(size_t)min(m, last) * 64]; };
    auto visit = [&](const uint4 &jj, uint32_t q, const auto &d0, const auto &d1, const auto &d2,
                     const auto &d3) {
        fn(jj.x, d0);
        if (q + 1 < c) fn(jj.y, d1);
        if (q + 2 < c) fn(jj.z, d2);
        if (q + 3 < c) fn(jj.w, d3);
    };
    uint4 ja = nb[0];
    uint4 jb = chunk(1);
    auto a0 = gather(ja.x);
    auto a1 = gather(ja.y);
    auto a2 = gather(ja.z);
    auto a3 = gather(ja.w);
    for (uint32_t m = 0;; m += 2) {
        // records of chunk m are in a*; chunk m+1 goes in flight
        const uint4 jc = chunk(m + 2);
        auto b0 = gather(jb.x);
        auto b1 = gather(jb.y);
        auto b2 = gather(jb.z);
        auto b3 = gather(jb.w);
        visit(ja, m << 2, a0, a1, a2, a3);
        if (m + 1 > last) break;
        // chunk m+2 goes in flight into a*; chunk m+1 is processed
        const uint4 jd = chunk(m + 3);
        a0 = gather(jc.x);
        a1 = gather(jc.y);
        a2 = gather(jc.z);
        a3 = gather(jc.w);
        visit(jb, (m + 1) << 2, b0, b1, b2, b3);
        if (m + 2 > last) break;
        ja = jc;
        jb = jd;
    }
}

// The fluid sums (density, λ's gradient sums, the Δp sum) are formed in two
// halves, h = 0 over the table's chunks 0, 2, 4, ... and h = 1 over chunks
// 1, 3, 5, ..., each in table order, and then added as half 0 + half 1.  The
// staged kernels give each half its own lane (two lanes per particle, twice
// the wavefronts for the same LDS); the others visit both halves in turn.
// Visits the entries of half h of particle i's first c neighbours, pipelined
// as for_neighbors.  GUARD = false visits the padding of the last chunk as
// well: for tables whose padding entries refer to the far-away sentinel
// record of a staged unit (kStageSentinel), which fails every pair test.
template <bool GUARD = true, typename G, typename Fn>
__device__ __forceinline__ void for_neighbors_half(const DevBuffers &b, uint32_t i, uint32_t c, uint32_t h,
                                                   G gather, Fn fn)
{
    const uint32_t nch = (c + 3u) >> 2;  // chunks in use
    if (h >= nch) return;
    const uint4 *nb = nbr_lane(b, i);
    const uint32_t last = ((nch - h + 1u) >> 1) - 1u;  // this half's chunks: 2m + h, m <= last
    auto chunk = [&](uint32_t m) { return nb[(size_t)(2u * min(m, last) + h) * 64]; };
    auto visit = [&](const uint4 &jj, uint32_t m, const auto &d0, const auto &d1, const auto &d2,
                     const auto &d3) {
        const uint32_t q = (2u * m + h) << 2;  // table position of jj.x
        fn(jj.x, d0);
        if (!GUARD || q + 1 < c) fn(jj.y, d1);
        if (!GUARD || q + 2 < c) fn(jj.z, d2);
        if (!GUARD || q + 3 < c) fn(jj.w, d3);
    };
    uint4 ja = chunk(0);
    uint4 jb = chunk(1);
    auto a0 = gather(ja.x);
    auto a1 = gather(ja.y);
    auto a2 = gather(ja.z);
    auto a3 = gather(ja.w);
    for (uint32_t m = 0;; m += 2) {
        const uint4 jc = chunk(m + 2);
        auto b0 = gather(jb.x);
        auto b1 = gather(jb.y);
        auto b2 = gather(jb.z);
        auto b3 = gather(jb.w);
        visit(ja, m, a0, a1, a2, a3);
        if (m + 1 > last) break;
        const uint4 jd = chunk(m + 3);
        a0 = gather(jc.x);
        a1 = gather(jc.y);
        a2 = gather(jc.z);
        a3 = gather(jc.w);
        visit(jb, m + 1, b0, b1, b2, b3);
        if (m + 2 > last) break;
        ja = jc;
        jb = jd;
    }
}

// Half h of particle i's first c neighbours, one chunk at a time as
// for_neighbors_chunked (four records in flight): for a kernel whose
// records or arithmetic are large that gives each half its own lane.
template <typename G, typename Fn>
__device__ __forceinline__ void for_neighbors_half_chunked(const DevBuffers &b, uint32_t i, uint32_t c,
                                                           uint32_t h, G gather, Fn fn)
{
    const uint32_t nch = (c + 3u) >> 2;  // chunks in use; half h has 2m + h
    if (h >= nch) return;
    const uint4 *nb = nbr_lane(b, i);
    uint4 jj = nb[(size_t)h * 64];
    for (uint32_t m = h; m < nch; m += 2u) {
        const uint4 jn = nb[(size_t)(m + 2u < nch ? m + 2u : m) * 64];
        auto d0 = gather(jj.x);
        auto d1 = gather(jj.y);
        auto d2 = gather(jj.z);
        auto d3 = gather(jj.w);
        const uint32_t q = m << 2;
        fn(jj.x, d0);
        if (q + 1 < c) fn(jj.y, d1);
        if (q + 2 < c) fn(jj.z, d2);
        if (q + 3 < c) fn(jj.w, d3);
        jj = jn;
    }
}

// The chunks first, first + stride, first + 2 stride, ... of particle i's
// first c neighbours, one chunk at a time (four records in flight), each in
// table order: quarter q of the table is first = q, stride = 4.
template <typename G, typename Fn>
__device__ __forceinline__ void for_neighbors_strided_chunked(const DevBuffers &b, uint32_t i, uint32_t c,
                                                              uint32_t first, uint32_t stride, G gather,
                                                              Fn fn)
{
    const uint32_t nch = (c + 3u) >> 2;  // chunks in use
    if (first >= nch) return;
    const uint4 *nb = nbr_lane(b, i);
    uint4 jj = nb[(size_t)first * 64];
    for (uint32_t m = first; m < nch; m += stride) {
        const uint4 jn = nb[(size_t)(m + stride < nch ? m + stride : m) * 64];
        auto d0 = gather(jj.x);
        auto d1 = gather(jj.y);
        auto d2 = gather(jj.z);
        auto d3 = gather(jj.w);
        const uint32_t q = m << 2;
        fn(jj.x, d0);
        if (q + 1 < c) fn(jj.y, d1);
        if (q + 2 < c) fn(jj.z, d2);
        if (q + 3 < c) fn(jj.w, d3);
        jj = jn;
    }
}

// As for_neighbors_strided_chunked with two records in flight instead of
// four: the gathers of a chunk's entries 2 and 3 go out once entries 0 and
// 1 are done.  For records so large (tens of bytes) that four of them in
// registers would cost a wavefront per SIMD.
template <typename G, typename Fn>
__device__ __forceinline__ void for_neighbors_strided_pairs(const DevBuffers &b, uint32_t i, uint32_t c,
                                                            uint32_t first, uint32_t stride, G gather,
                                                            Fn fn)
{
    const uint32_t nch = (c + 3u) >> 2;  // chunks in use
    if (first >= nch) return;
    const uint4 *nb = nbr_lane(b, i);
    uint4 jj = nb[(size_t)first * 64];
    for (uint32_t m = first; m < nch; m += stride) {
        const uint4 jn = nb[(size_t)(m + stride < nch ? m + stride : m) * 64];
        const uint32_t q = m << 2;
        {
            const auto d0 = gather(jj.x);
            const auto d1 = gather(jj.y);
            fn(jj.x, d0);
            if (q + 1 < c) fn(jj.y, d1);
        }
        if (q + 2 < c) {
            const auto d2 = gather(jj.z);
            const auto d3 = gather(jj.w);
            fn(jj.z, d2);
            if (q + 3 < c) fn(jj.w, d3);
        }
        jj = jn;
    }
}

// The entries of half h of particle i's first c neighbours in table order,
// one at a time: for the rarely taken recomputation of one half.
template <typename G, typename Fn>
__device__ __forceinline__ void for_neighbors_half_plain(const DevBuffers &b, uint32_t i, uint32_t c,
                                                         uint32_t h, G gather, Fn fn)
{
    for (uint32_t q0 = 4u * h; q0 < c; q0 += 8u)
        for (uint32_t q = q0; q < q0 + 4u && q < c; q++) {
            const uint32_t j = reinterpret_cast<const uint32_t *>(nbr_lane(b, i) + (size_t)(q >> 2) * 64)[q & 3u];
            fn(j, gather(j));
        }
}

// Both halves in one pass, as for_neighbors: fn0 gets the entries of the
// even chunks (half 0), fn1 those of the odd chunks (half 1), each half in
// table order.  The pipelined loop's two record sets hold an even and an
// odd chunk, so the half of every call is fixed where it is made.
template <typename G, typename Fn0, typename Fn1>
__device__ __forceinline__ void for_neighbors_split(const DevBuffers &b, uint32_t i, uint32_t c, G gather,
                                                    Fn0 fn0, Fn1 fn1)
{
    if (c == 0) return;
    const uint4 *nb = nbr_lane(b, i);
    const uint32_t last = (c - 1) >> 2;
    auto chunk = [&](uint32_t m) { return nb[(size_t)min(m, last) * 64]; };
    auto visit = [&](auto &fn, const uint4 &jj, uint32_t q, const auto &d0, const auto &d1,
                     const auto &d2, const auto &d3) {
        fn(jj.x, d0);
        if (q + 1 < c) fn(jj.y, d1);
        if (q + 2 < c) fn(jj.z, d2);
        if (q + 3 < c) fn(jj.w, d3);
    };
    uint4 ja = nb[0];
    uint4 jb = chunk(1);
    auto a0 = gather(ja.x);
    auto a1 = gather(ja.y);
    auto a2 = gather(ja.z);
    auto a3 = gather(ja.w);
    for (uint32_t m = 0;; m += 2) {
        // a*: chunk m (even), b*: chunk m+1 (odd)
        const uint4 jc = chunk(m + 2);
        auto b0 = gather(jb.x);
        auto b1 = gather(jb.y);
        auto b2 = gather(jb.z);
        auto b3 = gather(jb.w);
        visit(fn0, ja, m << 2, a0, a1, a2, a3);
        if (m + 1 > last) break;
        const uint4 jd = chunk(m + 3);
        a0 = gather(jc.x);
        a1 = gather(jc.y);
        a2 = gather(jc.z);
        a3 = gather(jc.w);
        visit(fn1, jb, (m + 1) << 2, b0, b1, b2, b3);
        if (m + 2 > last) break;
        ja = jc;
        jb = jd;
    }
}

// As for_neighbors, one chunk at a time: the four gathers of chunk m are
// issued together (and the index load of chunk m+1), then processed.  Half
// the records in flight, for kernels whose records or arithmetic are large.
template <typename G, typename Fn>
__device__ __forceinline__ void for_neighbors_chunked(const DevBuffers &b, uint32_t i, uint32_t c, G gather,
                                                      Fn fn)
{
    if (c == 0) return;
    const uint4 *nb = nbr_lane(b, i);
    const uint32_t last = (c - 1) >> 2;
    uint4 jj = nb[0];
    for (uint32_t m = 0; m <= last; m++) {
        const uint4 jn = nb[(size_t)min(m + 1, last) * 64];
        auto d0 = gather(jj.x);
        auto d1 = gather(jj.y);
        auto d2 = gather(jj.z);
        auto d3 = gather(jj.w);
        const uint32_t q = m << 2;
        fn(jj.x, d0);
        if (q + 1 < c) fn(jj.y, d1);
        if (q + 2 < c) fn(jj.z, d2);
        if (q + 3 < c) fn(jj.w, d3);
        jj = jn;
    }
}

// As for_neighbors_chunked, and fn(h, j, record) is told the half h of the
// table the entry lies in (its chunk's parity).
template <typename G, typename Fn>
__device__ __forceinline__ void for_neighbors_chunked_h(const DevBuffers &b, uint32_t i, uint32_t c, G gather,
                                                        Fn fn)
{
    if (c == 0) return;
    const uint4 *nb = nbr_lane(b, i);
    const uint32_t last = (c - 1) >> 2;
    uint4 jj = nb[0];
    for (uint32_t m = 0; m <= last; m++) {
        const uint4 jn = nb[(size_t)min(m + 1, last) * 64];
        auto d0 = gather(jj.x);
        auto d1 = gather(jj.y);
        auto d2 = gather(jj.z);
        auto d3 = gather(jj.w);
        const uint32_t q = m << 2, h = m & 1u;
        fn(h, jj.x, d0);
        if (q + 1 < c) fn(h, jj.y, d1);
        if (q + 2 < c) fn(h, jj.z, d2);
        if (q + 3 < c) fn(h, jj.w, d3);
        jj = jn;
    }
}

struct PA {
    float4 p;
    Attr a;
};

struct PX {
    float4 p, x;
};

// position and neighbour density mass (DevBuffers::mj)
struct PM {
    float4 p;
    float m;
};

// the same and the phase (a staged unit's contact records)
struct PMH {
    float4 p;
    float m;
    int32_t ph;
};

// a staged position and phase
struct PH {
    float4 p;
    int32_t ph;
};

// entry q of particle i (single-entry access for the less travelled loops)
__device__ __forceinline__ uint32_t nbr_at(const DevBuffers &b, uint32_t i, uint32_t q)
{
    return reinterpret_cast<const uint32_t *>(nbr_lane(b, i) + (size_t)(q >> 2) * 64)[q & 3u];
}

// ---------------------------------------------------------------------------
// Pair arithmetic shared with the CPU oracle: squared distances and sums use
// fused multiply-add, and 1/r comes from rsq() instead of sqrt + division
// (an IEEE square root and division cost ~30 VALU per pair, rsq 9).

// |d|^2 = fma(dz, dz, fma(dy, dy, dx dx))
__device__ __forceinline__ float dist2(float dx, float dy, float dz)
{
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// 1/sqrt(x), x > 0 normal: integer seed y0 = 0x5f375a86 - (bits(x) >> 1)
// (the "fast inverse square root" constant of C. Lomont, 2003; relative
// error up to 3.4e-2), then one fourth-order correction.  With
// e = 1 - x y0^2, 1/sqrt(x) = y0 (1 - e)^(-1/2) = y0 (1 + e p(e)), p the
// binomial series 1/2 + 3/8 e + 5/16 e^2 + 35/128 e^3 + ..., truncated
// after e^3 and its four coefficients re-fitted minimax over the seed's
// range of e (|e| < 0.07): y = fma(y0 e, p(e), y0), p by Horner in FMA.
// 9 operations; relative error below 1e-7 over [1e-8, 16] (the three
// Newton steps it replaces took 12 for 1.3e-7).
__device__ __forceinline__ float rsq(float x)
{
    const float y0 = __uint_as_float(0x5f375a86u - (__float_as_uint(x) >> 1));
    const float xy = x * y0;
    const float e = __builtin_fmaf(-xy, y0, 1.0f);
    float p = __builtin_fmaf(__uint_as_float(0x3e8c5420u), e, __uint_as_float(0x3ea0bc22u));
    p = __builtin_fmaf(p, e, __uint_as_float(0x3ec0000cu));
    p = __builtin_fmaf(p, e, __uint_as_float(0x3effffc8u));
    return __builtin_fmaf(y0 * e, p, y0);
}

// below this squared distance a pair has no direction
constexpr float kDegenerateR2 = 1e-8f;

// ---------------------------------------------------------------------------
// helpers of the solve kernel

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Wall-jitter uniform in [0,1): SplitMix64 finaliser of seed + golden*counter,
// counter = (iteration, insertion index, wall), top 24 bits.
__device__ __forceinline__ float wall_uniform(uint64_t seed, uint64_t iter, uint32_t orig,
                                              uint32_t wall)
{
    uint64_t ctr = (iter << 35) | ((uint64_t)orig << 3) | (uint64_t)wall;
    uint64_t z = mix64(seed + ctr * 0x9E3779B97F4A7C15ull);
    return (float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f);
}

// e^x: Cody-Waite reduction x = k ln2 + r, degree-7 Taylor polynomial of
// e^r (Horner), exact scaling by 2^k.
__device__ __forceinline__ float ps_exp(float x)
{
    if (x > 80.0f) x = 80.0f;
    if (x < -80.0f) x = -80.0f;
    float kk = floorf((x * 1.44269504088896341f) + 0.5f);
    float r = (x - (kk * 0.693145751953125f)) - (kk * 1.42860682030941723e-6f);
    float p = 1.98412698412698413e-4f;
    p = (p * r) + 1.38888888888888889e-3f;
    p = (p * r) + 8.33333333333333333e-3f;
    p = (p * r) + 4.16666666666666667e-2f;
    p = (p * r) + 1.66666666666666667e-1f;
    p = (p * r) + 0.5f;
    p = (p * r) + 1.0f;
    p = (p * r) + 1.0f;
    return ldexpf(p, (int)kk);
}

__device__ __forceinline__ float ipow(float x, uint32_t n)
{
    float r = 1.0f, bb = x;
    bool first = true;
    while (n) {
        if (n & 1u) {
            r = first ? bb : (r * bb);
            first = false;
        }
        n >>= 1;
        if (n) bb = bb * bb;
    }
    return r;
}

// a8: world box (collide_world_functor, integration_kernel.cuh:57-157).
struct WallBox {
    float lo[3], hi[3];
    uint64_t seed;
    uint32_t o;  // index of the particle inside its scene
};

__device__ __forceinline__ WallBox wall_box(const KConst &k, const DevBuffers &b, uint32_t i,
                                            uint32_t orig)
{
    WallBox w;
    if (k.nscenes > 1u) {
        const SceneDev &sc = b.scenes[b.sid[i]];
        for (int a = 0; a < 3; a++) {
            w.lo[a] = sc.box_min[a];
            w.hi[a] = sc.box_max[a];
        }
        w.seed = sc.seed;
        w.o = orig - sc.first;
    } else {
        for (int a = 0; a < 3; a++) {
            w.lo[a] = k.box_min[a];
            w.hi[a] = k.box_max[a];
        }
        w.seed = k.seed;
        w.o = orig;
    }
    return w;
}

// Returns what it did, for the adjoint: bits 0-2 the coordinates a wall
// set, bit 3 wall friction acted, bit 4 the kinetic branch with f < 1 (the
// solve kernel adds bit 5 for a pin, with bits 0-2).  When friction acts,
// rec = (position before friction, (n_x+1)*9 + (n_y+1)*3 + (n_z+1)).
constexpr uint32_t kWallFriction = 8u, kWallKinetic = 16u, kPinned = 32u;

__device__ __forceinline__ uint32_t walls(const KConst &k, const WallBox &wb, int32_t ph, float4 xs,
                                          float &px, float &py, float &pz, uint64_t iter, float4 &rec)
{
    uint32_t set = 0u;
    const float r = k.radius;
    const uint32_t o = wb.o;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    const float efl = ph < PS_PHASE_SOLID ? k.jitter : 0.0f;
    if (py < wb.lo[1] + r) {
        py = (wb.lo[1] + r) + (wall_uniform(wb.seed, iter, o, 5) * efl);
        set |= 2u;
        ny = ny + 1.0f;
    }
    if (px > wb.hi[0] - r) {
        px = wb.hi[0] - (r + (wall_uniform(wb.seed, iter, o, 0) * k.jitter));
        set |= 1u;
        nx = nx - 1.0f;
    }
    if (px < wb.lo[0] + r) {
        px = wb.lo[0] + (r + (wall_uniform(wb.seed, iter, o, 1) * k.jitter));
        set |= 1u;
        nx = nx + 1.0f;
    }
    if (py > wb.hi[1] - r) {
        py = wb.hi[1] - (r + (wall_uniform(wb.seed, iter, o, 2) * k.jitter));
        set |= 2u;
        ny = ny - 1.0f;
    }
    if (k.dims != 2u) {
        if (pz > wb.hi[2] - r) {
            pz = wb.hi[2] - (r + (wall_uniform(wb.seed, iter, o, 3) * k.jitter));
            set |= 4u;
            nz = nz - 1.0f;
        }
        if (pz < wb.lo[2] + r) {
            pz = wb.lo[2] + (r + (wall_uniform(wb.seed, iter, o, 4) * k.jitter));
            set |= 4u;
            nz = nz + 1.0f;
        }
    }
    if (ph < PS_PHASE_CLOTH) return set;
    if (nx == 0.0f && ny == 0.0f && nz == 0.0f) return set;
    float ex = px - xs.x, ey = py - xs.y, ez = pz - xs.z;
    float dn = ((ex * nx) + (ey * ny)) + (ez * nz);
    float tx = ex - (dn * nx), ty = ey - (dn * ny), tz = ez - (dn * nz);
    float l = sqrtf(((tx * tx) + (ty * ty)) + (tz * tz));
    if (l < 0.001f) return set;
    set |= kWallFriction;
    rec = make_float4(px, py, pz,
                      (float)((((int)nx + 1) * 9 + ((int)ny + 1) * 3) + ((int)nz + 1)));
    if (l < k.wall_s) {
        px = px - tx;
        py = py - ty;
        pz = pz - tz;
    } else {
        float f = k.wall_k / l;
        if (f > 1.0f)
            f = 1.0f;
        else
            set |= kWallKinetic;
        px = px - (tx * f);
        py = py - (ty * f);
        pz = pz - (tz * f);
    }
    return set;
}

}  // namespace ps
