// ps_kernels.hip — gfx950 kernels of the particle-solver substep.
//
// Layout: every per-particle array is in cell-sorted storage order; a
// particle's insertion index travels in Attr::orig.  Positions and the PBF
// multiplier share one float4 (x, y, z, lambda), so each neighbour visit in
// the lambda and delta-p kernels is a single 16-byte gather.  The neighbour
// table is a sliced ELL in chunks of four: for each group of 64 consecutive
// particles, entries 4m..4m+3 of every lane form one uint4, and the 64 lanes'
// uint4s are contiguous.  A wavefront therefore reads four neighbour indices
// per lane with one coalesced 1-KiB load and issues the four position
// gathers back to back; the builder writes 16 bytes per lane at a time.
//
// Arithmetic: built with -ffp-contract=off and IEEE division/sqrt, and every
// expression is parenthesised as in the CPU oracle, so FP32 results are
// reproducible operation by operation.  Accumulations run over the neighbour
// table in ascending slot order; no floating-point atomics anywhere.
//
// Launch mapping: blocks of the neighbour-gathering kernels are remapped so
// that each of the 8 XCDs (which receive workgroups round-robin) owns one
// contiguous range of sorted particles, keeping a spatial region's
// neighbour data inside one XCD's L2.

#include "ps_device.hpp"

namespace ps {

static constexpr uint32_t kBlock = 256;     // fixed-shape kernels (scan, scatter, shape)
static constexpr uint32_t kMaxBlock = 512;  // tunable kernels run with 64..512 threads
static constexpr uint32_t kScanPerThread = 16;
static constexpr uint32_t kScanTile = kBlock * kScanPerThread;
static_assert(kScanTile == kScanTileCells, "scan tile size");


__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        uint32_t t = __shfl_xor(v, d, 64);
        v = t > v ? t : v;
    }
    return v;
}


// ---------------------------------------------------------------------------
// a1 + a2: predict (integrate_functor, integration_kernel.cuh:159-184) and
// cell hash (calcGridPos/calcGridHash/calcHashD :187-225) with a dense,
// clamped grid.  The atomic's return value is the particle's arrival rank in
// its cell, used only to place it provisionally; the final in-cell order is
// fixed by insertion index in reorder_k.

template <bool PRED, bool HASH>
__global__ __launch_bounds__(kMaxBlock) void predict_hash_k(DevBuffers b, KConst k, uint32_t n,
                                                         float dt)
{
    const uint32_t s = xcd_block() * blockDim.x + threadIdx.x;
    const bool valid = s < n;
    const bool multi = k.nscenes > 1u;
    const SceneDev *sc = (multi && valid) ? &b.scenes[b.sid[s]] : nullptr;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (valid) {
        if (PRED) {
            // the three loads go out together (a static particle's velocity
            // is read and not used)
            const float4 x = b.X[s];
            const Attr a = b.A[s];
            const float4 v = b.V[s];
            px = x.x;
            py = x.y;
            pz = x.z;
            if (a.w != 0.0f) {
                const float gx = multi ? sc->gravity[0] : k.gravity[0];
                const float gy = multi ? sc->gravity[1] : k.gravity[1];
                const float gz = multi ? sc->gravity[2] : k.gravity[2];
                float vx = v.x + (gx * dt);
                float vy = v.y + (gy * dt);
                float vz = v.z + (gz * dt);
                px = x.x + (vx * dt);
                py = x.y + (vy * dt);
                pz = k.dims == 2u ? x.z : (x.z + (vz * dt));
            }
            b.P[s] = make_float4(px, py, pz, 0.0f);
        } else {
            const float4 p = b.P[s];
            px = p.x;
            py = p.y;
            pz = p.z;
        }
    }
    if (HASH) {
        uint32_t flags = 0, key = 0xffffffffu;
        if (valid) {
            if (multi) {
                uint32_t cx = cell_axis(px, sc->origin[0], k.cell, sc->grid[0], flags);
                uint32_t cy = cell_axis(py, sc->origin[1], k.cell, sc->grid[1], flags);
                uint32_t cz = cell_axis(pz, sc->origin[2], k.cell, sc->grid[2], flags);
                key = sc->cell_base + (cz * sc->grid[1] + cy) * sc->grid[0] + cx;
            } else {
                uint32_t cx = cell_axis(px, k.origin[0], k.cell, k.grid[0], flags);
                uint32_t cy = cell_axis(py, k.origin[1], k.cell, k.grid[1], flags);
                uint32_t cz = cell_axis(pz, k.origin[2], k.cell, k.grid[2], flags);
                key = (cz * k.grid[1] + cy) * k.grid[0] + cx;
            }
        }
        // Storage is in last substep's cell order, so equal keys sit in runs
        // of adjacent lanes: one atomic per run hands out the run's ranks.
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t prev = __shfl_up(key, 1, 64);
        const bool head = valid && (lane == 0 || key != prev);
        const uint64_t heads = __ballot(head);
        const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const uint64_t mine = heads & upto;
        const uint32_t seg = mine ? 63u - (uint32_t)__clzll((unsigned long long)mine) : lane;
        const uint64_t later = heads & ~upto;
        const uint64_t vmask = __ballot(valid);
        const uint32_t vend = vmask == ~0ull ? 64u : (uint32_t)__ffsll((unsigned long long)~vmask) - 1u;
        const uint32_t next = later ? (uint32_t)__ffsll((unsigned long long)later) - 1u : vend;
        uint32_t base = 0;
        if (head) base = atomicAdd(&b.cell_count[key], next - seg);
        base = __shfl(base, seg, 64);
        if (valid) {
            b.key[s] = key;
            b.rank[s] = base + (lane - seg);
        }
        // status bits: one atomic per wavefront, not per particle (a splash
        // can put thousands of predictions outside the grid at once), and
        // none once the bits are set (atomics on one address serialise)
        const uint64_t nf = __ballot((flags & PS_STATUS_NONFINITE) != 0u);
        const uint64_t og = __ballot((flags & PS_STATUS_OUTSIDE_GRID) != 0u);
        if (lane == 0 && (nf | og)) {
            const uint32_t want = (nf ? PS_STATUS_NONFINITE : 0u) | (og ? PS_STATUS_OUTSIDE_GRID : 0u);
            if ((status_peek(&b.status->flags) & want) != want) atomicOr(&b.status->flags, want);
        }
    }
}

// ---------------------------------------------------------------------------
// a4 part 1: exclusive scan of the per-cell counts into cell offsets
// (cell c = sorted slots [off[c], off[c+1]); the reference's cellStart/End,
// reorderDataAndFindCellStartD integration_kernel.cuh:229-299), two passes
// over tiles of 4096 cells; the counts are cleared for the next build.

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *lds4)
{
    v = wave_sum(v);
    const uint32_t w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    return lds4[0] + lds4[1] + lds4[2] + lds4[3];
}

__global__ __launch_bounds__(kBlock) void scan_partials_k(const uint32_t *cnt, uint32_t *part,
                                                          uint32_t C)
{
    __shared__ uint32_t lds4[4];
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < kScanPerThread; q++) {
        uint32_t c = base + q * kBlock + threadIdx.x;
        if (c < C) sum += cnt[c];
    }
    sum = block_sum(sum, lds4);
    if (threadIdx.x == 0) part[blockIdx.x] = sum;
}

__global__ __launch_bounds__(kBlock) void scan_final_k(uint32_t *cnt, const uint32_t *part,
                                                       uint32_t *off, uint32_t C)
{
    __shared__ uint32_t tile[kScanTile + kScanTile / 16];
    __shared__ uint32_t lds4[4];
    const uint32_t tid = threadIdx.x;
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t pre = 0;
    for (uint32_t t = tid; t < blockIdx.x; t += kBlock) pre += part[t];
    pre = block_sum(pre, lds4);
    // coalesced load (and clear) into LDS; padded index avoids bank conflicts
#pragma unroll
    for (uint32_t q = 0; q < kScanPerThread; q++) {
        uint32_t idx = q * kBlock + tid;
        uint32_t c = base + idx;
        uint32_t v = 0;
        if (c < C) {
            v = cnt[c];
            cnt[c] = 0;
        }
        tile[idx + (idx >> 4)] = v;
    }
    __syncthreads();
    uint32_t v[kScanPerThread];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t q = 0; q < kScanPerThread; q++) {
        uint32_t idx = tid * kScanPerThread + q;
        v[q] = tile[idx + (idx >> 4)];
        local += v[q];
    }
    // block exclusive scan of the per-thread totals
    uint32_t incl = wave_incl_scan(local);
    const uint32_t w = tid >> 6;
    __shared__ uint32_t wtot[4];
    if ((tid & 63) == 63) wtot[w] = incl;
    __syncthreads();
    uint32_t wpre = 0;
    for (uint32_t q = 0; q < w; q++) wpre += wtot[q];
    uint32_t run = pre + wpre + (incl - local);
#pragma unroll
    for (uint32_t q = 0; q < kScanPerThread; q++) {
        uint32_t idx = tid * kScanPerThread + q;
        tile[idx + (idx >> 4)] = run;
        run += v[q];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kScanPerThread; q++) {
        uint32_t idx = q * kBlock + tid;
        uint32_t c = base + idx;
        if (c < C) off[c] = tile[idx + (idx >> 4)];
    }
    if (blockIdx.x == gridDim.x - 1 && tid == kBlock - 1) off[C] = run;
}

// a4 part 2: provisional placement by arrival rank (insertion indices only)
__global__ __launch_bounds__(kBlock) void scatter_k(DevBuffers b, uint32_t n)
{
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const uint32_t t = b.cell_off[b.key[s]] + b.rank[s];
    if (t < n)
        b.tmp_orig[t] = b.A[s].orig;
    else
        atomicOr(&b.status->flags, PS_STATUS_INTERNAL);
}

// a3 + a4 part 3: final slot = cell start + number of particles in the same
// cell with a smaller insertion index, i.e. a sort by (cell, insertion
// index): a stable sort of insertion-ordered input, as thrust::sort_by_key
// of (hash, index) in sortParticles (integration.cu:270-275) is meant to
// give.  Every per-particle array moves with the particle.
// mj_solid >= 0: also write mj (the neighbour density mass) with s_solid =
// mj_solid.
__global__ __launch_bounds__(kMaxBlock) void reorder_k(DevBuffers b, uint32_t n, float mj_solid)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t c = b.key[s];
    const Attr a = b.A[s];
    // the particle's own records go in flight before the count
    const float4 p = b.P[s], xs = b.X[s], vs = b.V[s];
    const uint32_t b0 = b.cell_off[c], b1 = b.cell_off[c + 1];
    if (b1 > n || b0 >= b1 || a.orig >= n) {
        atomicOr(&b.status->flags, PS_STATUS_INTERNAL);
        return;
    }
    // the count in groups of four independent loads (one wait per group
    // instead of one per member)
    uint32_t r = 0;
    for (uint32_t u = b0; u < b1; u += 4) {
        uint32_t t[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) t[k] = b.tmp_orig[min(u + k, b1 - 1u)];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) r += (u + k < b1 && t[k] < a.orig) ? 1u : 0u;
    }
    const uint32_t d = b0 + r;
    // scenes that are not a uniform fluid: x*.w carries the particle's XSPH
    // mass as a neighbour, 1/w for a fluid particle, -1 for any other
    b.X2[d] = mj_solid >= 0.0f
                  ? make_float4(xs.x, xs.y, xs.z, a.phase == PS_PHASE_FLUID ? 1.0f / a.w : -1.0f)
                  : xs;
    b.Q[d] = make_float4(p.x, p.y, p.z, 0.0f);
    b.V2[d] = vs;
    b.A2[d] = a;
    b.key2[d] = c;
    if (mj_solid >= 0.0f) {
        // the neighbour mass of the fluid sums, as they form it per pair
        float m = -1.0f;
        if (a.w != 0.0f) {
            m = 1.0f / a.w;
            if (a.phase != PS_PHASE_FLUID) m = m * mj_solid;
        }
        b.mj[d] = m;
        // the phase as the staged contact walk reads it
        b.phs[d] = a.w != 0.0f ? a.phase : (int32_t)((uint32_t)a.phase | 0x80000000u);
    }
    if (b.sid) b.sid2[d] = b.sid[s];
    // a scatter by insertion index: skipped when no kernel reads it
    if (b.track_slots) b.slot_of[a.orig] = d;
    // recording for the reverse pass: the same records into the tape slot
    if (b.tape_x) {
        b.tape_x[d] = xs;
        b.tape_a[d] = a;
        if (b.track_slots) b.tape_s[a.orig] = d;
    }
}


// ---------------------------------------------------------------------------
// Neighbour table (collideCell/collideCellRadius, integration_kernel.cuh:
// 303-356,481-518): every j != i with r^2 < search_r2, scanning the 3x3 rows
// of x-adjacent cells; a row's three cells are one contiguous slot range.

// Scan the nine stencil rows of particle i in order (t = (dz+1)*3 + dy+1)
// and write its table: every slot j != i of range(t) = [j0, j1) with
// |p - P[j]|^2 < search_r2, in ascending j, stored as emit(t, j) (the slot
// itself, or its LDS index in a staged unit); pad fills the last chunk: i's
// own slot, or a staged unit's sentinel record.  Returns the number found
// (the table keeps the first kmax).
// A row is tested in windows of 32 candidates without branches: each test
// sets one bit of a mask (i's own bit is cleared), and the set bits are then
// emitted lowest first, so the entries come out in ascending j as from a
// plain loop over the row, and only accepted candidates pay for the table.
// p is taken by value: a reference into b.P would be re-read after every
// table store (the compiler cannot rule out aliasing).  emit(t, j) is
// j + emit(t, 0) (an offset per stencil row), taken once per row.
// GLOB: also write the table with global slots (entries j, padding i) to
// b.nbr_glob in the same layout (the copy a recording for the reverse pass
// keeps while the staged kernels read LDS indices).
// WAVE (the staged build; every lane of the wavefront calls it, valid or
// not): per stencil row the wavefront's lanes, consecutive sorted particles,
// mostly share one grid row, so their ranges lie in [lo, hi), lo the range
// start of lane 0 and hi the range end of the last valid lane (lastv).
// When that span fits kWaveRowSlots the wavefront copies it into its own
// LDS (wl) with coalesced loads, and a lane whose range lies inside it
// reads its candidates there instead of one global load per candidate and
// lane (the loads that bound this kernel).  Any other lane or row reads
// global memory as without WAVE; the table is the same.
template <bool GLOB = false, bool WAVE = false, typename Range, typename Emit>
__device__ __forceinline__ uint32_t scan_stencil(const DevBuffers &b, const KConst &k, uint32_t i,
                                                 uint32_t pad, const float4 p, uint32_t kmax,
                                                 Range range, Emit emit, bool valid = true,
                                                 float *wl = nullptr, uint32_t lastv = 63u)
{
    uint4 *out = const_cast<uint4 *>(nbr_lane(b, i));
    uint4 *out2 = GLOB ? reinterpret_cast<uint4 *>(b.nbr_glob) + (size_t)(i >> 6) * b.kq * 64 + (i & 63)
                       : nullptr;
    uint4 buf = make_uint4(pad, pad, pad, pad);
    uint4 buf2 = make_uint4(i, i, i, i);
    uint32_t cnt = 0;
    const float r2max = k.search_r2;
    // an accepted candidate: shifted into buf, which after the 4th entry of
    // a chunk holds it in order (x first); chunks past kmax are not stored
    auto accept = [&](uint32_t e, uint32_t j) {
        buf.x = buf.y;
        buf.y = buf.z;
        buf.z = buf.w;
        buf.w = e;
        if (GLOB) {
            buf2.x = buf2.y;
            buf2.y = buf2.z;
            buf2.z = buf2.w;
            buf2.w = j;
        }
        if ((cnt & 3u) == 3u && cnt < kmax) {
            out[(size_t)(cnt >> 2) * 64] = buf;
            if (GLOB) out2[(size_t)(cnt >> 2) * 64] = buf2;
        }
        cnt++;
    };
    // the candidates [j0, j1) of one stencil row; at(j - j0) reads slot j
    auto scan_row = [&](uint32_t j0, uint32_t j1, uint32_t eoff, auto at) {
        for (uint32_t w0 = j0; w0 < j1; w0 += 32u) {
            const uint32_t we = min(w0 + 32u, j1);
            uint32_t m = 0;
            uint32_t jb = w0;
            // whole groups of four: no bound test; the loads go ahead of
            // the tests
            for (; jb + 4u <= we; jb += 4u) {
                // one address per group, immediate offsets for the four
                float4 qv[4];
#pragma unroll
                for (uint32_t u = 0; u < 4; u++) qv[u] = at(jb - j0 + u);
                uint32_t g = 0;
#pragma unroll
                for (uint32_t u = 0; u < 4; u++)
                    g |= (dist2(p.x - qv[u].x, p.y - qv[u].y, p.z - qv[u].z) < r2max ? 1u : 0u) << u;
                m |= g << (jb - w0);
            }
            // the last one to three
            if (jb < we) {
                float4 qv[3];
#pragma unroll
                for (uint32_t u = 0; u < 3; u++) qv[u] = at(min(jb + u, we - 1u) - j0);
                uint32_t g = 0;
#pragma unroll
                for (uint32_t u = 0; u < 3; u++)
                    g |= (jb + u < we && dist2(p.x - qv[u].x, p.y - qv[u].y, p.z - qv[u].z) < r2max
                              ? 1u : 0u) << u;
                m |= g << (jb - w0);
            }
            // i itself (in its own row only)
            if (i - w0 < 32u) m &= ~(1u << (i - w0));
            while (m != 0u) {
                const uint32_t u = (uint32_t)__builtin_ctz(m);
                m &= m - 1u;
                accept(w0 + u + eoff, w0 + u);
            }
        }
    };
    for (uint32_t t = 0; t < 9; t++) {
        uint32_t j0 = 0, j1 = 0;
        const bool ok = valid && range(t, j0, j1);
        if constexpr (WAVE) {
            // every lane is here: the span and the copy are the wavefront's
            const uint32_t lo = __shfl(j0, 0, 64), hi = __shfl(j1, (int)lastv, 64);
            const bool rowl = lo < hi && hi - lo <= kWaveRowSlots;
            if (rowl) {
                // the previous row's LDS reads were issued before these
                // writes, and a wavefront's LDS operations run in order
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                for (uint32_t s = threadIdx.x & 63u; s < hi - lo; s += 64u) {
                    const float4 v = b.P[lo + s];
                    wl[3u * s] = v.x;
                    wl[3u * s + 1u] = v.y;
                    wl[3u * s + 2u] = v.z;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            if (!ok) continue;
            const uint32_t eoff = emit(t, 0u);
            if (rowl && j0 >= lo && j1 <= hi) {
                const float *w3 = wl + 3u * (j0 - lo);
                scan_row(j0, j1, eoff, [&](uint32_t d) {
                    return make_float4(w3[3u * d], w3[3u * d + 1u], w3[3u * d + 2u], 0.0f);
                });
            } else {
                const float4 *q0 = b.P + j0;
                scan_row(j0, j1, eoff, [&](uint32_t d) { return q0[d]; });
            }
        } else {
            if (!ok) continue;
            const float4 *q0 = b.P + j0;
            scan_row(j0, j1, emit(t, 0u), [&](uint32_t d) { return q0[d]; });
        }
    }
    const uint32_t kept = cnt < kmax ? cnt : kmax;
    // a partial last chunk: its r entries sit at the right end of buf
    const uint32_t r = kept & 3u;
    if (r == 1u) out[(size_t)(kept >> 2) * 64] = make_uint4(buf.w, pad, pad, pad);
    else if (r == 2u) out[(size_t)(kept >> 2) * 64] = make_uint4(buf.z, buf.w, pad, pad);
    else if (r == 3u) out[(size_t)(kept >> 2) * 64] = make_uint4(buf.y, buf.z, buf.w, pad);
    if (GLOB) {
        if (r == 1u) out2[(size_t)(kept >> 2) * 64] = make_uint4(buf2.w, i, i, i);
        else if (r == 2u) out2[(size_t)(kept >> 2) * 64] = make_uint4(buf2.z, buf2.w, i, i);
        else if (r == 3u) out2[(size_t)(kept >> 2) * 64] = make_uint4(buf2.y, buf2.z, buf2.w, i);
    }
    if (valid) b.ncnt[i] = kept;
    return cnt;
}

// one status update per block: max count and overflow
__device__ __forceinline__ void neighbor_status(const DevBuffers &b, uint32_t cnt, uint32_t kmax,
                                                uint32_t *ldsw)
{
    uint32_t m = wave_max(cnt);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) ldsw[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t mm = ldsw[0];
        for (uint32_t q = 1; q < (blockDim.x >> 6); q++) mm = ldsw[q] > mm ? ldsw[q] : mm;
        // one address for every block: skip the atomics that change nothing
        if (mm > 0 && mm > status_peek(&b.status->max_neighbors)) atomicMax(&b.status->max_neighbors, mm);
        if (mm > kmax && !(status_peek(&b.status->flags) & PS_STATUS_NEIGHBOR_OVERFLOW))
            atomicOr(&b.status->flags, (uint32_t)PS_STATUS_NEIGHBOR_OVERFLOW);
    }
}

__global__ __launch_bounds__(kMaxBlock) void neighbors_k(DevBuffers b, KConst k, uint32_t n,
                                                      uint32_t kmax)
{
    __shared__ uint32_t ldsw[kMaxBlock / 64];
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    uint32_t cnt = 0;
    if (i < n) {
        uint32_t gx = k.grid[0], gy = k.grid[1], gz = k.grid[2], base = 0;
        if (k.nscenes > 1u) {
            const SceneDev &sc = b.scenes[b.sid[i]];
            gx = sc.grid[0];
            gy = sc.grid[1];
            gz = sc.grid[2];
            base = sc.cell_base;
        }
        const uint32_t key = b.key[i] - base;
        const uint32_t cx = key % gx, cy = (key / gx) % gy, cz = key / (gx * gy);
        const uint32_t x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < gx ? cx + 1 : gx - 1;
        auto range = [&](uint32_t t, uint32_t &j0, uint32_t &j1) {
            const int z = (int)cz + (int)(t / 3) - 1, y = (int)cy + (int)(t % 3) - 1;
            if (z < 0 || z >= (int)gz || y < 0 || y >= (int)gy) return false;
            const uint32_t row = base + ((uint32_t)z * gy + (uint32_t)y) * gx;
            j0 = b.cell_off[row + x0];
            j1 = min(b.cell_off[row + x1 + 1], n);
            return j0 < j1;
        };
        cnt = scan_stencil(b, k, i, i, b.P[i], kmax, range, [](uint32_t, uint32_t j) { return j; });
    }
    neighbor_status(b, cnt, kmax, ldsw);
}

// Copy a staged unit's chunks of the current data into LDS.  With THREADS
// threads, thread t handles element t % 16 of chunks c = (THREADS/16) k +
// t / 16, BATCH chunks' loads in flight at a time; a chunk's elements past
// its interval read a clamped slot and fill padding that no table entry
// refers to.  load(j) reads slot j, store(l, v) writes LDS index l.  Every
// thread of the block calls it.
template <uint32_t BATCH, uint32_t THREADS = kStageUnit, typename Load, typename Store>
__device__ __forceinline__ void stage_copy(const StageDesc *d, uint32_t n, Load load, Store store)
{
    static_assert(THREADS % kStageChunk == 0 && kStageChunks % (THREADS / kStageChunk) == 0,
                  "whole chunks per step");
    constexpr uint32_t kPerStep = THREADS / kStageChunk;  // chunks per step
    constexpr uint32_t kSteps = kStageChunks / kPerStep;
    const uint32_t nch = d->nchunks, sub = threadIdx.x % kStageChunk;
    const uint32_t c0 = threadIdx.x / kStageChunk;
#pragma unroll
    for (uint32_t k0 = 0; k0 < kSteps; k0 += BATCH) {
        if (k0 * kPerStep >= nch) break;
        uint32_t src[BATCH];
#pragma unroll
        for (uint32_t k = 0; k < BATCH; k++) {
            const uint32_t c = (k0 + k) * kPerStep + c0;
            src[k] = c < nch ? d->src[c] : 0u;
        }
        decltype(load(0u)) v[BATCH];
#pragma unroll
        for (uint32_t k = 0; k < BATCH; k++) v[k] = load(min(src[k] + sub, n - 1u));
#pragma unroll
        for (uint32_t k = 0; k < BATCH; k++) {
            const uint32_t c = (k0 + k) * kPerStep + c0;
            if (c < nch) store(c * kStageChunk + sub, v[k]);
        }
    }
    __syncthreads();
}

// Staged build (the unit's scene grid): plan the unit's intervals and its
// chunk table for the kernels that copy them into LDS, then scan the
// stencil in global memory (full occupancy; this kernel reads each
// candidate once per particle, so LDS would not save it traffic) and store
// each neighbour as its LDS index j + base - start of its interval.  A unit
// spanning more than kStageRows rows or kStageSlots slots keeps global slots
// (staged = 0).
template <bool GLOB>
__global__ __launch_bounds__(kStageUnit) void neighbors_staged_k(DevBuffers b, KConst k, uint32_t n,
                                                              uint32_t kmax)
{
    __shared__ uint32_t s_start[kStageIv], s_base[kStageIv + 1], s_len[kStageIv];
    __shared__ uint32_t s_row[kStageRows], s_xa[kStageRows], s_xb[kStageRows];
    __shared__ uint32_t s_wheads[kStageUnit / 64], s_wlast[kStageUnit / 64], s_wwalk[kStageUnit / 64];
    __shared__ uint32_t s_staged;
    __shared__ StageDesc s_desc;
    __shared__ uint32_t ldsw[kStageUnit / 64];
    __shared__ float s_wrow[kStageUnit / 64][3 * kWaveRowSlots];  // each wavefront's stencil row (x, y, z)
    const uint32_t unit = xcd_block();
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t i = unit * kStageUnit + tid;
    const bool valid = i < n;  // the valid threads are a prefix of the unit
    // the particle's grid: in a batch its scene's (cells cbase ..), and a
    // unit whose particles lie in two scenes keeps global slots
    uint32_t gx = k.grid[0], gy = k.grid[1], gz = k.grid[2], cbase = 0;
    bool one_scene = true;
    if (k.nscenes > 1u) {
        const uint32_t first = unit * kStageUnit, lastv = min(first + kStageUnit, n) - 1u;
        const uint32_t su = b.sid[first];
        one_scene = b.sid[lastv] == su;
        const SceneDev &sc = b.scenes[valid ? b.sid[i] : su];
        gx = sc.grid[0];
        gy = sc.grid[1];
        gz = sc.grid[2];
        cbase = sc.cell_base;
    }
    const uint32_t key = valid ? b.key[i] - cbase : 0u;
    // whether the particle walks its table for contacts (phase from cloth
    // up, w != 0: the sign bit of phs clear), loaded beside the key
    const bool walker = valid && b.contacts && b.phs[i] >= PS_PHASE_CLOTH;
    const uint32_t row = key / gx, x = key % gx;  // row = z gy + y
    const uint32_t ry = row % gy, rz = row / gy;
    const uint32_t x0 = x > 0 ? x - 1 : 0, x1 = x + 1 < gx ? x + 1 : gx - 1;

    // rows of the unit: runs of equal row in the (sorted) keys
    uint32_t prev = __shfl_up(row, 1, 64);
    if (lane == 63) s_wlast[w] = row;
    if (tid < kStageRows) {
        s_xa[tid] = 0xffffffffu;
        s_xb[tid] = 0u;
    }
    __syncthreads();
    if (lane == 0 && w > 0) prev = s_wlast[w - 1];
    const bool head = valid && (tid == 0 || row != prev);
    const uint64_t hm = __ballot(head);
    const uint64_t wm = __ballot(walker);
    if (lane == 0) {
        s_wheads[w] = (uint32_t)__popcll(hm);
        s_wwalk[w] = wm != 0ull ? 1u : 0u;
    }
    __syncthreads();
    uint32_t before = 0, nrows = 0, walk = 0;
    for (uint32_t q = 0; q < kStageUnit / 64; q++) {
        before += q < w ? s_wheads[q] : 0u;
        nrows += s_wheads[q];
        walk |= s_wwalk[q];
    }
    if (!one_scene) nrows = kStageRows + 1u;  // not staged
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const uint32_t q = before + (uint32_t)__popcll(hm & upto) - 1u;  // this particle's row
    if (valid && nrows <= kStageRows) {
        if (head) s_row[q] = row;
        atomicMin(&s_xa[q], x);
        atomicMax(&s_xb[q], x);
    }
    __syncthreads();

    // intervals: thread 9q + t plans target row t of unit row q
    const uint32_t niv = nrows <= kStageRows ? 9u * nrows : 0u;
    if (tid < niv) {
        const uint32_t qq = tid / 9u, t = tid % 9u;
        const uint32_t r = s_row[qq], rry = r % gy, rrz = r / gy;
        const int z = (int)rrz + (int)(t / 3) - 1, y = (int)rry + (int)(t % 3) - 1;
        uint32_t st = 0, len = 0;
        if (z >= 0 && z < (int)gz && y >= 0 && y < (int)gy) {
            const uint32_t r2 = cbase + ((uint32_t)z * gy + (uint32_t)y) * gx;
            const uint32_t xa = s_xa[qq], xb = s_xb[qq];
            st = b.cell_off[r2 + (xa > 0 ? xa - 1 : 0)];
            const uint32_t en = min(b.cell_off[r2 + (xb + 1 < gx ? xb + 1 : gx - 1) + 1], n);
            len = en > st ? en - st : 0u;
        }
        s_start[tid] = st;
        s_len[tid] = len;
    }
    __syncthreads();
    if (tid == 0) {
        // LDS bases: intervals padded to whole chunks
        uint32_t acc = 0;
        for (uint32_t iv = 0; iv < niv; iv++) {
            s_base[iv] = acc;
            acc += (s_len[iv] + kStageChunk - 1) / kStageChunk * kStageChunk;
        }
        s_base[niv] = acc;
        s_staged = (niv > 0 && acc <= kStageSlots) ? 1u : 0u;
    }
    __syncthreads();
    const bool staged = s_staged != 0u;
    const uint32_t nch = staged ? s_base[niv] / kStageChunk : 0u;
    if (tid == 0) {
        s_desc.staged = staged ? 1u : 0u;
        s_desc.nchunks = nch;
    }
    // chunk table: chunk c lies in the interval whose padded region holds it
    for (uint32_t c = tid; c < nch; c += kStageUnit) {
        uint32_t iv = 0;
        while (iv + 1 < niv && s_base[iv + 1] <= c * kStageChunk) iv++;
        s_desc.src[c] = s_start[iv] + (c * kStageChunk - s_base[iv]);
    }
    __syncthreads();
    // the descriptor the other staged kernels of this substep read
    StageDesc *desc = b.stage + unit;
    if (tid == 0) {
        desc->staged = s_desc.staged;
        desc->nchunks = nch;
        desc->walk = walk;
    }
    for (uint32_t c = tid; c < nch; c += kStageUnit) desc->src[c] = s_desc.src[c];

    uint32_t cnt = 0;
    // the wavefront's first particle and its last valid lane
    const uint32_t wfirst = unit * kStageUnit + (tid & ~63u);
    if (wfirst < n) {
        auto range = [&](uint32_t t, uint32_t &j0, uint32_t &j1) {
            const int z = (int)rz + (int)(t / 3) - 1, y = (int)ry + (int)(t % 3) - 1;
            if (z < 0 || z >= (int)gz || y < 0 || y >= (int)gy) return false;
            const uint32_t r2 = cbase + ((uint32_t)z * gy + (uint32_t)y) * gx;
            j0 = b.cell_off[r2 + x0];
            j1 = min(b.cell_off[r2 + x1 + 1], n);
            if (j0 >= j1) return false;
            if (staged) {
                // the row range lies inside its interval (by construction)
                const uint32_t iv = 9u * q + t;
                if (j0 < s_start[iv] || j1 - s_start[iv] > s_len[iv]) {
                    atomicOr(&b.status->flags, PS_STATUS_INTERNAL);
                    return false;
                }
            }
            return true;
        };
        // LDS index = slot + (base - start) of the stencil row's interval
        auto emit = [&](uint32_t t, uint32_t j) {
            const uint32_t iv = 9u * q + t;
            return staged ? j + (s_base[iv] - s_start[iv]) : j;
        };
        // a staged unit pads with its sentinel record, a global one with i
        const uint32_t lastv = min(63u, n - 1u - wfirst);
        cnt = scan_stencil<GLOB, true>(b, k, i, staged ? kStageSentinel : i,
                                       valid ? b.P[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f), kmax,
                                       range, emit, valid, s_wrow[w], lastv);
    }
    neighbor_status(b, cnt, kmax, ldsw);
}

// ---------------------------------------------------------------------------
// a5: PBF multiplier (findLambdasD, integration_kernel.cuh:521-593)
//   rho_i = W(0)/w_i + sum_j s_j W(r)/w_j,  grad_j = a d,
//   a = (spiky/rho0_i) (h-r)^2 / r,
//   lambda_i = -(rho_i/rho0_i - 1) / (sum |grad_j|^2 + |sum grad_j|^2 + eps)
// Per pair, with t = h^2 - r^2, 1/r = rsq(r^2) and q = h - r =
// fma(-r^2, 1/r, h): the density term t^3 m_j, the gradient term q^2 (1/r) d
// and |grad_j|^2 as q^4; the constants poly6 and sk = spiky/rho0_i multiply
// the finished sums (rho = fma(poly6, sum, W(0)/w_i), grad sum sk S,
// sum |grad_j|^2 = sk^2 sum q^4).  Degenerate pairs (r^2 < 1e-8) add density
// only.
// UNIFORM: every particle is fluid with inverse mass wu (no Attr gathers).

// TAPE (reverse-pass recording): also store S, D and rho per particle.
template <bool UNIFORM, bool STAGED, bool TAPE = false>
__global__ __launch_bounds__(kMaxBlock) void lambda_k(DevBuffers b, KConst k, uint32_t n, float mu)
{
    // staged: a uniform fluid stages neighbour positions only, 12 bytes
    // per slot (36 KB, so four blocks fit in a CU's LDS where 16-byte slots
    // allow three); other scenes stage (x, y, z, mj), a particle with w = 0
    // as a position at infinity, which every pair test rejects as the
    // global path leaves it out.  Then the sentinel record; two lanes per
    // particle (2 kStageUnit threads), lane hh summing half hh
    __shared__ float lds3[STAGED && UNIFORM ? 3 * (kStageSlots + 1) : 1];
    __shared__ float4 lds4[STAGED && !UNIFORM ? kStageSlots + 1 : 1];
    const uint32_t unit = xcd_block();
    const uint32_t hh = STAGED ? (threadIdx.x & 1u) : 0u;
    const uint32_t i = STAGED ? unit * kStageUnit + (threadIdx.x >> 1) : unit * blockDim.x + threadIdx.x;
    bool staged = false;
    if (STAGED) {
        const StageDesc *d = b.stage + unit;
        staged = d->staged != 0u;
        if (UNIFORM) {
            if (staged && threadIdx.x < 3) lds3[3 * kStageSentinel + threadIdx.x] = __builtin_inff();
            if (staged)
                stage_copy<6, 2 * kStageUnit>(d, n, [&](uint32_t j) { return b.P[j]; },
                                              [&](uint32_t l, const float4 &v) {
                                                  lds3[3 * l] = v.x;
                                                  lds3[3 * l + 1] = v.y;
                                                  lds3[3 * l + 2] = v.z;
                                              });
        } else {
            const float inf = __builtin_inff();
            if (staged && threadIdx.x == 0) lds4[kStageSentinel] = make_float4(inf, inf, inf, 0.0f);
            if (staged)
                stage_copy<6, 2 * kStageUnit>(d, n, [&](uint32_t j) { return PM{b.P[j], b.mj[j]}; },
                                              [&](uint32_t l, const PM &v) {
                                                  lds4[l] = v.m < 0.0f ? make_float4(inf, inf, inf, 0.0f)
                                                                       : make_float4(v.p.x, v.p.y, v.p.z, v.m);
                                              });
        }
    }
    if (i >= n) return;
    float *Pw = reinterpret_cast<float *>(b.P);
    const Attr a = b.A[i];
    if (!UNIFORM && a.phase != PS_PHASE_FLUID) {
        Pw[4 * (size_t)i + 3] = 0.0f;
        if (TAPE) {
            const float4 pn = b.P[i];
            b.tape_lq[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            b.tape_rho[i] = 0.0f;
            b.tape_pit[i] = make_float4(pn.x, pn.y, pn.z, 0.0f);
        }
        return;
    }
    const float4 p = b.P[i];
    const float r0 = a.rho0;
    const float sk = k.spiky * (1.0f / r0);
    const uint32_t c = b.ncnt[i];
    // the sums of one half of the table
    struct Sums {
        float rho, gix, giy, giz, den;
    };
    const float self = k.w_zero * (UNIFORM ? mu : (1.0f / a.w));
    // mj: 1/w_j, times s_solid for a non-fluid neighbour
    auto body = [&](Sums &s, const float4 &pj, float mj) {
        float dx = p.x - pj.x, dy = p.y - pj.y, dz = p.z - pj.z;
        float r2 = dist2(dx, dy, dz);
        if (!(r2 < k.h2)) return;
        float t = k.h2 - r2;
        s.rho = __builtin_fmaf((t * t) * t, mj, s.rho);
        if (r2 < kDegenerateR2) return;  // density only
        const float ri = rsq(r2);
        const float qq = __builtin_fmaf(-r2, ri, k.h);
        const float q2 = qq * qq;
        const float aa = q2 * ri;
        s.den = __builtin_fmaf(q2, q2, s.den);
        s.gix = __builtin_fmaf(aa, dx, s.gix);
        s.giy = __builtin_fmaf(aa, dy, s.giy);
        s.giz = __builtin_fmaf(aa, dz, s.giz);
    };
    Sums s0 = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, s1 = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (STAGED) {
        // this lane's half (from LDS, or from global memory for a unit that
        // was not staged); the partner lane (lane ^ 1) holds the other.
        // Both lanes form half 0 + half 1 (addition commutes)
        Sums mine = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (UNIFORM && staged)
            for_neighbors_half<false>(
                b, i, c, hh,
                [&](uint32_t l) {
                    // byte offset 12 l by the full-rate 24-bit multiply (l <= kStageSentinel)
                    const float *q = reinterpret_cast<const float *>(reinterpret_cast<const char *>(lds3) +
                                                                     __umul24(l, 12u));
                    return make_float4(q[0], q[1], q[2], 0.0f);
                },
                [&](uint32_t, const float4 &pj) { body(mine, pj, mu); });
        else if (UNIFORM)
            for_neighbors_half(
                b, i, c, hh, [&](uint32_t j) { return b.P[j]; },
                [&](uint32_t, const float4 &pj) { body(mine, pj, mu); });
        else if (staged)
            for_neighbors_half<false>(b, i, c, hh, [&](uint32_t l) { return lds4[l]; },
                                      [&](uint32_t, const float4 &pj) { body(mine, pj, pj.w); });
        else
            for_neighbors_half(
                b, i, c, hh, [&](uint32_t j) { return PM{b.P[j], b.mj[j]}; },
                [&](uint32_t, const PM &d) {
                    if (d.m >= 0.0f) body(mine, d.p, d.m);
                });
        const Sums other = {__shfl_xor(mine.rho, 1, 64), __shfl_xor(mine.gix, 1, 64),
                            __shfl_xor(mine.giy, 1, 64), __shfl_xor(mine.giz, 1, 64),
                            __shfl_xor(mine.den, 1, 64)};
        s0 = hh ? other : mine;
        s1 = hh ? mine : other;
    } else if (UNIFORM) {
        for_neighbors_split(
            b, i, c, [&](uint32_t j) { return b.P[j]; },
            [&](uint32_t, const float4 &pj) { body(s0, pj, mu); },
            [&](uint32_t, const float4 &pj) { body(s1, pj, mu); });
    } else {
        // the neighbour's mass from the sort's mj (-1: w = 0, left out)
        auto gen = [&](Sums &s, const PM &d) {
            if (d.m < 0.0f) return;
            body(s, d.p, d.m);
        };
        for_neighbors_split(
            b, i, c, [&](uint32_t j) { return PM{b.P[j], b.mj[j]}; },
            [&](uint32_t, const PM &d) { gen(s0, d); }, [&](uint32_t, const PM &d) { gen(s1, d); });
    }
    const float rho = __builtin_fmaf(k.poly6, s0.rho + s1.rho, self);
    const float gix = sk * (s0.gix + s1.gix), giy = sk * (s0.giy + s1.giy), giz = sk * (s0.giz + s1.giz);
    const float den = __builtin_fmaf(sk * sk, s0.den + s1.den, dist2(gix, giy, giz));
    float C = (rho / r0) - 1.0f;
    if (STAGED && hh) return;  // the pair's lane 0 writes
    const float lam = -C / (den + k.relax);
    Pw[4 * (size_t)i + 3] = lam;
    if (TAPE) {
        b.tape_pit[i] = make_float4(p.x, p.y, p.z, lam);
        b.tape_lq[i] = make_float4(gix, giy, giz, den);
        b.tape_rho[i] = rho;
    }
}


// ---------------------------------------------------------------------------
// One Jacobi iteration for every particle: contacts (a7), fluid delta-p
// (a6), distance (a9), shape goal (a12) summed in that order from the same
// snapshot P, then walls (a8) and pins (a10); result to Q.

// TAPE (reverse-pass recording): also store the fluid sum f (before the
// division by rho0) and the contact count per particle.
template <uint32_t F, bool UNIFORM, uint32_t NC, bool STAGED, bool TAPE = false>
__global__ __launch_bounds__(kMaxBlock) void solve_k(DevBuffers b, KConst k, uint32_t n, float wu,
                                                  uint32_t it)
{
    // staged: two lanes per particle (2 kStageUnit threads), lane hh summing
    // half hh of the fluid sum; lane 0 finishes the particle.  The LDS holds
    // the unit's slots and the sentinel record
    // In a scene that is not a uniform fluid a particle with w = 0 is
    // staged as a position at infinity (the fluid sum leaves it out, as the
    // global path does).  With contacts, in a unit where some particle
    // walks its table for them (the neighbour build's StageDesc::walk), each
    // slot's phase is staged as well (lds_ph, from the sort's contiguous phs:
    // the sign bit set for a particle with w = 0), so the counting walk runs
    // from LDS, reading the true position of such a particle from global
    // memory; the corrections gather through the unit's chunk table (s_src).
    __shared__ float4 lds[STAGED ? kStageSlots + 1 : 1];
    __shared__ int32_t lds_ph[STAGED && (F & F_CONTACT) ? kStageSlots + 1 : 1];
    __shared__ uint32_t s_src[STAGED && (F & F_CONTACT) ? kStageChunks : 1];
    // the contact lists of the block (global-slot variants with contacts):
    // kContactList entries per thread, entry q of thread t at q blockDim + t
    extern __shared__ uint32_t contact_lds[];
    const uint32_t unit = xcd_block();
    const uint32_t hh = STAGED ? (threadIdx.x & 1u) : 0u;
    const uint32_t i = STAGED ? unit * kStageUnit + (threadIdx.x >> 1) : unit * blockDim.x + threadIdx.x;
    bool staged = false;
    if (STAGED) {
        const StageDesc *d = b.stage + unit;
        // one class of units per launch (the other class's blocks leave)
        if (b.solve_units && ((d->walk != 0u) != (b.solve_units == 2u))) return;
        staged = d->staged != 0u;
        const float inf = __builtin_inff();
        if (staged && threadIdx.x == 0) lds[kStageSentinel] = make_float4(inf, inf, inf, 0.0f);
        if ((F & F_CONTACT) && staged && threadIdx.x == 0) lds_ph[kStageSentinel] = 0;
        if ((F & F_CONTACT) && staged)
            for (uint32_t q = threadIdx.x; q < d->nchunks; q += blockDim.x) s_src[q] = d->src[q];
        if (UNIFORM && staged)
            stage_copy<6, 2 * kStageUnit>(d, n, [&](uint32_t j) { return b.P[j]; },
                                          [&](uint32_t l, const float4 &v) { lds[l] = v; });
        else if ((F & F_CONTACT) && staged && d->walk)
            stage_copy<6, 2 * kStageUnit>(d, n, [&](uint32_t j) { return PMH{b.P[j], b.mj[j], b.phs[j]}; },
                                          [&](uint32_t l, const PMH &v) {
                                              lds[l] = v.m < 0.0f ? make_float4(inf, inf, inf, 0.0f) : v.p;
                                              lds_ph[l] = v.ph;
                                          });
        else if (staged)
            stage_copy<6, 2 * kStageUnit>(d, n, [&](uint32_t j) { return PM{b.P[j], b.mj[j]}; },
                                          [&](uint32_t l, const PM &v) {
                                              lds[l] = v.m < 0.0f ? make_float4(inf, inf, inf, 0.0f) : v.p;
                                          });
    }
    // the storage slot of table entry e (an LDS index in a staged unit; its
    // sentinel padding, never visited, maps to i)
    auto slot = [&](uint32_t e) -> uint32_t {
        if (!(STAGED && (F & F_CONTACT)) || !staged) return e;
        return e >= kStageSlots ? i : s_src[e >> 4] + (e & (kStageChunk - 1u));
    };
    if (i >= n) return;
    const float4 p = b.P[i];
    const Attr a = b.A[i];
    const float px = p.x, py = p.y, pz = p.z;
    const int32_t ph = UNIFORM ? PS_PHASE_FLUID : a.phase;
    const float wi = UNIFORM ? wu : a.w;
    const uint32_t c = b.ncnt[i];
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    float4 tape = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // (f, contact count) when TAPE

    // a7: particle contacts + friction (collideD, integration_kernel.cuh:358-462)
    if (F & F_CONTACT) {
        if (ph >= PS_PHASE_CLOTH && wi != 0.0f) {
            // The corrections are summed in the two halves of the table, each
            // in table order, then added half 0 + half 1 (the oracle's
            // order).  The counting walk lists the contacts in the block's
            // LDS (entry m of thread t at contact_lds[m blockDim + t]) and
            // the corrections visit only those, or walk the table again when
            // the list overflowed.  Global-slot variants: one lane walks both
            // halves (pipelined two chunks at a time), listing slot | half
            // << 31.  Staged: each of the pair's two lanes counts and
            // corrects its own half, listing LDS indices (global slots in a
            // unit that was not staged), and the halves meet by a shuffle.
            auto pa = [&](uint32_t e) {
                const uint32_t j = slot(e);
                return PA{b.P[j], b.A[j]};
            };
            auto is_contact = [&](const PA &d) {
                if (ph > PS_PHASE_SOLID && ph == d.a.phase) return false;
                const float4 qj = d.p;
                float dx = px - qj.x, dy = py - qj.y, dz = pz - qj.z;
                float r2 = ((dx * dx) + (dy * dy)) + (dz * dz);
                return r2 < k.collide_r2;
            };
            uint32_t *clist = contact_lds + threadIdx.x;
            const uint32_t cstride = blockDim.x;
            constexpr uint32_t kList = STAGED ? kHalfContactList : kContactList;
            uint32_t nc = 0;
            if constexpr (STAGED) {
                if (staged) {
                    // from the staged records: position and phase (a
                    // particle with w = 0 is staged at infinity and marked
                    // by the phase's sign bit: its position from memory)
                    for_neighbors_half_chunked(
                        b, i, c, hh, [&](uint32_t e) { return PH{lds[e], lds_ph[e]}; },
                        [&](uint32_t e, const PH &d) {
                            const int32_t phj = d.ph & 0x7fffffff;
                            if (ph > PS_PHASE_SOLID && ph == phj) return;
                            float qx = d.p.x, qy = d.p.y, qz = d.p.z;
                            if (d.ph < 0) {
                                // a rare path; read through a volatile
                                // pointer, which the compiler cannot merge
                                // with the staged record's (merged, the
                                // records went through scratch memory)
                                const volatile float *g = &b.P[slot(e)].x;
                                qx = g[0];
                                qy = g[1];
                                qz = g[2];
                            }
                            float dx = px - qx, dy = py - qy, dz = pz - qz;
                            float r2 = ((dx * dx) + (dy * dy)) + (dz * dz);
                            if (r2 < k.collide_r2) {
                                if (nc < kList) clist[nc * cstride] = e;
                                nc++;
                            }
                        });
                } else {
                    for_neighbors_half_chunked(b, i, c, hh, pa, [&](uint32_t e, const PA &d) {
                        if (is_contact(d)) {
                            if (nc < kList) clist[nc * cstride] = e;
                            nc++;
                        }
                    });
                }
            } else {
                for_neighbors_chunked_h(b, i, c, pa, [&](uint32_t h, uint32_t j, const PA &d) {
                    if (is_contact(d)) {
                        if (nc < kList) clist[nc * cstride] = j | (h << 31);
                        nc++;
                    }
                });
            }
            const uint32_t nmine = nc;  // contacts this lane found (and listed up to kList)
            if (STAGED) nc = nc + __shfl_xor(nc, 1, 64);
            if (TAPE) tape.w = __uint_as_float(nc);
            if (nc > 0) {
                const float fnc = (float)nc;
                const float ewi = wi * ps_exp(py);
                const float4 xi = b.X[i];
                struct C3 {
                    float x, y, z;
                };
                auto correct = [&](C3 &cs, uint32_t e, const PA &d) {
                    const Attr aj = d.a;
                    const int32_t pj = aj.phase;
                    if (ph > PS_PHASE_SOLID && ph == pj) return;
                    const float4 qj = d.p;
                    float dx = px - qj.x, dy = py - qj.y, dz = pz - qj.z;
                    float r2 = ((dx * dx) + (dy * dy)) + (dz * dz);
                    if (!(r2 < k.collide_r2)) return;
                    float dist = sqrtf(r2);
                    if (dist < 1e-6f) return;
                    const float wj = aj.w;
                    float cw = wi, cw2 = wj;
                    const bool solid = ph >= PS_PHASE_SOLID && pj >= PS_PHASE_SOLID;
                    if (solid) {
                        cw = ewi;
                        cw2 = wj * ps_exp(qj.y);
                    }
                    float wsum = cw + cw2;
                    if (wsum == 0.0f) return;
                    float mag = dist - k.collide_dist;
                    float sc = (((mag / wsum) / dist) * cw) / fnc;
                    float ex = -(sc * dx), ey = -(sc * dy), ez = -(sc * dz);
                    cs.x = cs.x + ex;
                    cs.y = cs.y + ey;
                    cs.z = cs.z + ez;
                    if (!solid) return;
                    float sc2 = (((mag / wsum) / dist) * cw2) / fnc;
                    float fx = sc2 * dx, fy = sc2 * dy, fz = sc2 * dz;
                    const float4 xj = b.X[slot(e)];
                    float rx = ((px + ex) - xi.x) - ((qj.x + fx) - xj.x);
                    float ry = ((py + ey) - xi.y) - ((qj.y + fy) - xj.y);
                    float rz = ((pz + ez) - xi.z) - ((qj.z + fz) - xj.z);
                    float nx = dx / dist, ny = dy / dist, nz = dz / dist;
                    float dn = ((rx * nx) + (ry * ny)) + (rz * nz);
                    float tx = rx - (dn * nx), ty = ry - (dn * ny), tz = rz - (dn * nz);
                    float lt = sqrtf(((tx * tx) + (ty * ty)) + (tz * tz));
                    if (lt < 0.001f) return;
                    float share = cw / wsum;
                    if (lt < (k.mu_s * dist)) {
                        cs.x = cs.x - (tx * share);
                        cs.y = cs.y - (ty * share);
                        cs.z = cs.z - (tz * share);
                    } else {
                        float f = (k.mu_k * dist) / lt;
                        if (f > 1.0f) f = 1.0f;
                        cs.x = cs.x - ((tx * f) * share);
                        cs.y = cs.y - ((ty * f) * share);
                        cs.z = cs.z - ((tz * f) * share);
                    }
                };
                C3 ch0 = {0.0f, 0.0f, 0.0f}, ch1 = {0.0f, 0.0f, 0.0f};  // the halves
                if (nmine <= kList) {
                    // the listed contacts in order, four records in flight
                    for (uint32_t q0 = 0; q0 < nmine; q0 += 4u) {
                        uint32_t jq[4];
                        PA dq[4];
#pragma unroll
                        for (uint32_t u = 0; u < 4; u++) jq[u] = clist[min(q0 + u, nmine - 1u) * cstride];
#pragma unroll
                        for (uint32_t u = 0; u < 4; u++) dq[u] = pa(STAGED ? jq[u] : (jq[u] & 0x7fffffffu));
#pragma unroll
                        for (uint32_t u = 0; u < 4; u++)
                            if (q0 + u < nmine) {
                                if (STAGED || !(jq[u] >> 31))
                                    correct(ch0, jq[u] & (STAGED ? ~0u : 0x7fffffffu), dq[u]);
                                else
                                    correct(ch1, jq[u] & 0x7fffffffu, dq[u]);
                            }
                    }
                } else if (STAGED) {
                    for_neighbors_half_chunked(b, i, c, hh, pa,
                                               [&](uint32_t e, const PA &d) { correct(ch0, e, d); });
                } else {
                    for_neighbors_chunked_h(b, i, c, pa, [&](uint32_t h, uint32_t j, const PA &d) {
                        if (h)
                            correct(ch1, j, d);
                        else
                            correct(ch0, j, d);
                    });
                }
                if (STAGED) {
                    // this lane's half is in ch0; the partner lane holds the other
                    const C3 other = {__shfl_xor(ch0.x, 1, 64), __shfl_xor(ch0.y, 1, 64), __shfl_xor(ch0.z, 1, 64)};
                    if (hh) {
                        ch1 = ch0;
                        ch0 = other;
                    } else {
                        ch1 = other;
                    }
                }
                ax = ax + (ch0.x + ch1.x);
                ay = ay + (ch0.y + ch1.y);
                az = az + (ch0.z + ch1.z);
            }
        }
    }

    // a6: PBF delta-p with tensile correction s_corr = -k h^2 (W/W(dq h))^n
    // (solveFluidsD :596-642); k.k_corr already holds k h^2 (or k).  Per
    // pair, with t = h^2 - r^2, 1/r = rsq(r^2) and q = h - r =
    // fma(-r^2, 1/r, h): W/W(dq h) = s^3 with s = (h^2 - r^2)/(h^2 - (dq h)^2)
    // = fma(-w_scale, r^2, w_scale_h2), coef = fma(-k, (W/W(dq h))^n,
    // l_i + l_j), and the term -coef q^2 (1/r) d; spiky multiplies the
    // finished sum
    if (F & F_FLUID) {
        if (ph == PS_PHASE_FLUID) {
            const float li = p.w, ir0 = 1.0f / a.rho0;
            // r^2 < 1e-8: gradient (0, -0.001 spiky h^2, 0) (integration_kernel.cuh:625-626)
            const float gdeg = (k.h * k.h) * 0.001f;
            // the sum of one half of the table; deg: a degenerate pair was
            // left out, and the half must be formed again by exact()
            struct Sums {
                float fx, fy, fz;
                bool deg;
            };
            // w_scale_h2 held in a vector register for the loop: an FMA may
            // read one scalar operand only, and w_scale is the other
            float wsh2 = k.w_scale_h2;
            asm volatile("" : "+v"(wsh2));
            auto coef_of = [&](float r2, float lj) {
                const float sw = __builtin_fmaf(-k.w_scale, r2, wsh2);
                const float x = (sw * sw) * sw;
                return __builtin_fmaf(-k.k_corr, ipow(x, NC ? NC : k.n_corr), li + lj);
            };
            auto term = [&](Sums &s, float coef, float r2, float dx, float dy, float dz) {
                const float ri = rsq(r2);
                const float qq = __builtin_fmaf(-r2, ri, k.h);
                const float g = coef * ((qq * qq) * ri);
                s.fx = __builtin_fmaf(-g, dx, s.fx);
                s.fy = __builtin_fmaf(-g, dy, s.fy);
                s.fz = __builtin_fmaf(-g, dz, s.fz);
            };
            // every pair but a degenerate one (essentially never present)
            auto fast = [&](Sums &s, const float4 &pj) {
                float dx = px - pj.x, dy = py - pj.y, dz = pz - pj.z;
                float r2 = dist2(dx, dy, dz);
                if (!(r2 < k.h2)) return;
                const bool dg = r2 < kDegenerateR2;
                s.deg = s.deg || dg;
                if (dg) return;
                term(s, coef_of(r2, pj.w), r2, dx, dy, dz);
            };
            // the reference's sum, degenerate pairs in their place
            auto exact = [&](Sums &s, const float4 &pj) {
                float dx = px - pj.x, dy = py - pj.y, dz = pz - pj.z;
                float r2 = dist2(dx, dy, dz);
                if (!(r2 < k.h2)) return;
                const float coef = coef_of(r2, pj.w);
                if (r2 < kDegenerateR2)
                    s.fy = s.fy - (coef * gdeg);
                else
                    term(s, coef, r2, dx, dy, dz);
            };
            Sums s0 = {0.0f, 0.0f, 0.0f, false}, s1 = {0.0f, 0.0f, 0.0f, false};
            if constexpr (STAGED) {
                // this lane's half; the partner lane (lane ^ 1) holds the other
                Sums mine = {0.0f, 0.0f, 0.0f, false};
                auto from_lds = [&](uint32_t l) { return lds[l]; };
                auto from_glob = [&](uint32_t j) { return b.P[j]; };
                // a unit that was not staged, in a scene that is not a uniform fluid
                auto from_glob_m = [&](uint32_t j) { return PM{b.P[j], b.mj[j]}; };
                if (staged)
                    for_neighbors_half<false>(b, i, c, hh, from_lds,
                                              [&](uint32_t, const float4 &pj) { fast(mine, pj); });
                else if (UNIFORM)
                    for_neighbors_half(b, i, c, hh, from_glob,
                                       [&](uint32_t, const float4 &pj) { fast(mine, pj); });
                else
                    for_neighbors_half(b, i, c, hh, from_glob_m, [&](uint32_t, const PM &d) {
                        if (d.m >= 0.0f) fast(mine, d.p);
                    });
                if (mine.deg) {
                    mine = {0.0f, 0.0f, 0.0f, false};
                    auto ex = [&](uint32_t, const float4 &pj) { exact(mine, pj); };
                    if (staged)
                        for_neighbors_half_plain(b, i, c, hh, from_lds, ex);
                    else if (UNIFORM)
                        for_neighbors_half_plain(b, i, c, hh, from_glob, ex);
                    else
                        for_neighbors_half_plain(b, i, c, hh, from_glob_m, [&](uint32_t, const PM &d) {
                            if (d.m >= 0.0f) exact(mine, d.p);
                        });
                }
                const Sums other = {__shfl_xor(mine.fx, 1, 64), __shfl_xor(mine.fy, 1, 64),
                                    __shfl_xor(mine.fz, 1, 64), false};
                s0 = hh ? other : mine;
                s1 = hh ? mine : other;
            } else if (UNIFORM) {
                auto gather = [&](uint32_t j) { return b.P[j]; };
                for_neighbors_split(
                    b, i, c, gather, [&](uint32_t, const float4 &pj) { fast(s0, pj); },
                    [&](uint32_t, const float4 &pj) { fast(s1, pj); });
                if (s0.deg || s1.deg) {
                    s0 = {0.0f, 0.0f, 0.0f, false};
                    s1 = {0.0f, 0.0f, 0.0f, false};
                    for_neighbors_half_plain(b, i, c, 0u, gather,
                                             [&](uint32_t, const float4 &pj) { exact(s0, pj); });
                    for_neighbors_half_plain(b, i, c, 1u, gather,
                                             [&](uint32_t, const float4 &pj) { exact(s1, pj); });
                }
            } else {
                // neighbours with w = 0 are left out (mj < 0)
                auto gather = [&](uint32_t j) { return PM{b.P[j], b.mj[j]}; };
                for_neighbors_split(
                    b, i, c, gather,
                    [&](uint32_t, const PM &d) {
                        if (d.m >= 0.0f) fast(s0, d.p);
                    },
                    [&](uint32_t, const PM &d) {
                        if (d.m >= 0.0f) fast(s1, d.p);
                    });
                if (s0.deg || s1.deg) {
                    s0 = {0.0f, 0.0f, 0.0f, false};
                    s1 = {0.0f, 0.0f, 0.0f, false};
                    for_neighbors_half_plain(b, i, c, 0u, gather, [&](uint32_t, const PM &d) {
                        if (d.m >= 0.0f) exact(s0, d.p);
                    });
                    for_neighbors_half_plain(b, i, c, 1u, gather, [&](uint32_t, const PM &d) {
                        if (d.m >= 0.0f) exact(s1, d.p);
                    });
                }
            }
            const float fx = k.spiky * (s0.fx + s1.fx), fy = k.spiky * (s0.fy + s1.fy),
                        fz = k.spiky * (s0.fz + s1.fz);
            ax = ax + (fx * ir0);
            ay = ay + (fy * ir0);
            az = az + (fz * ir0);
            if (TAPE) {
                tape.x = fx;
                tape.y = fy;
                tape.z = fz;
            }
        }
    }

    // a9: distance constraints, gathered over this particle's incident list
    // in order.  Each is a chain of dependent loads (index, pair, slot,
    // position, attributes); staged, the particle's two lanes take two at a
    // time, kR such rounds at once, each level of the chains going out for
    // all of them together, and lane 0 subtracts the terms in constraint
    // order
    if (F & F_DIST) {
        const uint32_t o = a.orig;
        const uint32_t a0 = b.dist_off[o], a1 = b.dist_off[o + 1];
        if (a1 > a0 && wi != 0.0f) {
            float cx = 0.0f, cy = 0.0f, cz = 0.0f;
            constexpr uint32_t kStep = STAGED ? 2u : 1u, kR = 3;
            for (uint32_t q0 = a0; q0 < a1; q0 += kStep * kR) {  // the same trip count in both lanes
                uint32_t cid[kR], j[kR];
                bool ok[kR];
#pragma unroll
                for (uint32_t r = 0; r < kR; r++) {
                    const uint32_t q = q0 + kStep * r + (STAGED ? hh : 0u);
                    ok[r] = q < a1;
                    cid[r] = b.dist_adj[ok[r] ? q : a0];
                }
                uint2 pr[kR];
#pragma unroll
                for (uint32_t r = 0; r < kR; r++) pr[r] = b.dist_pairs[cid[r]];
#pragma unroll
                for (uint32_t r = 0; r < kR; r++) j[r] = b.slot_of[pr[r].x == o ? pr[r].y : pr[r].x];
                float4 pj[kR];
                float wj[kR];
                float2 rs[kR];
#pragma unroll
                for (uint32_t r = 0; r < kR; r++) {
                    pj[r] = b.P[j[r]];
                    wj[r] = b.A[j[r]].w;
                    rs[r] = b.dist_rs[cid[r]];
                }
#pragma unroll
                for (uint32_t r = 0; r < kR; r++) {
                    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
                    bool use = false;
                    float dx = px - pj[r].x, dy = py - pj[r].y, dz = pz - pj[r].z;
                    float dist = sqrtf(((dx * dx) + (dy * dy)) + (dz * dz));
                    float wsum = wi + wj[r];
                    if (ok[r] && !(dist < 1e-6f || wsum == 0.0f)) {
                        float sc = ((((dist - rs[r].x) / wsum) / dist) * wi) * rs[r].y;
                        tx = sc * dx;
                        ty = sc * dy;
                        tz = sc * dz;
                        use = true;
                    }
                    if (use) {
                        cx = cx - tx;
                        cy = cy - ty;
                        cz = cz - tz;
                    }
                    if (STAGED) {
                        // lane 0: then its partner's constraint of the round
                        const float ox = __shfl_xor(tx, 1, 64), oy = __shfl_xor(ty, 1, 64),
                                    oz = __shfl_xor(tz, 1, 64);
                        if (__shfl_xor(use ? 1 : 0, 1, 64)) {
                            cx = cx - ox;
                            cy = cy - oy;
                            cz = cz - oz;
                        }
                    }
                }
            }
            const float occ = (float)b.occ[o];
            ax = ax + (cx / occ);
            ay = ay + (cy / occ);
            az = az + (cz / occ);
        }
    }

    // staged: the pair's lane 0 finishes the particle
    if (STAGED && hh) return;

    // a12: shape-matching goal pull
    if (F & F_SHAPE) {
        const uint32_t o = a.orig;
        const uint32_t bo = b.body_of[o];
        if (bo) {
            const float4 g = b.body_goal[b.member_of[o]];
            const float st = b.body_stiff[bo - 1];
            ax = ax + ((g.x - px) * st);
            ay = ay + ((g.y - py) * st);
            az = az + ((g.z - pz) * st);
        }
    }

    float nxp = px + ax, nyp = py + ay, nzp = pz + az;
    if (k.dims == 2u) nzp = pz;
    const float4 xs = (ph >= PS_PHASE_CLOTH) ? b.X[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint64_t iter = b.status->iter + it;
    float4 wrec;
    uint32_t set = walls(k, wall_box(k, b, i, a.orig), ph, xs, nxp, nyp, nzp, iter, wrec);
    if (F & F_PIN) {
        const uint32_t pin = b.pin_of[a.orig];
        if (pin) {
            const float4 t = b.pin_xyz[pin - 1];
            nxp = t.x;
            nyp = t.y;
            nzp = t.z;
            set |= 7u | kPinned;
        }
    }
    if (TAPE) {
        b.clamp_rec[i] = (uint8_t)set;
        if (set & kWallFriction) b.wall_rec[i] = wrec;
        b.tape_f[i] = tape;
    }
    b.Q[i] = make_float4(nxp, nyp, nzp, 0.0f);
    if (TAPE && b.tape_q) b.tape_q[i] = make_float4(nxp, nyp, nzp, 0.0f);
}

// ---------------------------------------------------------------------------
// a12: shape matching per body (one wavefront per body): centre of mass, A_pq,
// rotation by the fixed-point quaternion iteration (Mueller, Bender,
// Chentanez, Macklin, MIG 2016) with a first-order quaternion step, goals.

__device__ __forceinline__ void quat_to_mat(float w, float x, float y, float z, float *m)
{
    m[0] = 1.0f - (2.0f * ((y * y) + (z * z)));
    m[1] = 2.0f * ((x * y) - (w * z));
    m[2] = 2.0f * ((x * z) + (w * y));
    m[3] = 2.0f * ((x * y) + (w * z));
    m[4] = 1.0f - (2.0f * ((x * x) + (z * z)));
    m[5] = 2.0f * ((y * z) - (w * x));
    m[6] = 2.0f * ((x * z) - (w * y));
    m[7] = 2.0f * ((y * z) + (w * x));
    m[8] = 1.0f - (2.0f * ((x * x) + (y * y)));
}

// The sum of v over the 64 lanes in pairwise order, v[l] + v[l + s] for
// s = 32, 16, ..., 1 (the oracle's lane_tree_sum), returned to every lane.
__device__ __forceinline__ float wave_tree_sum(float v)
{
#pragma unroll
    for (uint32_t s = 32; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
    return __shfl(v, 0, 64);
}

// One wavefront per body: member q (of a0..a1) belongs to lane (q - a0) %
// 64, each lane sums its members in order and the lanes are added pairwise;
// every lane then runs the same rotation iteration on the same sums, and
// the members' goals are written lane-parallel.
__global__ __launch_bounds__(kBlock) void shape_k(DevBuffers b, uint32_t nbody)
{
    const uint32_t bd = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (bd >= nbody) return;  // the whole wavefront leaves together
    const uint32_t a0 = b.body_off[bd], a1 = b.body_off[bd + 1];
    const float M = b.body_mass[bd];
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    for (uint32_t q = a0 + lane; q < a1; q += 64u) {
        const uint32_t t = b.slot_of[b.body_idx[q]];
        const float m = 1.0f / b.A[t].w;
        const float4 p = b.P[t];
        cx = cx + (m * p.x);
        cy = cy + (m * p.y);
        cz = cz + (m * p.z);
    }
    cx = wave_tree_sum(cx) / M;
    cy = wave_tree_sum(cy) / M;
    cz = wave_tree_sum(cz) / M;
    float A[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (uint32_t q = a0 + lane; q < a1; q += 64u) {
        const uint32_t t = b.slot_of[b.body_idx[q]];
        const float m = 1.0f / b.A[t].w;
        const float4 p = b.P[t];
        const float ux = m * (p.x - cx), uy = m * (p.y - cy), uz = m * (p.z - cz);
        const float4 r = b.body_rest[q];
        A[0] = A[0] + (ux * r.x); A[1] = A[1] + (ux * r.y); A[2] = A[2] + (ux * r.z);
        A[3] = A[3] + (uy * r.x); A[4] = A[4] + (uy * r.y); A[5] = A[5] + (uy * r.z);
        A[6] = A[6] + (uz * r.x); A[7] = A[7] + (uz * r.y); A[8] = A[8] + (uz * r.z);
    }
#pragma unroll
    for (int e = 0; e < 9; e++) A[e] = wave_tree_sum(A[e]);
    float4 qt = b.body_q[bd];
    float qw = qt.x, qx = qt.y, qy = qt.z, qz = qt.w;
    for (int it = 0; it < 10; it++) {
        float m[9];
        quat_to_mat(qw, qx, qy, qz, m);
        float ox = 0.0f, oy = 0.0f, oz = 0.0f, dot = 0.0f;
        for (int col = 0; col < 3; col++) {
            float rx = m[col], ry = m[3 + col], rz = m[6 + col];
            float ax_ = A[col], ay_ = A[3 + col], az_ = A[6 + col];
            ox = ox + ((ry * az_) - (rz * ay_));
            oy = oy + ((rz * ax_) - (rx * az_));
            oz = oz + ((rx * ay_) - (ry * ax_));
            dot = dot + (((rx * ax_) + (ry * ay_)) + (rz * az_));
        }
        float den = (dot < 0.0f ? -dot : dot) + 1e-9f;
        ox = ox / den;
        oy = oy / den;
        oz = oz / den;
        float wl2 = ((ox * ox) + (oy * oy)) + (oz * oz);
        if (wl2 < 1e-18f) break;  // the same on every lane
        float dw = 1.0f, dx = ox * 0.5f, dy = oy * 0.5f, dz = oz * 0.5f;
        float nw = (((dw * qw) - (dx * qx)) - (dy * qy)) - (dz * qz);
        float nx = (((dw * qx) + (dx * qw)) + (dy * qz)) - (dz * qy);
        float ny = (((dw * qy) - (dx * qz)) + (dy * qw)) + (dz * qx);
        float nz = (((dw * qz) + (dx * qy)) - (dy * qx)) + (dz * qw);
        float nl = sqrtf((((nw * nw) + (nx * nx)) + (ny * ny)) + (nz * nz));
        qw = nw / nl;
        qx = nx / nl;
        qy = ny / nl;
        qz = nz / nl;
    }
    if (lane == 0) b.body_q[bd] = make_float4(qw, qx, qy, qz);
    float m[9];
    quat_to_mat(qw, qx, qy, qz, m);
    for (uint32_t q = a0 + lane; q < a1; q += 64u) {
        const float4 r = b.body_rest[q];
        b.body_goal[q] = make_float4((((m[0] * r.x) + (m[1] * r.y)) + (m[2] * r.z)) + cx,
                                     (((m[3] * r.x) + (m[4] * r.y)) + (m[5] * r.z)) + cy,
                                     (((m[6] * r.x) + (m[7] * r.y)) + (m[8] * r.z)) + cz, 0.0f);
    }
}

// ---------------------------------------------------------------------------
// a11: v = (x - x*)/dt (subtract_functor, integration_kernel.cuh:465-475)
// plus XSPH over fluid neighbours from displacements u = x - x*:
//   v_i = u_i/dt + ((c/rho0_i) sum_j (u_j - u_i) m_j W_ij)/dt
// with the pair term (u_j - u_i) m_j t^3, t = h^2 - r^2, and poly6
// multiplying the finished sum.  Four lanes per particle: lane hh sums
// quarter hh of the table (its chunks hh, hh + 4, ..., in table order), and
// two shuffles add the quarters as (q0 + q2) + (q1 + q3).  Its one sum is
// short (about 17 VALU per pair), so the kernel is bound by latency more
// than by issue: four lanes give the 256 particles of a staged unit (24
// bytes of LDS per slot, two units per CU) eight wavefronts per SIMD.

static constexpr uint32_t kFinalizeStagedBlock = 4u * kStageUnit;

template <bool UNIFORM, bool STAGED, bool MASS = false>
__global__ __launch_bounds__(STAGED ? kFinalizeStagedBlock : kMaxBlock) void finalize_k(
    DevBuffers b, KConst k, uint32_t n, float mu, float dt, uint32_t its)
{
    // staged: neighbour position and displacement u = p - x* (x, y, z, u.x)
    // and (u.y, u.z), formed with the same subtractions as the global path
    // (and the sentinel record).  In a scene that is not a uniform fluid a
    // particle that is not fluid is staged at infinity: the XSPH sum leaves
    // it out, as the global path does by its x*.w < 0.  With every fluid
    // particle of one mass (mu >= 0) the mass is mu; with fluid masses that
    // differ (MASS) each slot's mass (x*.w) is staged as well, 12 KB more
    // LDS (one block per CU instead of two).  A unit that was not staged
    // gathers from global memory through the unit's chunk table (s_src).
    __shared__ float4 lds_pu[STAGED ? kStageSlots + 1 : 1];
    __shared__ float2 lds_u[STAGED ? kStageSlots + 1 : 1];
    __shared__ float lds_m[STAGED && MASS ? kStageSlots + 1 : 1];
    __shared__ uint32_t s_src[STAGED && !UNIFORM ? kStageChunks : 1];
    const uint32_t unit = xcd_block();
    const uint32_t hh = threadIdx.x & 3u;
    const uint32_t i = STAGED ? unit * kStageUnit + (threadIdx.x >> 2) : (unit * blockDim.x + threadIdx.x) >> 2;
    bool staged = false, mapped = false;
    if (STAGED) {
        const StageDesc *d = b.stage + unit;
        staged = d->staged != 0u;
        mapped = !UNIFORM && !MASS && staged && mu < 0.0f;
        if (mapped) {
            for (uint32_t q = threadIdx.x; q < d->nchunks; q += blockDim.x) s_src[q] = d->src[q];
            __syncthreads();
        }
        staged = staged && !mapped;
        const float inf = __builtin_inff();
        if (staged && threadIdx.x == 0) {
            lds_pu[kStageSentinel] = make_float4(inf, inf, inf, 0.0f);
            lds_u[kStageSentinel] = make_float2(0.0f, 0.0f);
            if (MASS) lds_m[kStageSentinel] = 0.0f;
        }
        if (MASS && staged)
            stage_copy<3, kFinalizeStagedBlock>(d, n, [&](uint32_t j) { return PX{b.P[j], b.X[j]}; },
                                                [&](uint32_t l, const PX &v) {
                                                    lds_pu[l] = v.x.w < 0.0f
                                                                    ? make_float4(inf, inf, inf, 0.0f)
                                                                    : make_float4(v.p.x, v.p.y, v.p.z, v.p.x - v.x.x);
                                                    lds_u[l] = make_float2(v.p.y - v.x.y, v.p.z - v.x.z);
                                                    lds_m[l] = v.x.w;
                                                });
        else if (staged)
            stage_copy<3, kFinalizeStagedBlock>(d, n, [&](uint32_t j) { return PX{b.P[j], b.X[j]}; },
                                                [&](uint32_t l, const PX &v) {
                                                    lds_pu[l] = !UNIFORM && v.x.w < 0.0f
                                                                    ? make_float4(inf, inf, inf, 0.0f)
                                                                    : make_float4(v.p.x, v.p.y, v.p.z, v.p.x - v.x.x);
                                                    lds_u[l] = make_float2(v.p.y - v.x.y, v.p.z - v.x.z);
                                                });
    }
    // every solve of this substep has completed
    if (i == 0 && hh == 0) b.status->iter += its;
    if (i >= n) return;  // the four lanes of a particle leave together
    const float4 p = b.P[i];
    const float4 x = b.X[i];
    const float ux = p.x - x.x, uy = p.y - x.y, uz = p.z - x.z;
    float vx = ux / dt, vy = uy / dt, vz = uz / dt;
    const Attr a = b.A[i];
    if ((UNIFORM || a.phase == PS_PHASE_FLUID) && k.xsph_c != 0.0f) {
        const uint32_t c = b.ncnt[i];
        // the XSPH sum of this lane's quarter of the table
        struct Sums {
            float sx, sy, sz;
        };
        Sums s = {0.0f, 0.0f, 0.0f};
        // qj: neighbour position, e: its displacement over the substep
        auto body = [&](const float4 &qj, float ex, float ey, float ez, float mj) {
            float dx = p.x - qj.x, dy = p.y - qj.y, dz = p.z - qj.z;
            float r2 = dist2(dx, dy, dz);
            if (!(r2 < k.h2)) return;
            float t = k.h2 - r2;
            float m = ((t * t) * t) * mj;
            s.sx = __builtin_fmaf(ex - ux, m, s.sx);
            s.sy = __builtin_fmaf(ey - uy, m, s.sy);
            s.sz = __builtin_fmaf(ez - uz, m, s.sz);
        };
        struct PU {
            float4 pu;
            float2 u;
        };
        auto glob = [&](const PX &d) { body(d.p, d.p.x - d.x.x, d.p.y - d.x.y, d.p.z - d.x.z, mu); };
        // the neighbour's mass from x*.w (the sort's 1/w for a fluid
        // particle, -1 for any other: left out)
        auto gen = [&](const PX &d) {
            if (d.x.w < 0.0f) return;
            body(d.p, d.p.x - d.x.x, d.p.y - d.x.y, d.p.z - d.x.z, d.x.w);
        };
        struct PUM {
            float4 pu;
            float2 u;
            float m;
        };
        if (STAGED && MASS && staged)
            for_neighbors_strided_chunked(
                b, i, c, hh, 4u, [&](uint32_t l) { return PUM{lds_pu[l], lds_u[l], lds_m[l]}; },
                [&](uint32_t, const PUM &d) { body(d.pu, d.pu.w, d.u.x, d.u.y, d.m); });
        else if (STAGED && staged)
            for_neighbors_strided_chunked(
                b, i, c, hh, 4u, [&](uint32_t l) { return PU{lds_pu[l], lds_u[l]}; },
                [&](uint32_t, const PU &d) { body(d.pu, d.pu.w, d.u.x, d.u.y, mu); });
        else if (STAGED && mapped)
            // table entries are LDS indices (the sentinel padding, never
            // visited, maps to i)
            for_neighbors_strided_chunked(
                b, i, c, hh, 4u,
                [&](uint32_t e) {
                    const uint32_t j = e >= kStageSlots ? i : s_src[e >> 4] + (e & (kStageChunk - 1u));
                    return PX{b.P[j], b.X[j]};
                },
                [&](uint32_t, const PX &d) { gen(d); });
        else
            for_neighbors_strided_chunked(
                b, i, c, hh, 4u, [&](uint32_t j) { return PX{b.P[j], b.X[j]}; },
                [&](uint32_t, const PX &d) {
                    if (UNIFORM)
                        glob(d);
                    else
                        gen(d);
                });
        // lanes 0 and 1 form q0 + q2 and q1 + q3, lane 0 their sum
        s.sx = s.sx + __shfl_xor(s.sx, 2, 64);
        s.sy = s.sy + __shfl_xor(s.sy, 2, 64);
        s.sz = s.sz + __shfl_xor(s.sz, 2, 64);
        const Sums o = {__shfl_xor(s.sx, 1, 64), __shfl_xor(s.sy, 1, 64), __shfl_xor(s.sz, 1, 64)};
        const float sx = k.poly6 * (s.sx + o.sx), sy = k.poly6 * (s.sy + o.sy), sz = k.poly6 * (s.sz + o.sz);
        const float cr = k.xsph_c / a.rho0;
        vx = vx + ((cr * sx) / dt);
        vy = vy + ((cr * sy) / dt);
        vz = vz + ((cr * sz) / dt);
    }
    if (hh) return;  // the particle's lane 0 writes
    b.V[i] = make_float4(vx, vy, vz, 0.0f);
}

// Sleeping (Particle::confirmGuess, cpu/src/particle.h:60-65): a particle
// that moved less than 1e-4 keeps its start position and gets v = 0.
__global__ __launch_bounds__(kBlock) void sleep_k(DevBuffers b, uint32_t n)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = b.P[i];
    const float4 x = b.X[i];
    float ex = p.x - x.x, ey = p.y - x.y, ez = p.z - x.z;
    float l = sqrtf(((ex * ex) + (ey * ey)) + (ez * ez));
    if (l < 1e-4f) {
        b.V[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.P[i] = make_float4(x.x, x.y, x.z, 0.0f);
    }
}

// ---------------------------------------------------------------------------
// launchers

thread_local LaunchEvents g_launch_events;

static inline uint32_t blocks_for(uint32_t n, uint32_t bs = kBlock) { return (n + bs - 1) / bs; }

uint32_t scan_tiles(uint32_t num_cells) { return (num_cells + kScanTile - 1) / kScanTile; }

hipError_t launch_predict_hash(hipStream_t st, const DevBuffers &b, const KConst &k, uint32_t n,
                               float dt, bool predict, bool hash, uint32_t bs)
{
    if (n == 0) return hipSuccess;
    const dim3 g(blocks_for(n, bs)), t(bs);
    if (predict && hash)
        PS_LAUNCH((predict_hash_k<true, true>), g, t, 0, st, b, k, n, dt);
    else if (predict)
        PS_LAUNCH((predict_hash_k<true, false>), g, t, 0, st, b, k, n, dt);
    else if (hash)
        PS_LAUNCH((predict_hash_k<false, true>), g, t, 0, st, b, k, n, dt);
    return hipGetLastError();
}

hipError_t launch_scan(hipStream_t st, const DevBuffers &b, uint32_t num_cells)
{
    const uint32_t tiles = scan_tiles(num_cells);
    hipLaunchKernelGGL(scan_partials_k, dim3(tiles), dim3(kBlock), 0, st, b.cell_count,
                       b.scan_part, num_cells);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(scan_final_k, dim3(tiles), dim3(kBlock), 0, st, b.cell_count, b.scan_part,
                       b.cell_off, num_cells);
    return hipGetLastError();
}

hipError_t launch_scatter(hipStream_t st, const DevBuffers &b, uint32_t n)
{
    if (n == 0) return hipSuccess;
    PS_LAUNCH(scatter_k, dim3(blocks_for(n)), dim3(kBlock), 0, st, b, n);
    return hipGetLastError();
}

hipError_t launch_reorder(hipStream_t st, const DevBuffers &b, uint32_t n, uint32_t bs, float mj_solid)
{
    if (n == 0) return hipSuccess;
    PS_LAUNCH(reorder_k, dim3(blocks_for(n, bs)), dim3(bs), 0, st, b, n, mj_solid);
    return hipGetLastError();
}

hipError_t launch_neighbors(hipStream_t st, const DevBuffers &b, const KConst &k, uint32_t n,
                            uint32_t kmax, uint32_t bs)
{
    if (n == 0) return hipSuccess;
    if (b.staging && b.nbr_glob)  // recording: the global-slot copy as well
        PS_LAUNCH(neighbors_staged_k<true>, dim3(blocks_for(n, kStageUnit)), dim3(kStageUnit), 0,
                           st, b, k, n, kmax);
    else if (b.staging)
        PS_LAUNCH(neighbors_staged_k<false>, dim3(blocks_for(n, kStageUnit)), dim3(kStageUnit), 0,
                           st, b, k, n, kmax);
    else
        PS_LAUNCH(neighbors_k, dim3(blocks_for(n, bs)), dim3(bs), 0, st, b, k, n, kmax);
    return hipGetLastError();
}

hipError_t launch_lambda(hipStream_t st, const DevBuffers &b, const KConst &k, uint32_t n,
                         bool uniform, float wu, uint32_t bs)
{
    if (n == 0) return hipSuccess;
    const dim3 g(blocks_for(n, bs)), t(bs);
    const float mu = 1.0f / wu;
    const dim3 gs(blocks_for(n, kStageUnit)), ts(2 * kStageUnit);  // staged: two lanes per particle
    if (b.tape_lq && b.staging && uniform)  // recording for the reverse pass
        PS_LAUNCH((lambda_k<true, true, true>), gs, ts, 0, st, b, k, n, mu);
    else if (b.tape_lq && b.staging)
        PS_LAUNCH((lambda_k<false, true, true>), gs, ts, 0, st, b, k, n, mu);
    else if (b.tape_lq && uniform)
        PS_LAUNCH((lambda_k<true, false, true>), g, t, 0, st, b, k, n, mu);
    else if (b.tape_lq)
        PS_LAUNCH((lambda_k<false, false, true>), g, t, 0, st, b, k, n, mu);
    else if (uniform && b.staging)
        PS_LAUNCH((lambda_k<true, true>), dim3(blocks_for(n, kStageUnit)), dim3(2 * kStageUnit),
                           0, st, b, k, n, mu);
    else if (b.staging)
        PS_LAUNCH((lambda_k<false, true>), dim3(blocks_for(n, kStageUnit)), dim3(2 * kStageUnit),
                           0, st, b, k, n, mu);
    else if (uniform)
        PS_LAUNCH((lambda_k<true, false>), g, t, 0, st, b, k, n, mu);
    else
        PS_LAUNCH((lambda_k<false, false>), g, t, 0, st, b, k, n, mu);
    return hipGetLastError();
}

hipError_t launch_shape(hipStream_t st, const DevBuffers &b, uint32_t nbody)
{
    if (nbody == 0) return hipSuccess;
    // one wavefront per body
    PS_LAUNCH(shape_k, dim3(blocks_for(nbody, kBlock / 64)), dim3(kBlock), 0, st, b, nbody);
    return hipGetLastError();
}

// NC: tensile exponent known at compile time (4, as in the reference's
// E_P, integration_kernel.cuh:34) or 0 = read from the parameters
template <uint32_t F>
static void solve_variant(hipStream_t st, const DevBuffers &b, const KConst &k, uint32_t n,
                          bool uniform, float wu, uint32_t iter, uint32_t bs)
{
    const dim3 g(blocks_for(n, bs)), t(bs);
    const dim3 gs(blocks_for(n, kStageUnit)), ts(2 * kStageUnit);  // two lanes per particle
    const bool n4 = k.n_corr == 4u;
    // the contact lists of a block
    const size_t cl = (F & F_CONTACT) ? sizeof(uint32_t) * kContactList * bs : 0;
    const size_t cls = (F & F_CONTACT) ? sizeof(uint32_t) * kHalfContactList * 2 * kStageUnit : 0;  // per lane
    // staged variants for every composition (uniform fluid or not)
    constexpr bool CAN_STAGE = true;
    if (b.tape_f && b.staging) {  // recording for the reverse pass
        if (uniform && n4)
            PS_LAUNCH((solve_k<F, true, 4, true, true>), gs, ts, cls, st, b, k, n, wu, iter);
        else if (uniform)
            PS_LAUNCH((solve_k<F, true, 0, true, true>), gs, ts, cls, st, b, k, n, wu, iter);
        else if (n4)
            PS_LAUNCH((solve_k<F, false, 4, true, true>), gs, ts, cls, st, b, k, n, wu, iter);
        else
            PS_LAUNCH((solve_k<F, false, 0, true, true>), gs, ts, cls, st, b, k, n, wu, iter);
        return;
    }
    if (b.tape_f) {  // recording for the reverse pass
        if (uniform && n4)
            PS_LAUNCH((solve_k<F, true, 4, false, true>), g, t, cl, st, b, k, n, wu, iter);
        else if (uniform)
            PS_LAUNCH((solve_k<F, true, 0, false, true>), g, t, cl, st, b, k, n, wu, iter);
        else if (n4)
            PS_LAUNCH((solve_k<F, false, 4, false, true>), g, t, cl, st, b, k, n, wu, iter);
        else
            PS_LAUNCH((solve_k<F, false, 0, false, true>), g, t, cl, st, b, k, n, wu, iter);
        return;
    }
    if (CAN_STAGE && uniform && b.staging && n4)
        PS_LAUNCH((solve_k<F, true, 4, CAN_STAGE>), gs, ts, 0, st, b, k, n, wu, iter);
    else if (CAN_STAGE && uniform && b.staging)
        PS_LAUNCH((solve_k<F, true, 0, CAN_STAGE>), gs, ts, 0, st, b, k, n, wu, iter);
    else if (CAN_STAGE && b.staging && n4)
        PS_LAUNCH((solve_k<F, false, 4, CAN_STAGE>), gs, ts, cls, st, b, k, n, wu, iter);
    else if (CAN_STAGE && b.staging)
        PS_LAUNCH((solve_k<F, false, 0, CAN_STAGE>), gs, ts, cls, st, b, k, n, wu, iter);
    else if (uniform && n4)
        PS_LAUNCH((solve_k<F, true, 4, false>), g, t, cl, st, b, k, n, wu, iter);
    else if (uniform)
        PS_LAUNCH((solve_k<F, true, 0, false>), g, t, cl, st, b, k, n, wu, iter);
    else if (n4)
        PS_LAUNCH((solve_k<F, false, 4, false>), g, t, cl, st, b, k, n, wu, iter);
    else
        PS_LAUNCH((solve_k<F, false, 0, false>), g, t, cl, st, b, k, n, wu, iter);
}

hipError_t launch_solve(hipStream_t st, const DevBuffers &b, const KConst &k, uint32_t n,
                        uint32_t features, bool uniform, float wu, uint32_t iter, uint32_t bs)
{
    if (n == 0) return hipSuccess;
    // uniform-fluid scenes have no contacts, constraints or bodies
    switch (features) {
    case F_FLUID: solve_variant<F_FLUID>(st, b, k, n, uniform, wu, iter, bs); break;
    case F_FLUID | F_PIN: solve_variant<F_FLUID | F_PIN>(st, b, k, n, false, wu, iter, bs); break;
    default:
        if (b.staging && !b.tape_f && b.split_solve) {
            // Scenes where contact-walking particles are rare (ropes in a
            // fluid): units without one run the variant without contacts
            // (their particles take no contact branch, so the results are
            // the same), which has no contact lists or staged phases in LDS,
            // 49 KB against 78 KB, three blocks per CU instead of two.  The
            // units with one run the full variant.  Each launch covers every
            // unit and the other class's blocks leave at once, so where most
            // units hold such a particle (cloth, rigid bodies) the split
            // does not pay and is not made.  Timed as one: the event pair
            // brackets both launches.
            const LaunchEvents ev = g_launch_events;
            g_launch_events = LaunchEvents{};
            DevBuffers b1 = b, b2 = b;
            b1.solve_units = 1u;
            b2.solve_units = 2u;
            g_launch_events = LaunchEvents{ev.start, nullptr};
            solve_variant<F_FLUID | F_DIST | F_SHAPE | F_PIN>(st, b1, k, n, false, wu, iter, bs);
            g_launch_events = LaunchEvents{nullptr, ev.stop};
            solve_variant<F_FLUID | F_CONTACT | F_DIST | F_SHAPE | F_PIN>(st, b2, k, n, false, wu, iter, bs);
            g_launch_events = LaunchEvents{};
        } else {
            solve_variant<F_FLUID | F_CONTACT | F_DIST | F_SHAPE | F_PIN>(st, b, k, n, false, wu, iter, bs);
        }
        break;
    }
    return hipGetLastError();
}

hipError_t launch_finalize(hipStream_t st, const DevBuffers &b, const KConst &k, uint32_t n,
                           float dt, bool sleeping, bool uniform, float wu, float wf, uint32_t its,
                           uint32_t bs)
{
    if (n == 0) return hipSuccess;
    // four lanes per particle (bs is a multiple of 64)
    const dim3 g(blocks_for(n, bs / 4u)), t(bs);
    const float mu = 1.0f / wu;
    if (uniform && b.staging)
        PS_LAUNCH((finalize_k<true, true>), dim3(blocks_for(n, kStageUnit)),
                           dim3(kFinalizeStagedBlock), 0, st, b, k, n, mu, dt, its);
    else if (b.staging && wf >= 0.0f)  // every fluid particle of inverse mass wf
        PS_LAUNCH((finalize_k<false, true>), dim3(blocks_for(n, kStageUnit)),
                           dim3(kFinalizeStagedBlock), 0, st, b, k, n, 1.0f / wf, dt, its);
    else if (b.staging)  // fluid masses that differ: staged with each slot's mass
        PS_LAUNCH((finalize_k<false, true, true>), dim3(blocks_for(n, kStageUnit)),
                           dim3(kFinalizeStagedBlock), 0, st, b, k, n, -1.0f, dt, its);
    else if (uniform)
        PS_LAUNCH((finalize_k<true, false>), g, t, 0, st, b, k, n, mu, dt, its);
    else
        PS_LAUNCH((finalize_k<false, false>), g, t, 0, st, b, k, n, mu, dt, its);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !sleeping) return e;
    hipLaunchKernelGGL(sleep_k, dim3(blocks_for(n)), dim3(kBlock), 0, st, b, n);
    return hipGetLastError();
}

}  // namespace ps
