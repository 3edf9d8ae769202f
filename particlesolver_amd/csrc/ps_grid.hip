// ps_grid.hip — the stateless grid sort over caller device arrays
// (ps_grid_sort): cell hash, a stable sort by (cell, input index) and the
// cell bounds, the reference's calcHash + sortParticles +
// reorderDataAndFindCellStart (gpu/src/cuda/integration.cu:161-275) in one
// call, with the cells and order of ps_step's grid build.
//
// A counting sort in five launches, all gathers or one-slot scatters:
//   1. clear the cell counts (a fill kernel, so the call can sit in a
//      captured HIP graph);
//   2. grid_hash_k: cell key of every input and its arrival rank in the
//      cell (the return value of one atomicAdd on the cell's count);
//   3. the exclusive scan of the counts into cell offsets (the scan kernels
//      of ps_step's build, ps_kernels.hip);
//   4. grid_place_k: provisional slot = offset + arrival rank;
//   5. grid_order_k: final slot = offset + the number of inputs in the same
//      cell with a smaller index (a stable sort), the outputs gathered there;
//   then grid_bounds_k turns the offsets into cell_start / cell_end.
// Arrival order depends on scheduling; the final order does not.
//
// ps_grid_sort_noalloc runs the same launches over `scratch`, a static
// device array of the library (one copy per GPU), so that it allocates
// nothing; its address on each GPU is looked up once with
// hipGetSymbolAddress and kept.

#include <cmath>
#include <mutex>

#include "ps_device.hpp"

namespace ps {
namespace {

constexpr uint32_t kGridBlock = 256;

struct GridK {
    float origin[3];
    float cell;
    uint32_t grid[3];
};

// workspace: three per-input arrays, the counts, the offsets and the scan
// partials, each 256-byte aligned
constexpr size_t align256(size_t x) { return (x + 255u) & ~(size_t)255u; }
constexpr size_t grid_tiles(size_t cells) { return (cells + kScanTileCells - 1u) / kScanTileCells; }
constexpr size_t grid_work_bytes(size_t n, size_t cells)
{
    return 3u * align256(4u * n) + align256(4u * cells) + align256(4u * (cells + 1u)) +
           align256(4u * grid_tiles(cells));
}

struct GridWork {
    uint32_t *key, *rank, *tmp;  // n each
    uint32_t *cnt;               // cells
    uint32_t *off;               // cells + 1
    uint32_t *part;              // scan tiles
};

GridWork carve(void *base, uint32_t n, uint32_t cells)
{
    char *p = static_cast<char *>(base);
    GridWork w;
    w.key = reinterpret_cast<uint32_t *>(p);
    p += align256(4u * (size_t)n);
    w.rank = reinterpret_cast<uint32_t *>(p);
    p += align256(4u * (size_t)n);
    w.tmp = reinterpret_cast<uint32_t *>(p);
    p += align256(4u * (size_t)n);
    w.cnt = reinterpret_cast<uint32_t *>(p);
    p += align256(4u * (size_t)cells);
    w.off = reinterpret_cast<uint32_t *>(p);
    p += align256(4u * ((size_t)cells + 1u));
    w.part = reinterpret_cast<uint32_t *>(p);
    return w;
}

__global__ __launch_bounds__(kGridBlock) void grid_fill_k(uint32_t *dst, uint32_t count, uint32_t v)
{
    const uint32_t i = blockIdx.x * kGridBlock + threadIdx.x;
    if (i < count) dst[i] = v;
}

__global__ __launch_bounds__(kGridBlock) void grid_hash_k(const float4 *pos, uint32_t n, GridK g,
                                                          uint32_t *key, uint32_t *rank, uint32_t *cnt)
{
    const uint32_t i = blockIdx.x * kGridBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    uint32_t flags = 0;  // clamping is the defined behaviour here; nothing reads the bits
    const uint32_t cx = cell_axis(p.x, g.origin[0], g.cell, g.grid[0], flags);
    const uint32_t cy = cell_axis(p.y, g.origin[1], g.cell, g.grid[1], flags);
    const uint32_t cz = cell_axis(p.z, g.origin[2], g.cell, g.grid[2], flags);
    const uint32_t c = (cz * g.grid[1] + cy) * g.grid[0] + cx;
    key[i] = c;
    rank[i] = atomicAdd(&cnt[c], 1u);
}

// provisional slot: the cell's offset plus the arrival rank, < off[c + 1]
__global__ __launch_bounds__(kGridBlock) void grid_place_k(const uint32_t *key, const uint32_t *rank,
                                                           const uint32_t *off, uint32_t *tmp, uint32_t n)
{
    const uint32_t i = blockIdx.x * kGridBlock + threadIdx.x;
    if (i >= n) return;
    tmp[off[key[i]] + rank[i]] = i;
}

// final slot: the cell's offset plus the members with a smaller input index
__global__ __launch_bounds__(kGridBlock) void grid_order_k(const float4 *pos, const uint32_t *key,
                                                           const uint32_t *off, const uint32_t *tmp,
                                                           uint32_t n, uint32_t *cell_key,
                                                           uint32_t *sorted_index, float4 *sorted_pos)
{
    const uint32_t i = blockIdx.x * kGridBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = key[i];
    const uint32_t b0 = off[c], b1 = off[c + 1];
    uint32_t r = 0;
    for (uint32_t u = b0; u < b1; u += 4) {
        uint32_t t[4];
#pragma unroll
        for (uint32_t q = 0; q < 4; q++) t[q] = tmp[min(u + q, b1 - 1u)];
#pragma unroll
        for (uint32_t q = 0; q < 4; q++) r += (u + q < b1 && t[q] < i) ? 1u : 0u;
    }
    const uint32_t d = b0 + r;
    cell_key[d] = c;
    sorted_index[d] = i;
    if (sorted_pos) sorted_pos[d] = pos[i];
}

// cell c holds sorted slots [off[c], off[c + 1]); empty: both 0xffffffff
__global__ __launch_bounds__(kGridBlock) void grid_bounds_k(const uint32_t *off, uint32_t cells,
                                                            uint32_t *start, uint32_t *end)
{
    const uint32_t c = blockIdx.x * kGridBlock + threadIdx.x;
    if (c >= cells) return;
    const uint32_t s = off[c], e = off[c + 1];
    start[c] = s < e ? s : 0xffffffffu;
    end[c] = s < e ? e : 0xffffffffu;
}

uint32_t grid_blocks(uint32_t count) { return (count + kGridBlock - 1) / kGridBlock; }

// Checks the arguments; on success sets g and the cell count.
int grid_check(const ps_grid_desc *grid, uint32_t n, const float *pos, const uint32_t *cell_key,
               const uint32_t *sorted_index, const uint32_t *cell_start, const uint32_t *cell_end,
               GridK &g, uint32_t &cells)
{
    if (!grid || !cell_start || !cell_end || (n > 0 && (!pos || !cell_key || !sorted_index)))
        return fail_stateless(PS_ERR_INVALID_ARGUMENT, "ps_grid_sort: null pointer");
    if (!(grid->cell > 0.0f) || !std::isfinite(grid->cell) || !std::isfinite(grid->origin[0]) ||
        !std::isfinite(grid->origin[1]) || !std::isfinite(grid->origin[2]))
        return fail_stateless(PS_ERR_INVALID_ARGUMENT, "ps_grid_sort: cell must be finite and > 0, origin finite");
    uint64_t c = 1;
    for (int a = 0; a < 3; a++) {
        if (grid->grid[a] == 0)
            return fail_stateless(PS_ERR_INVALID_ARGUMENT, "ps_grid_sort: every grid dimension must be > 0");
        c *= grid->grid[a];
        if (c >= (1ull << 31))
            return fail_stateless(PS_ERR_INVALID_ARGUMENT, "ps_grid_sort: 2^31 cells or more");
    }
    if (n >= (1u << 31)) return fail_stateless(PS_ERR_INVALID_ARGUMENT, "ps_grid_sort: 2^31 inputs or more");
    for (int a = 0; a < 3; a++) {
        g.origin[a] = grid->origin[a];
        g.grid[a] = grid->grid[a];
    }
    g.cell = grid->cell;
    cells = (uint32_t)c;
    return PS_OK;
}

// The launches over a carved workspace.
hipError_t grid_run(hipStream_t st, const GridK &g, uint32_t cells, uint32_t n, const float4 *pos,
                    const GridWork &w, uint32_t *cell_key, uint32_t *sorted_index, uint32_t *cell_start,
                    uint32_t *cell_end, float4 *sorted_pos)
{
    const dim3 t(kGridBlock);
    hipLaunchKernelGGL(grid_fill_k, dim3(grid_blocks(cells)), t, 0, st, w.cnt, cells, 0u);
    if (n > 0) hipLaunchKernelGGL(grid_hash_k, dim3(grid_blocks(n)), t, 0, st, pos, n, g, w.key, w.rank, w.cnt);
    // the scan kernels of the grid build read cell_count / scan_part and
    // write cell_off (and clear the counts again)
    DevBuffers view{};
    view.cell_count = w.cnt;
    view.cell_off = w.off;
    view.scan_part = w.part;
    hipError_t e = launch_scan(st, view, cells);
    if (e != hipSuccess) return e;
    if (n > 0) {
        hipLaunchKernelGGL(grid_place_k, dim3(grid_blocks(n)), t, 0, st, w.key, w.rank, w.off, w.tmp, n);
        hipLaunchKernelGGL(grid_order_k, dim3(grid_blocks(n)), t, 0, st, pos, w.key, w.off, w.tmp, n, cell_key,
                           sorted_index, sorted_pos);
    }
    hipLaunchKernelGGL(grid_bounds_k, dim3(grid_blocks(cells)), t, 0, st, w.off, cells, cell_start, cell_end);
    return hipGetLastError();
}

}  // namespace
}  // namespace ps

// The allocation-free variant's workspace, sized for
// PS_GRID_NOALLOC_MAX_PARTICLES inputs on PS_GRID_NOALLOC_MAX_CELLS cells
// (28 MiB).  In host code `scratch` names a host-side stand-in, not the
// device address: only the pointer from hipGetSymbolAddress is used.
static constexpr size_t kNoAllocBytes =
    ps::grid_work_bytes(PS_GRID_NOALLOC_MAX_PARTICLES, PS_GRID_NOALLOC_MAX_CELLS);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace ps {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on the current device, looked up on the first
// call there (with that device current) and kept.
int scratch_on_current_device(void **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return fail_stateless(PS_ERR_NO_DEVICE, std::string("ps_grid_sort_noalloc: hipGetDevice: ") +
                                                    hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return fail_stateless(PS_ERR_DEVICE, "ps_grid_sort_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail_stateless(PS_ERR_DEVICE, std::string("ps_grid_sort_noalloc: hipGetSymbolAddress: ") +
                                                     hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PS_OK;
}

}  // namespace
}  // namespace ps

using namespace ps;

extern "C" int ps_grid_sort(const ps_grid_desc *grid, uint32_t n, const float *pos, uint32_t *cell_key,
                            uint32_t *sorted_index, uint32_t *cell_start, uint32_t *cell_end,
                            float *sorted_pos, void *stream)
{
    GridK g;
    uint32_t cells = 0;
    if (int rc = grid_check(grid, n, pos, cell_key, sorted_index, cell_start, cell_end, g, cells)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t bytes = grid_work_bytes(n, cells);
    void *ws = nullptr;
    hipError_t e = hipMallocAsync(&ws, bytes, st);
    if (e != hipSuccess)
        return fail_stateless(e == hipErrorOutOfMemory ? PS_ERR_OUT_OF_MEMORY : PS_ERR_DEVICE,
                              std::string("ps_grid_sort: hipMallocAsync: ") + hipGetErrorString(e));
    e = grid_run(st, g, cells, n, reinterpret_cast<const float4 *>(pos), carve(ws, n, cells), cell_key,
                 sorted_index, cell_start, cell_end, reinterpret_cast<float4 *>(sorted_pos));
    const hipError_t ef = hipFreeAsync(ws, st);
    if (e == hipSuccess) e = ef;
    if (e != hipSuccess)
        return fail_stateless(PS_ERR_DEVICE, std::string("ps_grid_sort: ") + hipGetErrorString(e));
    return PS_OK;
}

extern "C" int ps_grid_sort_noalloc(const ps_grid_desc *grid, uint32_t n, const float *pos, uint32_t *cell_key,
                                    uint32_t *sorted_index, uint32_t *cell_start, uint32_t *cell_end,
                                    float *sorted_pos, void *stream)
{
    GridK g;
    uint32_t cells = 0;
    if (int rc = grid_check(grid, n, pos, cell_key, sorted_index, cell_start, cell_end, g, cells)) return rc;
    if (n > PS_GRID_NOALLOC_MAX_PARTICLES)
        return fail_stateless(PS_ERR_CAPACITY, "ps_grid_sort_noalloc: more than PS_GRID_NOALLOC_MAX_PARTICLES "
                                               "(1048576) inputs; use ps_grid_sort");
    if (cells > PS_GRID_NOALLOC_MAX_CELLS)
        return fail_stateless(PS_ERR_CAPACITY, "ps_grid_sort_noalloc: more than PS_GRID_NOALLOC_MAX_CELLS "
                                               "(2097152) cells; use ps_grid_sort");
    void *ws = nullptr;
    if (int rc = scratch_on_current_device(&ws)) return rc;
    // the layout for (n, cells) is a prefix-sized carve of the one for the
    // limits: grid_work_bytes grows with both
    const hipError_t e = grid_run(static_cast<hipStream_t>(stream), g, cells, n,
                                  reinterpret_cast<const float4 *>(pos), carve(ws, n, cells), cell_key,
                                  sorted_index, cell_start, cell_end, reinterpret_cast<float4 *>(sorted_pos));
    if (e != hipSuccess)
        return fail_stateless(PS_ERR_DEVICE, std::string("ps_grid_sort_noalloc: ") + hipGetErrorString(e));
    return PS_OK;
}
