// ps_export.hip — device-side export of the particle state in insertion
// order (the replacement of the reference's position VBO, which the solver
// writes through a CUDA-GL mapping: gpu/src/particlesystem.cpp:95-96,
// gpu/src/cuda/util.cu:39-62).
//
// Storage may be cell-sorted; A[t].orig is the insertion index of slot t,
// so each thread reads one slot (coalesced) and writes the record of its
// insertion index (a permutation: no two threads write the same record).

#include "ps_internal.hpp"

namespace ps {

namespace {

constexpr uint32_t kExportBlock = 256;

// dst[orig - first] = (src.xyz, w) for the slots whose insertion index lies
// in [first, first + count)
__global__ __launch_bounds__(kExportBlock) void export_vec_k(const float4 *__restrict__ src,
                                                             const Attr *__restrict__ A, uint32_t n,
                                                             uint32_t first, uint32_t count,
                                                             float4 *__restrict__ dst, float w)
{
    const uint32_t t = blockIdx.x * kExportBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t o = A[t].orig - first;
    if (o >= count) return;
    const float4 v = src[t];
    dst[o] = make_float4(v.x, v.y, v.z, w);
}

__global__ __launch_bounds__(kExportBlock) void export_phase_k(const Attr *__restrict__ A, uint32_t n,
                                                               uint32_t first, uint32_t count,
                                                               int32_t *__restrict__ dst)
{
    const uint32_t t = blockIdx.x * kExportBlock + threadIdx.x;
    if (t >= n) return;
    const Attr a = A[t];
    const uint32_t o = a.orig - first;
    if (o >= count) return;
    dst[o] = a.phase;
}

}  // namespace

hipError_t launch_export(hipStream_t st, const DevBuffers &b, uint32_t n, uint32_t field,
                         uint32_t first, uint32_t count, void *dst)
{
    if (!n || !count) return hipSuccess;
    const dim3 grid((n + kExportBlock - 1) / kExportBlock);
    switch (field) {
    case PS_FIELD_POSITIONS:
        PS_LAUNCH(export_vec_k, grid, dim3(kExportBlock), 0, st, b.X, b.A, n, first, count,
                           (float4 *)dst, 1.0f);
        break;
    case PS_FIELD_VELOCITIES:
        PS_LAUNCH(export_vec_k, grid, dim3(kExportBlock), 0, st, b.V, b.A, n, first, count,
                           (float4 *)dst, 0.0f);
        break;
    case PS_FIELD_PHASES:
        PS_LAUNCH(export_phase_k, grid, dim3(kExportBlock), 0, st, b.A, n, first, count,
                           (int32_t *)dst);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ps
