// ps_trajectory.hip — binary trajectory files (host code over the public
// C-ABI: ps_export_device fills a device staging buffer in insertion order,
// one copy brings it to pinned host memory, the xyz/phase records go to the
// file).  Format: the comment above ps_trajectory_open in ps.h.

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "ps.h"

struct ps_trajectory {
    FILE *f = nullptr;
    uint32_t fields = 0, dims = 0;
    size_t cap = 0;               // records the staging buffers hold
    void *dev = nullptr;          // device staging, 16 bytes per record
    void *host = nullptr;         // pinned host copy of it
    std::vector<float> xyz;       // 3 floats per record, written to the file
};

namespace {

void release(ps_trajectory *t)
{
    if (t->dev) (void)hipFree(t->dev);
    if (t->host) (void)hipHostFree(t->host);
    t->dev = t->host = nullptr;
    t->cap = 0;
}

int grow(ps_trajectory *t, size_t n)
{
    if (n <= t->cap) return PS_OK;
    release(t);
    size_t cap = n < 1024 ? 1024 : n + n / 4;
    if (hipMalloc(&t->dev, 16 * cap) != hipSuccess) return PS_ERR_OUT_OF_MEMORY;
    if (hipHostMalloc(&t->host, 16 * cap, hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(t->dev);
        t->dev = nullptr;
        return PS_ERR_OUT_OF_MEMORY;
    }
    t->cap = cap;
    return PS_OK;
}

bool put(ps_trajectory *t, const void *p, size_t bytes)
{
    return fwrite(p, 1, bytes, t->f) == bytes;
}

}  // namespace

extern "C" {

int ps_trajectory_open(const char *path, uint32_t fields, uint32_t dims, ps_trajectory **out)
{
    if (!path || !out || !fields || (fields & ~7u) || (dims != 2 && dims != 3))
        return PS_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    ps_trajectory *t = new ps_trajectory();
    t->f = fopen(path, "wb");
    if (!t->f) {
        delete t;
        return PS_ERR_IO;
    }
    t->fields = fields;
    t->dims = dims;
    unsigned char hdr[32] = {0};
    memcpy(hdr, "PSTRAJ01", 8);
    const uint32_t version = 1;
    memcpy(hdr + 8, &version, 4);
    memcpy(hdr + 12, &fields, 4);
    memcpy(hdr + 16, &dims, 4);
    if (!put(t, hdr, sizeof(hdr))) {
        fclose(t->f);
        delete t;
        return PS_ERR_IO;
    }
    *out = t;
    return PS_OK;
}

int ps_trajectory_write(ps_trajectory *t, ps_system *sys)
{
    if (!t || !t->f || !sys) return PS_ERR_INVALID_ARGUMENT;
    const uint32_t n = ps_num_particles(sys);
    ps_status st;
    if (int rc = ps_get_status(sys, &st)) return rc;
    if (int rc = grow(t, n)) return rc;
    unsigned char hdr[16];
    const uint64_t sub = st.substeps;
    memcpy(hdr, &sub, 8);
    memcpy(hdr + 8, &n, 4);
    memcpy(hdr + 12, &t->fields, 4);
    if (!put(t, hdr, sizeof(hdr))) return PS_ERR_IO;
    t->xyz.resize(3 * (size_t)n);
    for (uint32_t field = 1; field <= 4; field <<= 1) {
        if (!(t->fields & field)) continue;
        if (n) {
            if (int rc = ps_export_device(sys, field, 0, n, t->dev)) return rc;
            if (int rc = ps_synchronize(sys)) return rc;
            const size_t bytes = field == PS_FIELD_PHASES ? 4 * (size_t)n : 16 * (size_t)n;
            if (hipMemcpy(t->host, t->dev, bytes, hipMemcpyDeviceToHost) != hipSuccess)
                return PS_ERR_DEVICE;
        }
        if (field == PS_FIELD_PHASES) {
            if (!put(t, t->host, 4 * (size_t)n)) return PS_ERR_IO;
            continue;
        }
        const float *r = (const float *)t->host;
        for (uint32_t k = 0; k < n; k++) {
            t->xyz[3 * k] = r[4 * k];
            t->xyz[3 * k + 1] = r[4 * k + 1];
            t->xyz[3 * k + 2] = r[4 * k + 2];
        }
        if (!put(t, t->xyz.data(), 12 * (size_t)n)) return PS_ERR_IO;
    }
    return fflush(t->f) == 0 ? PS_OK : PS_ERR_IO;
}

int ps_trajectory_close(ps_trajectory *t)
{
    if (!t) return PS_OK;
    int rc = PS_OK;
    if (t->f && fclose(t->f) != 0) rc = PS_ERR_IO;
    release(t);
    delete t;
    return rc;
}

}  // extern "C"
