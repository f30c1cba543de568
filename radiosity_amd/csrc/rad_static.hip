// rad_static.hip - rad_process_runs_static: the run-form ProcessHemicube
// (the kernels of rad_process_runs) for callers that cannot allocate (code
// under HIP graph capture, tight serving loops): the caller passes the item
// buffer and the triplet arrays in device memory, and the per-work-item slot
// counts and offsets live in `scratch`, a fixed device array every GPU holds
// until the process ends.  The call allocates nothing and synchronises with
// nothing; it queues its kernels on the context's stream.
//
// Every call on one GPU shares that GPU's scratch, so calls that overlap on
// one GPU (two streams, two threads) race.  Nothing in it has to start at
// zero: the count kernel writes every slot count and the scan every offset.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "rad_ctx.h"
#include "rad_internal.h"

namespace {

// 4 MiB per GPU: slot counts (n words) and offsets (n + 1 words) of up to
// (kStaticWords - 1) / 2 work-items (k hemicubes x TEX_H rows x spans).
constexpr uint32_t kStaticWords = 1u << 20;
constexpr uint32_t kStaticItems = (kStaticWords - 1) / 2;
static __device__ __align__(256) uint32_t scratch[kStaticWords];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
// the device address of `scratch` on each GPU, from hipGetSymbolAddress with
// that GPU current (the host-side name `scratch` is no device address)
std::atomic<uint32_t *> g_scratch[kMaxDevices];

__global__ void k_copy_word(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst)
{
    *dst = *src;
}

} // namespace

namespace rad {

// Called by rad_create with the context's device current: looks the scratch
// address up once per device and reserves room for this entry's error
// messages, so that rad_process_runs_static allocates nothing, failures
// included.
int static_scratch_init(rad_ctx *c)
{
    c->err.reserve(256);
    if (c->device < 0 || c->device >= kMaxDevices)
        return RAD_OK; // rad_process_runs_static refuses such a device
    std::lock_guard<std::mutex> guard(g_lock);
    if (g_scratch[c->device].load())
        return RAD_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
        return RAD_E_HIP;
    g_scratch[c->device].store((uint32_t *)p);
    return RAD_OK;
}

} // namespace rad

extern "C" int rad_process_runs_static(rad_ctx *c, const uint32_t *d_itembuf, uint32_t k,
                                       uint32_t cap, uint32_t *d_hemicubes, uint32_t *d_ids,
                                       float *d_energies, uint32_t *d_count)
{
    if (!c)
        return RAD_E_ARG;
    if (!d_itembuf || !d_count || k == 0 || (cap && (!d_hemicubes || !d_ids || !d_energies))) {
        c->err = "process_runs_static: null device array or k = 0";
        return RAD_E_ARG;
    }
    if (c->device < 0 || c->device >= kMaxDevices) {
        c->err = "process_runs_static: device index beyond the supported 64";
        return RAD_E_LIMIT;
    }
    const uint32_t wx = c->cfg.ocl_workitems_x;
    const uint32_t nx = wx < c->g.TW ? wx : c->g.TW; // Config.cpp:43-45
    const uint64_t items = (uint64_t)k * c->g.TH * nx;
    if (items > kStaticItems) {
        c->err = "process_runs_static: k x TEX_H x workitems beyond the 524287 work-items its "
                 "fixed scratch holds";
        return RAD_E_LIMIT;
    }
    uint32_t *scr = g_scratch[c->device].load();
    if (!scr) {
        c->err = "process_runs_static: no scratch address for this device";
        return RAD_E_HIP;
    }
    const uint32_t n = (uint32_t)items;
    uint32_t *slots = scr, *offs = scr + n;
    rad::launch_runs_count(c->g, k, nx, d_itembuf, slots, offs, c->st);
    rad::launch_runs_write(c->g, k, nx, d_itembuf, c->dff, offs, cap, d_hemicubes, d_ids,
                           d_energies, c->st);
    k_copy_word<<<1, 1, 0, c->st>>>(offs + n, d_count);
    if (hipGetLastError() != hipSuccess) {
        c->err = "process_runs_static: kernel launch failed";
        return RAD_E_HIP;
    }
    return RAD_OK;
}
