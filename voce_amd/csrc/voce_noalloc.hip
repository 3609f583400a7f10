// voce_noalloc.hip -- voce_increment_noalloc: one strain increment whose
// per-increment workspace comes from a file-scope device array instead of
// the context's own memory, so the call allocates nothing.
//
// The workspace of an increment (pre-update stress, crystal-frame K_c,
// spectra, skew sums, block partials) is written before it is read inside
// the increment: the skew sums by the first FFT iteration, K_c by the first
// Newton launch, the spectra, partials and pre-update stress by every
// iteration.  So the array needs no clearing between calls, and what a
// previous call left in it is never used.
//
// One array per device, shared by every call on that device: calls that
// overlap on one device (two host threads, or two streams) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string.h>

#include "../../include/voce.h"
#include "voce_host.h"

// Host code must not use this name as a device address: it is a host-side
// stand-in.  The device address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) unsigned char scratch[VOCE_NOALLOC_MAX_BYTES];

namespace {

constexpr size_t kScratchBytes = (size_t)VOCE_NOALLOC_MAX_BYTES;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
void *g_scratch[kMaxDevices];   // address of `scratch` on each device

// The device address of this device's array, looked up once per device
// with that device current.
int scratch_address(int dev, char **out)
{
    if (dev < 0 || dev >= kMaxDevices) return VOCE_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return VOCE_ERR_DEVICE;
        g_scratch[dev] = p;
    }
    *out = (char *)g_scratch[dev];
    return 0;
}

}  // namespace

extern "C" int voce_increment_noalloc(voce_ctx *c, voce_incr_report *rep)
{
    // contexts of voce_ctx_create only: batch members and a server's ids
    // take their workspace from their owner
    if (!c || c->on_tab || !c->d.spec) return VOCE_ERR_ARG;
    if (c->iter_max < 1) return VOCE_ERR_ARG;
    if (voce_ws_bytes(c->d) > kScratchBytes) return VOCE_ERR_SCRATCH;
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) return VOCE_ERR_DEVICE;
    if (prev != c->device && hipSetDevice(c->device) != hipSuccess) return VOCE_ERR_DEVICE;
    int rc = 0;
    char *base = nullptr;
    if ((rc = scratch_address(c->device, &base)) == 0) {
        // the context's descriptor with its workspace fields pointing into
        // the array, for this increment only
        RveDev d = c->d;
        voce_ws_carve(d, base);
        *c->h_tab = d;
        if (hipMemcpyAsync(c->dtab, c->h_tab, sizeof(RveDev), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
            rc = VOCE_ERR_DEVICE;
        } else {
            rc = voce_run_increment(c, rep, 0);
            // back to the context's own workspace (voce_run_increment has
            // synchronised the stream, and waits again on any error path
            // below, so no kernel still reads the staged descriptor)
            (void)hipStreamSynchronize(c->stream);
            *c->h_tab = c->d;
            if (hipMemcpyAsync(c->dtab, c->h_tab, sizeof(RveDev), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                if (!rc) rc = VOCE_ERR_DEVICE;
            }
        }
    }
    if (prev != c->device) (void)hipSetDevice(prev);
    return rc;
}
