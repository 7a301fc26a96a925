// ps_render_noalloc: the single-kernel render (ps_render) for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid allocation). No counterpart in the reference.
//
// ps_render stages each frame's parameter block in a device ring the library
// allocates on first use. This entry point stages it instead in `scratch`, a
// file-scope device array of kScratchBytes (one parameter block) that every
// GPU holds from load to process exit, and copies the block there with small
// kernels whose arguments carry it (no host staging buffer either). It calls
// no hipMalloc, hipFree or other allocating routine.
//
// Every call on a GPU shares that GPU's array and leaves the last call's
// block in it: calls that overlap on one GPU (several contexts or streams at
// once) race. Calls on one stream are ordered and safe.
#include <cstring>
#include <mutex>
#include <string>

#include "ps_device.h"
#include "ps_host.h"
#include "ps_launch.h"

namespace {

constexpr size_t kScratchBytes = 4096;
static_assert(sizeof(RenderParams) <= kScratchBytes, "the parameter block must fit the scratch array");
static_assert(sizeof(RenderParams) % 4 == 0, "the parameter block is copied as 32-bit words");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// The array's device address on each GPU (host code must use these, never
// the array's name, which is a host-side stand-in).
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_mu;

// 1 KiB of a parameter block, passed by value.
struct Chunk {
    uint32_t w[256];
};

__global__ __launch_bounds__(256) void k_store_chunk(const Chunk c, uint32_t* __restrict__ dst, int nwords) {
    const int i = (int)threadIdx.x;
    if (i < nwords) dst[i] = c.w[i];
}

}  // namespace

ps_status psi::noalloc_init(int device) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (device < 0 || device >= kMaxDevices)
        return fail(PS_ERR_UNSUPPORTED, "scratch set-up: device index beyond the supported 64 GPUs");
    if (g_scratch[device]) return PS_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(PS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = p;
    return PS_OK;
}

extern "C" ps_status ps_render_noalloc(ps_ctx* c, const float* depth, const ps_image_out* out) {
    using namespace psi;
    if (!c || !out) return fail(PS_ERR_INVALID_ARGUMENT, "ps_render_noalloc: NULL argument");
    if (c->half_res || c->taa)
        return fail(PS_ERR_UNSUPPORTED, "ps_render_noalloc: renders the single-kernel path only (half resolution "
                                        "and TAA need the cloud RT the library allocates)");
    if (c->delta)
        return fail(PS_ERR_UNSUPPORTED, "ps_render_noalloc: the perturbation volume needs the bounds buffer the "
                                        "library allocates (use ps_render)");
    ps_status s;
    if ((s = check_render_ready(c))) return s;
    void* p;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        p = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
    }
    if (!p) return fail(PS_ERR_NOT_READY, "ps_render_noalloc: no scratch address for this context's GPU");
    RenderParams P;
    if ((s = fill_render_params(c, c->sc, c->luts, depth, out, P))) return s;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&P);
    const int words = (int)(sizeof(RenderParams) / 4);
    for (int off = 0; off < words; off += 256) {
        Chunk ch;
        const int n = words - off < 256 ? words - off : 256;
        std::memcpy(ch.w, src + off, sizeof(uint32_t) * n);
        hipLaunchKernelGGL(k_store_chunk, dim3(1), dim3(256), 0, c->stream, ch, static_cast<uint32_t*>(p) + off, n);
        PS_HIP(hipGetLastError());
    }
    PS_HIP(ps_launch_render(static_cast<const RenderParams*>(p), P.W, P.H, false, false, false, c->stream));
    return PS_OK;
}
