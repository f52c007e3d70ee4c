// Plan execution that allocates nothing: afft_plan_exec_device_noalloc.
//
// afft_plan_exec_device gives a plan's scratch roles (TMP0/TMP1 of the
// four-step, Bluestein and gather chains, WORK/WORK2 of the real N-D plans)
// device memory that it allocates on the plan's first execution.  Callers
// that cannot allocate at execution time (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) use this variant:
// the same steps and kernels, with the roles carved out of one static
// device array of kNoAllocScratchBytes per GPU.  Each GPU the library runs on
// holds that array until the process ends.
//
// The array's name in host code is a host-side stand-in, not a device
// address: its device address is looked up once per device with
// hipGetSymbolAddress (at context creation, else on the first call on that
// device) and only that address is handed to kernels.  Every call on one
// GPU shares the array; no kernel of the plans reads scratch it has not
// written in the same execution, so nothing needs clearing between calls,
// and calls that overlap in time on one GPU race on it.
#include <mutex>

#include "afft_internal.hpp"

// 64 MiB: the scratch of a 2^20-point complex64 four-step over a batch of 8,
// or of a complex64 Bluestein chain of padded length 2^21.
static constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace afft {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Byte offset of each scratch role inside the array, 256-byte aligned.
size_t carve(const Plan& p, size_t off[NTMP]) {
    size_t total = 0;
    for (int t = 0; t < NTMP; ++t) {
        off[t] = total;
        total += (p.tmp_bytes[t] + 255) & ~size_t(255);
    }
    return total;
}
}  // namespace

// Device address of this device's scratch array, looked up once per device
// with that device current.  The caller's current device is restored.
int noalloc_scratch(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices)
        return set_error(AFFT_ERR_INVALID_ARGUMENT, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        int prev = 0;
        AFFT_HIP(hipGetDevice(&prev));
        AFFT_HIP(hipSetDevice(device));
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        (void)hipSetDevice(prev);
        if (e != hipSuccess || !ptr)
            return set_error(AFFT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[device] = ptr;
    }
    *out = g_ptr[device];
    return AFFT_OK;
}

}  // namespace afft

using namespace afft;

extern "C" {

int64_t afft_noalloc_scratch_bytes(void) { return (int64_t)kNoAllocScratchBytes; }

int afft_plan_exec_device_noalloc(afft_plan p, int direction, void* dst, const void* src, void* stream) {
    if (!p) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "nil plan");
    if (!dst || !src) return set_error(AFFT_ERR_NIL_SLICE, "dst or src is nil");
    if (direction != AFFT_FORWARD && direction != AFFT_INVERSE)
        return set_error(AFFT_ERR_INVALID_ARGUMENT, "direction must be -1 or +1");
    const int di = direction == AFFT_INVERSE ? 1 : 0;
    if (p->steps[di].empty()) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "direction not available for this plan kind");
    size_t off[NTMP];
    const size_t need = carve(*p, off);
    if (need > kNoAllocScratchBytes)
        return set_error(AFFT_ERR_INVALID_LENGTH, "plan needs " + std::to_string(need) +
                                                      " bytes of scratch; the static scratch holds " +
                                                      std::to_string(kNoAllocScratchBytes));
    std::lock_guard<std::mutex> lock(p->mu);
    void* base = nullptr;
    int rc = noalloc_scratch(p->ctx->device, &base);
    if (rc) return rc;
    AFFT_HIP(hipSetDevice(p->ctx->device));
    void* tmp[NTMP];
    for (int t = 0; t < NTMP; ++t) tmp[t] = p->tmp_bytes[t] ? (char*)base + off[t] : nullptr;
    rc = launch_steps(*p, di, dst, src, tmp, stream ? (hipStream_t)stream : p->stream);
    if (rc == AFFT_OK) clear_error();
    return rc;
}

}  // extern "C"
