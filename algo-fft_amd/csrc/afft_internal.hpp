// Host-side internals of the algofft_hip library: error state, the kernel
// registry, and the context / buffer / stream / plan objects behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/algofft_hip.h"
#include "kernels.hpp"

namespace afft {

// ---- errors -------------------------------------------------------------------
int set_error(int status, const std::string& msg);
void clear_error();

#define AFFT_HIP(expr)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::afft::set_error(AFFT_ERR_DEVICE, std::string(#expr) + ": " +            \
                                                          hipGetErrorString(e_));            \
    } while (0)

// ---- kernel registry ------------------------------------------------------------
const KernelInfo* find_kernel(int prec, int n, int var, int dir);
int max_single_pass(int prec);        // largest power of two held in LDS

// ---- objects ----------------------------------------------------------------
// One wisdom record: the reference's WisdomEntry (planner/wisdom.go:17-30),
// keyed by size, precision and a device feature word.
struct WisdomEntry {
    int64_t size = 0;
    int precision = 0;
    uint64_t features = 0;
    std::string algorithm;
    int64_t timestamp = 0;
};

struct Mlp;

struct Context {
    int device = 0;
    hipDeviceProp_t props{};
    std::mutex mu;                                                   // guards wisdom and model
    std::map<std::tuple<int64_t, int, uint64_t>, WisdomEntry> wisdom;
    Mlp* model = nullptr;                                            // cost model, not owned
    uint64_t features() const;                                       // gfx arch number, e.g. 0x950
};

struct Buffer {
    Context* ctx = nullptr;
    int64_t len = 0;
    int precision = 0;
    void* ptr = nullptr;
};

struct Stream {
    Context* ctx = nullptr;
    hipStream_t s = nullptr;
};

// Where a step reads or writes.
// TMP0/TMP1 are the scratch of the four-step, Bluestein and gather paths;
// WORK holds the compact complex array of an N-D inverse real transform and
// WORK2 the half-length spectra of real rows without a fused kernel.
enum Role { ROLE_SRC = 0, ROLE_DST = 1, ROLE_TMP0 = 2, ROLE_TMP1 = 3, ROLE_WORK = 4, ROLE_WORK2 = 5, ROLE_COUNT = 6 };
constexpr int NTMP = ROLE_COUNT - 2;

// One kernel launch of a plan.  Pointers inside `args` are filled at exec
// time from the roles; `in_off`/`out_off` are byte offsets from the role base.
struct Step {
    const void* fn = nullptr;     // __global__ entry
    dim3 grid, block;
    int lds_bytes = 0;            // dynamic LDS (0: static only)
    int in_role = ROLE_SRC, out_role = ROLE_DST;
    int64_t in_off = 0, out_off = 0;
    std::vector<unsigned char> arg;   // kernel argument block (a POD struct)
    int in_field = 0, out_field = 0;  // byte offsets of the in/out pointers in arg
    int nt_field = -1;                // byte offset of KArgs::nt in arg, if any
    std::string name;
};

// Bluestein tables of one (n, padded m, filter layout), shared by a plan's
// two directions: the FP64 chirp on the host and the device tables.
struct BlueCache {
    std::vector<double> ch;
    const void* chirp = nullptr;
    const void* filt[2] = {nullptr, nullptr};
};

struct Plan {
    Context* ctx = nullptr;
    hipStream_t stream = nullptr;
    afft_plan_desc desc{};
    int64_t n = 0;                  // points of one transform (product of dims)
    int64_t in_elems[2] = {0, 0};   // [fwd, inv] input length (elements of that side)
    int64_t out_elems[2] = {0, 0};
    int in_esize[2] = {0, 0};       // bytes per element of each side
    int out_esize[2] = {0, 0};
    std::vector<Step> steps[2];     // [0] forward, [1] inverse
    std::vector<void*> owned;       // device tables
    size_t tmp_bytes[NTMP] = {0, 0, 0, 0};
    void* tmp[NTMP] = {nullptr, nullptr, nullptr, nullptr};
    void* stage_in = nullptr;
    void* stage_out = nullptr;
    void* full = nullptr;             // half spectrum of the full-spectrum transforms (full.hip)
    size_t full_bytes = 0;
    size_t stage_in_bytes = 0, stage_out_bytes = 0;
    struct HostPipe* pipe = nullptr;   // chunked host-slice execution (runtime.cpp)
    // Workspace ordering across streams: the last stream that used tmp[] and
    // an event recorded on it after that use.
    bool ws_used = false;
    hipStream_t ws_stream = nullptr;
    hipEvent_t ws_done = nullptr;
    std::string algo;
    std::map<std::tuple<int64_t, int64_t, int64_t>, BlueCache> blue_cache;
    std::mutex mu;
    ~Plan();
};
void destroy_host_pipe(HostPipe* h);

int build_plan(Plan& p);

// Helpers shared by the planners (plan.cpp).
int upload_table(Plan& p, const void* host, size_t bytes, const void** dev);
std::vector<int> smooth_schedule(int64_t n);   // empty unless 2/3/5/7-smooth
int64_t next_pow2(int64_t n);
struct BlueTables {
    const void* chirp = nullptr;                 // n entries, forward sign
    const void* filt[2] = {nullptr, nullptr};    // [fwd, inv], m entries, 1/m folded in
    const void* post[2] = {nullptr, nullptr};    // [fwd, inv]: chirp, conj(chirp)*inv_scale
};
// tr_m1 > 0 stores the filter spectra row-major as [k1][k2], k = k1 + tr_m1*k2
int blue_tables(Plan& p, int64_t n, int64_t m, double inv_scale, BlueTables& bt, int64_t tr_m1 = 0);
// W_L^i = hi[i >> shift] * lo[i & (2^shift - 1)], both FP64-generated
int split_twiddles(Plan& p, int prec, int64_t L, int* shift, const void** lo, const void** hi);
int max_single_pass(int prec);
int gen_cap(int prec);
int gen_row_cap(int prec);          // contiguous rows: one in-place pass
int64_t gen_fourstep_split(int64_t n, int prec);   // default n1 of a smooth four-step

// ---- planning knowledge (tune.cpp, mlp.hip) --------------------------------------
constexpr int NFEAT = 32;
struct Candidate {
    int strategy;
    int64_t split;
};
// Strategies worth timing for a 1-D transform of length n.
std::vector<Candidate> plan_candidates(const afft_plan_desc& d);
// Feature vector of (shape, candidate) for the cost model.
void plan_features(const afft_plan_desc& d, const Candidate& c, float* out);
std::string strategy_name(const Candidate& c);
bool parse_strategy(const std::string& algo, Candidate* c);
// Predicted log-runtimes of candidates (GPU inference); 0 on success.
int mlp_predict_host(Mlp* m, const float* x, int64_t n, float* y);
int exec_plan(Plan& p, int dir, void* dst, const void* src, hipStream_t s);
// The steps of direction di (0 forward, 1 inverse) with explicit workspace
// buffers; allocates nothing and records no event.
int launch_steps(Plan& p, int di, void* dst, const void* src, void* const* tmp, hipStream_t s);
// Device address of the static scratch array of afft_plan_exec_device_noalloc
// on `device`, looked up once per device (noalloc.hip).
int noalloc_scratch(int device, void** out);

// FP64 roots of unity W_L^m = exp(-2πi m / L), exact integer reduction of m
// (and the same in long double, before the final rounding).
void root_of_unity(int64_t m, int64_t L, double* re, double* im);
void root_of_unity_ld(int64_t m, int64_t L, long double* re, long double* im);

}  // namespace afft

// The opaque handle types of the C ABI are the internal objects.
struct afft_context_s : afft::Context {};
struct afft_buffer_s : afft::Buffer {};
struct afft_stream_s : afft::Stream {};
struct afft_plan_s : afft::Plan {};
