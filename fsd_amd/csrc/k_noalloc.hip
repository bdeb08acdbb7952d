// Allocation-free variant of fsd_rfu.
//
// fsd_rfu keeps the lubrication pair records in context buffers and regrows
// them with hipMalloc when some particle has more pairs than they hold.  Code
// that may not allocate (HIP graph capture, tight serving loops) calls
// fsd_rfu_noalloc instead: the records go into `scratch`, a file-scope device
// array of fixed size, and a particle with too many pairs is an error instead
// of a regrow.  Every device that runs the library holds its own copy of the
// array until the process ends; its device address is looked up once per
// device (noalloc_init_device, called by fsd_create) and only that address is
// handed to kernels.  Calls that overlap on one device share the array and race.
#include <mutex>

#include "fsd_internal.h"

namespace {
constexpr int NA_MAX_N = FSD_NOALLOC_MAX_N;
constexpr int NA_MAX_PAIRS = FSD_NOALLOC_MAX_PAIRS;
// layout in float4 units: pair records a, b; indices j; counts n; self blocks; flags
constexpr size_t NA_A = 0;
constexpr size_t NA_B = NA_A + (size_t)NA_MAX_N * NA_MAX_PAIRS;
constexpr size_t NA_J = NA_B + (size_t)NA_MAX_N * NA_MAX_PAIRS;            // int, 4 per float4
constexpr size_t NA_N = NA_J + (size_t)NA_MAX_N * NA_MAX_PAIRS / 4;        // int
constexpr size_t NA_SELF = NA_N + (size_t)NA_MAX_N / 4;                    // 9 float4 per particle
constexpr size_t NA_FLAGS = NA_SELF + (size_t)NA_MAX_N * 9;                // 16 ints
constexpr size_t NA_TOTAL = NA_FLAGS + 4;
static_assert(NA_MAX_N % 4 == 0 && NA_MAX_PAIRS % 4 == 0, "scratch carving needs multiples of 4");

constexpr int NA_MAX_DEVICES = 64;
void *g_scratch_ptr[NA_MAX_DEVICES] = {};
std::mutex g_scratch_mu;
}  // namespace

static __device__ __align__(256) float4 scratch[NA_TOTAL];

__global__ void na_clear_kernel(int *flags, int n)
{
    const int i = threadIdx.x;
    if (i < n) flags[i] = 0;
}

// Device address of this device's scratch (once per device, with the device current).
int noalloc_init_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) return FSD_E_HIP;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return 0;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return FSD_E_HIP;
    g_scratch_ptr[dev] = ptr;
    return 0;
}

static void *scratch_ptr()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) return nullptr;
    if (!g_scratch_ptr[dev] && noalloc_init_device() != 0) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[dev];
}

extern "C" int fsd_rfu_noalloc(fsd_ctx *ctx, const float *d_pos4, const float *d_U6, float *d_F6)
{
    if (!ctx) return FSD_E_PARAM;
    if (ctx->N > NA_MAX_N) {
        ctx->err = "fsd_rfu_noalloc: N exceeds the scratch limit of " + std::to_string(NA_MAX_N) + " particles";
        return FSD_E_PARAM;
    }
    if (!ctx->nlist_valid) { ctx->err = "neighbour list not built"; return FSD_E_STATE; }
    float4 *base = static_cast<float4 *>(scratch_ptr());
    if (!base) { ctx->err = "fsd_rfu_noalloc: device address of the scratch array not available"; return FSD_E_HIP; }
    float4 *lub_a = base + NA_A, *lub_b = base + NA_B;
    int *lub_j = reinterpret_cast<int *>(base + NA_J);
    int *lub_n = reinterpret_cast<int *>(base + NA_N);
    float *lub_self = reinterpret_cast<float *>(base + NA_SELF);
    int *flags = reinterpret_cast<int *>(base + NA_FLAGS);
    const float4 *pos = reinterpret_cast<const float4 *>(d_pos4);
    // the overflow flag keeps the previous call's value: clear it (a kernel, not a memset node)
    hipLaunchKernelGGL(na_clear_kernel, dim3(1), dim3(64), 0, ctx->stream, flags, 16);
    FSD_HIP(hipGetLastError());
    {
        const int r = k_lub_pairs_into(ctx, pos, NA_MAX_PAIRS, lub_n, lub_j, lub_a, lub_b, lub_self, flags);
        if (r) return r;
    }
    FSD_HIP(hipMemcpyAsync(ctx->h_flags + 7, flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->h_flags[7] > 0) {
        ctx->err = "fsd_rfu_noalloc: a particle has " + std::to_string(ctx->h_flags[7]) +
                   " lubrication pairs, more than the scratch limit of " + std::to_string(NA_MAX_PAIRS);
        return FSD_E_CAPACITY;
    }
    {
        const int r = k_rfu_apply_with(ctx, d_U6, d_F6, lub_n, lub_j, lub_a, lub_b, lub_self, NA_MAX_PAIRS);
        if (r) return r;
    }
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}
