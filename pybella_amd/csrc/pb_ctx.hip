// Context management of the pybella_amd C ABI.
#include <mutex>

#include "pb_common.hpp"

// The allocation-free solve entry point (pb_bicgstab_solve_noalloc) keeps its
// Krylov vectors, block partials and state records in this array instead of
// memory the library allocates.  Every device the library runs on holds its
// own copy until the process ends.  Host code never uses the array's name as
// a device address: the address is looked up once per device below (with the
// device current) and handed to the kernels as an argument.
namespace {
constexpr size_t kScratchBytes = size_t(40) << 20;   // 40 MiB per device
constexpr int kMaxDevices = 64;
}  // namespace
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {
std::mutex g_scratch_mu;
bool g_scratch_done[kMaxDevices] = {};
double* g_scratch_ptr[kMaxDevices] = {};
int g_scratch_err[kMaxDevices] = {};

// Look up this device's copy of ``scratch`` once; the caller has made
// ``device`` current.
void scratch_lookup(pb_ctx* ctx, int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device >= kMaxDevices) {
    ctx->scratch_err = (int)hipErrorInvalidDevice;
    return;
  }
  if (!g_scratch_done[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_done[device] = true;
    g_scratch_err[device] = (int)e;
    g_scratch_ptr[device] = e == hipSuccess ? static_cast<double*>(p) : nullptr;
  }
  ctx->scratch = g_scratch_ptr[device];
  ctx->scratch_bytes = ctx->scratch ? kScratchBytes : 0;
  ctx->scratch_err = g_scratch_err[device];
}
}  // namespace

extern "C" int pb_version(void) { return 100; }

extern "C" int pb_ctx_create(int device, pb_ctx** out) {
  if (!out) return PB_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PB_ERR_HIP;
  if (device < 0 || device >= ndev) return PB_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return PB_ERR_HIP;
  pb_ctx* ctx = new pb_ctx();
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return PB_ERR_HIP;
  }
  ctx->stream = ctx->own_stream;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->n_cu = prop.multiProcessorCount;
  // a failed lookup only disables the allocation-free entry point, which
  // reports it; every other entry point works without the scratch
  scratch_lookup(ctx, device);
  *out = ctx;
  return PB_OK;
}

extern "C" int pb_ctx_destroy(pb_ctx* ctx) {
  if (!ctx) return PB_OK;
  if (ctx->own_stream) {
    (void)hipStreamSynchronize(ctx->own_stream);
    (void)hipStreamDestroy(ctx->own_stream);
  }
  delete ctx;
  return PB_OK;
}

extern "C" int pb_ctx_set_stream(pb_ctx* ctx, void* hip_stream) {
  if (!ctx) return PB_ERR_ARG;
  ctx->stream = (hipStream_t)hip_stream;   // NULL = the device's null stream
  return PB_OK;
}

extern "C" int pb_ctx_synchronize(pb_ctx* ctx) {
  if (!ctx) return PB_ERR_ARG;
  PB_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PB_OK;
}

extern "C" const char* pb_last_error(const pb_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }
