// Context management of the pybella_amd C ABI.
#include "pb_common.hpp"

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
