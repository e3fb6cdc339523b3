// Device evaluation of the deterministic elementary functions
// (include/mtx_fmath.h) the path tracer uses, for host-side reproduction checks.
#include "../../include/mtx_fmath.h"
#include "mtx_internal.h"

namespace mtx {

__global__ void k_math_eval(uint32_t kind, const float *__restrict__ x, const float *__restrict__ y,
                            float *__restrict__ out, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    float a = x[i], b = y[i], r;
    switch (kind) {
      case 0: r = mtxfm::sinf(a); break;
      case 1: r = mtxfm::cosf(a); break;
      case 2: r = mtxfm::powf(a, b); break;
      case 3: r = mtxfm::atan2f(a, b); break;
      case 4: r = mtxfm::acosf(a); break;
      case 5: r = mtxfm::expf(a); break;
      case 6: r = mtxfm::sincosf(a).s; break;
      default: r = mtxfm::sincosf(a).c; break;
    }
    out[i] = r;
  }
}

}  // namespace mtx

extern "C" int32_t mtx_math_eval(uint32_t kind, const float *x, const float *y, float *out, uint64_t n,
                                 void *stream) {
  try {
    MTX_REQUIRE(kind <= 7, MTX_INVALID_ARGUMENT, "kind must be 0..7");
    MTX_REQUIRE(n == 0 || (x && y && out), MTX_INVALID_ARGUMENT, "NULL buffer");
    if (!n) return MTX_OK;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(mtx::k_math_eval, dim3((uint32_t)blocks), dim3(256), 0, (hipStream_t)stream, kind, x, y, out, n);
    MTX_HIP(hipGetLastError());
    return MTX_OK;
  } catch (const mtx::Error &e) {
    return e.code;
  }
}
