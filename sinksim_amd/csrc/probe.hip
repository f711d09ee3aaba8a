// libsinksim_probe.so: the device arithmetic the step relies on, evaluated over arrays, so that tests can hold it bit
// for bit against the CPU: the fdlibm routines of fdlibm_dev.hpp as the step calls them, f64 division and square
// root (correctly rounded in IEEE 754, which the CPU's SSE2 instructions are), f32 division, and the f64 -> f32
// conversion. Not part of the batch engine's C ABI; a test hook of the same build flags.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "step_dev.hpp"

namespace {

// op: 0 sin, 1 cos, 2 pow(a, b), 3 scalbn(a, (int)b), 4 a / b, 5 sqrt(a), 6 (double)((float)a / (float)b),
//     7 (double)(float)a, 8 the step's pow call pair (dpow), 9 pose trig as the step calls it (trig(a, b != 0))
__global__ void probe_kernel(int op, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = a[i], y = b[i];
  double r;
  switch (op) {
    case 0: r = ssa::fdl::sin(x); break;
    case 1: r = ssa::fdl::cos(x); break;
    case 2: r = ssa::fdl::pow(x, y); break;
    case 3: r = ssa::fdl::scalbn(x, static_cast<int>(y)); break;
    case 4: r = x / y; break;
    case 5: r = __builtin_sqrt(x); break;
    case 6: r = static_cast<double>(static_cast<float>(x) / static_cast<float>(y)); break;
    case 7: r = static_cast<double>(static_cast<float>(x)); break;
    case 8: r = ssa::dpow(x, y); break;
    default: r = ssa::trig(x, y != 0 ? 1 : 0); break;
  }
  out[i] = r;
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char* ssp_last_error(void) { return g_err.c_str(); }

// Evaluates op over n inputs (host arrays); returns 0 on success.
int ssp_eval(int op, const double* a, const double* b, double* out, long long n) {
  if (n <= 0) return 0;
  double *da = nullptr, *db = nullptr, *dout = nullptr;
  const size_t bytes = static_cast<size_t>(n) * sizeof(double);
  hipError_t e = hipMalloc(&da, bytes);
  if (e == hipSuccess) e = hipMalloc(&db, bytes);
  if (e == hipSuccess) e = hipMalloc(&dout, bytes);
  if (e == hipSuccess) e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, b, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const long long blocks = (n + 255) / 256;
    hipLaunchKernelGGL(probe_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, nullptr, op, da, db, dout, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost);
  if (da) (void)hipFree(da);
  if (db) (void)hipFree(db);
  if (dout) (void)hipFree(dout);
  if (e != hipSuccess) {
    g_err = hipGetErrorString(e);
    return 1;
  }
  return 0;
}

}  // extern "C"
