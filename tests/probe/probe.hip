// libsinksim_probe.so: the device arithmetic the step relies on, evaluated over arrays, so that tests can hold it bit
// for bit against the CPU: the fdlibm routines of fdlibm_dev.hpp as the step calls them, f64 division and square
// root (correctly rounded in IEEE 754, which the CPU's SSE2 instructions are), f32 division, and the f64 -> f32
// conversion. Not part of the batch engine's C ABI; a test hook of the same build flags. It also holds a plain
// device copy (ssp_copy_probe) whose known byte counts calibrate the hardware byte counters of a profiler run.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "step_dev.hpp"

namespace {

// op: 0 sin, 1 cos, 2 pow(a, b), 3 scalbn(a, (int)b), 4 a / b, 5 sqrt(a), 6 (double)((float)a / (float)b),
//     7 (double)(float)a, 8 the weir powers as the step and the adjoint call them (b = 1.5: pow_weir(a), b = 0.385:
//     pow_villemonte(a), else the Villemonte factor villemonte(a)), 9 pose trig as the step calls it (trig(a, b != 0))
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
    case 8: r = y == 1.5 ? ssa::pow_weir(x) : (y == 0.385 ? ssa::pow_villemonte(x) : ssa::villemonte(x)); break;
    default: r = ssa::trig(x, y != 0 ? 1 : 0); break;
  }
  out[i] = r;
}

// dst[i] = v over n doubles, and dst[i] = src[i]: 16-byte accesses, grid-stride
__global__ void ssp_fill(double2* __restrict__ dst, double v, long long n2) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n2; i += static_cast<long long>(gridDim.x) * blockDim.x)
    dst[i] = make_double2(v, v);
}
__global__ void ssp_copy(const double2* __restrict__ src, double2* __restrict__ dst, long long n2) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n2; i += static_cast<long long>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
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

// Allocates two device buffers of `bytes` (a multiple of 16), fills the first (ssp_fill: writes `bytes`), copies
// it to the second (ssp_copy: reads and writes `bytes`), frees both. Returns 0 on success.
int ssp_copy_probe(long long bytes) {
  if (bytes <= 0 || bytes % 16 != 0) {
    g_err = "ssp_copy_probe: bytes must be a positive multiple of 16";
    return 1;
  }
  double2 *a = nullptr, *b = nullptr;
  const long long n2 = bytes / 16;
  hipError_t e = hipMalloc(&a, static_cast<size_t>(bytes));
  if (e == hipSuccess) e = hipMalloc(&b, static_cast<size_t>(bytes));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(ssp_fill, dim3(4096), dim3(256), 0, nullptr, a, 1.0, n2);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(ssp_copy, dim3(4096), dim3(256), 0, nullptr, a, b, n2);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e != hipSuccess) {
    g_err = hipGetErrorString(e);
    return 1;
  }
  return 0;
}

}  // extern "C"
