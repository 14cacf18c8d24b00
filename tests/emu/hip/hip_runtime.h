// Host stand-in for <hip/hip_runtime.h>, used only by the CPU block emulator of the
// generic FFT kernels (tests/emu/fft_block_emu.cpp): kernels compile as plain C++
// functions, each emulated GPU thread runs them in its own std::thread, and
// __syncthreads is a workgroup barrier that reports a deadlock instead of hanging.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdlib>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__
#define __align__(x)

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
};
extern thread_local dim3 threadIdx;
extern thread_local dim3 blockIdx;
extern dim3 blockDim;
extern dim3 gridDim;
void emu_syncthreads();
#define __syncthreads() emu_syncthreads()

struct float2 {
    float x, y;
};
struct double2 {
    double x, y;
};
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline float __expf(float x) { return std::exp(x); }
using std::exp;

typedef int hipError_t;
constexpr hipError_t hipSuccess = 0;
constexpr unsigned hipHostMallocMapped = 1, hipHostMallocPortable = 2;
inline const char* hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipMalloc(void** p, size_t n) { *p = std::malloc(n); return hipSuccess; }
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned) { *p = std::malloc(n); return hipSuccess; }
inline hipError_t hipHostFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipHostGetDevicePointer(void** d, void* h, unsigned) { *d = h; return hipSuccess; }
