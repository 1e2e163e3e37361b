// TEST INFRASTRUCTURE: host emulation of the HIP subset smaug_amd/csrc
// uses, so the tests can compile the kernels for the CPU and check them
// there (tests/test_hostemu.py).  Every thread of a workgroup runs as a
// fiber; __syncthreads and the wavefront shuffle switch fibers; "device"
// memory is host memory with PROT_NONE guard pages on both sides, so an
// out-of-bounds access of a kernel stops the process.  Never part of the
// product library.
#pragma once
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ucontext.h>
#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <functional>
#include <map>
#include <vector>

#define __global__
#define __device__
#define __forceinline__ inline
#define __host__
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)

typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
enum { hipSuccess = 0 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };
#define hipStreamNonBlocking 1

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace emu {
struct Fiber {
  ucontext_t ctx;
  char* stack;
  bool done;
};
extern dim3 g_block, g_grid, g_blockIdx;
extern int g_cur;
extern bool g_in_fiber;
extern ucontext_t g_sched;
extern std::vector<Fiber> g_fibers;
extern double g_shfl[1024];
extern std::function<void()> g_body;
void launch(dim3 grid, dim3 block, std::function<void()> body);
inline dim3 tid() { return dim3((unsigned)g_cur, 0, 0); }
inline void barrier() {
  if (g_in_fiber) swapcontext(&g_fibers[g_cur].ctx, &g_sched);
}
}  // namespace emu

#define threadIdx (emu::tid())
#define blockIdx (emu::g_blockIdx)
#define blockDim (emu::g_block)
#define gridDim (emu::g_grid)
inline void __syncthreads() { emu::barrier(); }
inline double __shfl_xor(double v, int off, int w = 64) {
  const int t = emu::g_cur;
  emu::g_shfl[t] = v;
  emu::barrier();
  const double r = emu::g_shfl[(t & ~(w - 1)) | ((t & (w - 1)) ^ off)];
  emu::barrier();
  return r;
}
inline double __longlong_as_double(long long b) {
  double d;
  memcpy(&d, &b, 8);
  return d;
}
inline long long __double_as_longlong(double d) {
  long long b;
  memcpy(&b, &d, 8);
  return b;
}
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
  unsigned long long o = *p;
  if (v > o) *p = v;
  return o;
}
inline int atomicOr(int* p, int v) {
  int o = *p;
  *p = o | v;
  return o;
}

#define hipLaunchKernelGGL(K, G, B, S, ST, ...) emu::launch(dim3(G), dim3(B), [&]() { K(__VA_ARGS__); })

hipError_t hipMalloc(void** p, size_t n);
template <typename T>
hipError_t hipMalloc(T** p, size_t n) {
  return hipMalloc((void**)p, n);
}
hipError_t hipFree(void* p);
typedef void* hipDeviceptr_t;
// the allocation of this emulation's hipMalloc holding p; memory it did not
// allocate (host arrays the checks pass as device buffers) is reported as a
// range it cannot bound
hipError_t hipMemGetAddressRange(hipDeviceptr_t* base, size_t* size, hipDeviceptr_t p);
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  memmove(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(d, s, n, k);
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, int) {
  *s = (void*)0x1;
  return hipSuccess;
}
inline hipError_t hipStreamCreate(hipStream_t* s) { return hipStreamCreateWithFlags(s, 0); }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
// a file-scope "device" array is host memory here, its address its name's
#define HIP_SYMBOL(s) (&(s))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
inline const char* hipGetErrorString(hipError_t) { return "emu"; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
  *e = (void*)0x1;
  return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0;
  return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }

// the hardware reciprocal: exact on the host (the refinement steps that
// follow it in the kernels then change nothing)
#define __builtin_amdgcn_rcp(x) (1.0 / (x))
// wave priority: scheduling only, nothing to emulate
#define __builtin_amdgcn_s_setprio(p) ((void)0)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
