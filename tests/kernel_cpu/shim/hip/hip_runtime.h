/* TEST INFRASTRUCTURE ONLY.  Lets a generated kernel translation unit compile
 * as plain host C++ so its kernels can run on the CPU (one call per thread
 * index), e.g. under AddressSanitizer to check every memory access of every
 * generated schedule against exact-size buffers. */
#pragma once
#include <stddef.h>

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
typedef int hipError_t;
typedef void *hipStream_t;
#define __constant__
static const hipError_t hipSuccess = 0;
inline dim3 blockIdx, blockDim, gridDim;
inline thread_local dim3 threadIdx;
inline int min(int a, int b) { return a < b ? a : b; }
#define BHAH_PIN(x) ((void)0)
#define BHAH_PINS(...) ((void)0)
#define __builtin_nontemporal_load(p) (*(p))
#define __builtin_nontemporal_store(v, p) (*(p) = (v))
#define __shared__ static
inline hipError_t hipGetLastError() { return hipSuccess; }

#define __global__
#define __device__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __builtin_amdgcn_sched_barrier(x) ((void)0)

// Buffer resources: a base pointer and a byte size; loads outside it abort.
struct __amdgpu_buffer_rsrc_t {
  const char *base;
  unsigned long bytes;
};
inline __amdgpu_buffer_rsrc_t __builtin_amdgcn_make_buffer_rsrc(void *p, short, int n, int) {
  return {(const char *)p, (unsigned long)(unsigned)n};
}
inline double __builtin_amdgcn_raw_buffer_load_b64(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, int) {
  const unsigned long off = (unsigned long)vo + so;
  if (r.base == nullptr || off + 8 > r.bytes) __builtin_trap();
  double v;
  __builtin_memcpy(&v, r.base + off, 8);
  return v;
}

struct bhah_u4 {
  unsigned v[4];
  unsigned operator[](int i) const { return v[i]; }
};
inline bhah_u4 __builtin_amdgcn_raw_buffer_load_b128(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, int) {
  const unsigned long off = (unsigned long)vo + so;
  if (r.base == nullptr || off + 16 > r.bytes) __builtin_trap();
  bhah_u4 v;
  __builtin_memcpy(&v, r.base + off, 16);
  return v;
}

#ifdef BHAH_CPU_THREADED
/* Kernels with workgroup barriers: every thread of a block runs as its own host
 * thread and __syncthreads() is a barrier over the block's threads (blocks run
 * one after another, so __shared__ arrays may be plain statics). */
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>
struct bhah_cpu_barrier_t {
  std::mutex m;
  std::condition_variable cv;
  unsigned n = 0, count = 0, gen = 0;
  void wait() {
    std::unique_lock<std::mutex> l(m);
    const unsigned g = gen;
    if (++count == n) {
      count = 0;
      gen++;
      cv.notify_all();
    } else {
      cv.wait(l, [&] { return g != gen; });
    }
  }
};
inline bhah_cpu_barrier_t bhah_cpu_bar;
#define __syncthreads() bhah_cpu_bar.wait()
#define BHAH_BARRIER() bhah_cpu_bar.wait()
#define BHAH_BARRIER_PHASE(n) bhah_cpu_bar.wait()
#define BHAH_VMWAIT(n) ((void)0)
// every lane of a wave holds the same value where the kernels use it (the wave index)
#define __builtin_amdgcn_readfirstlane(x) (x)
// global-to-LDS load of 16 bytes per lane: lane l writes dst + 16 l (dst is the wave's base)
#define BHAH_DMA16(src, dst) bhah_cpu_dma16((src), (dst))
#define BHAH_DMA16O(base, off, dst) bhah_cpu_dma16((base) + (off), (dst))
inline void bhah_cpu_dma16(const void *src, double *dst) {
  __builtin_memcpy(dst + 2 * (threadIdx.x % 64), src, 16);
}

template <typename K, typename... A>
void cpu_launch(K kernel, dim3 g, dim3 b, A... args) {
  gridDim = g;
  blockDim = b;
  bhah_cpu_bar.n = b.x * b.y * b.z;
  for (unsigned bz = 0; bz < g.z; bz++)
    for (unsigned by = 0; by < g.y; by++)
      for (unsigned bx = 0; bx < g.x; bx++) {
        blockIdx = dim3(bx, by, bz);
        std::vector<std::thread> ts;
        for (unsigned tz = 0; tz < b.z; tz++)
          for (unsigned ty = 0; ty < b.y; ty++)
            for (unsigned tx = 0; tx < b.x; tx++)
              ts.emplace_back([=] {
                threadIdx = dim3(tx, ty, tz);
                kernel(args...);
              });
        for (auto &t : ts) t.join();
      }
}
#else
#define __syncthreads() __builtin_trap()
template <typename K, typename... A>
void cpu_launch(K kernel, dim3 g, dim3 b, A... args) {
  gridDim = g;
  blockDim = b;
  for (unsigned bz = 0; bz < g.z; bz++)
    for (unsigned by = 0; by < g.y; by++)
      for (unsigned bx = 0; bx < g.x; bx++)
        for (unsigned tz = 0; tz < b.z; tz++)
          for (unsigned ty = 0; ty < b.y; ty++)
            for (unsigned tx = 0; tx < b.x; tx++) {
              blockIdx = dim3(bx, by, bz);
              threadIdx = dim3(tx, ty, tz);
              kernel(args...);
            }
}
#endif
#define hipLaunchKernelGGL(k, g, b, shmem, stream, ...) cpu_launch(k, g, b, __VA_ARGS__)
