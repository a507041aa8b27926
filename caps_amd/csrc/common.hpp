// Internal: HIP error handling, device buffers and the context of the capsgpu library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace capsgpu {

// errors thrown inside the library and turned into return codes at the C boundary (as capi's guard(),
// capi/src/caps_c.cpp:373-383)
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess)
    throw Error(-3, std::string("HIP error in ") + what + " (" + file + ":" + std::to_string(line) + "): " + hipGetErrorString(e));
}
#define HIPCHECK(x) ::capsgpu::hip_check((x), #x, __FILE__, __LINE__)
#ifdef CAPSGPU_DEBUG_GUARD
// debug builds only: guarded allocations and a check of every guard after each launch
void* guard_alloc(size_t bytes);
void guard_free(void* p);
void guard_check(const char* file, int line);
#define LAUNCHCHECK() (::capsgpu::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__), ::capsgpu::guard_check(__FILE__, __LINE__))
#else
#define LAUNCHCHECK() ::capsgpu::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)
#endif

namespace {
// zeros nbytes from p: 16-byte words, then the tail bytes (device allocations are 256-byte aligned). A kernel rather
// than hipMemsetAsync, so work captured into a HIP graph holds no memset node.
__global__ void k_zero_bytes(unsigned char* __restrict__ p, size_t nbytes) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x, w = nbytes / 16;
  if (i < w) reinterpret_cast<uint4*>(p)[i] = make_uint4(0u, 0u, 0u, 0u);
  else if (i < w + (nbytes & 15)) p[16 * w + (i - w)] = 0;
}
}  // namespace

// A fixed block of device memory handed out front to back (256-byte aligned pieces), never freed piecewise: the
// memory of evaluations that may not allocate. Taking more than it holds is an error that names the limit.
struct Arena {
  unsigned char* base = nullptr;
  size_t cap = 0, used = 0;
  void* take(size_t bytes) {
    const size_t at = (used + 255) & ~size_t(255);
    if (at > cap || bytes > cap - at)
      throw Error(-1, "the allocation-free evaluation needs more than its " + std::to_string(cap >> 20) +
                          " MiB of device scratch memory (" + std::to_string((at + bytes + (1u << 20) - 1) >> 20) +
                          " MiB asked so far): use a smaller system or the allocating entry point");
    used = at + bytes;
    return base + at;
  }
};

// a device array that owns its memory and grows (never shrinks) on demand; bound to an Arena it takes its memory
// from there instead and frees nothing
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0, cap = 0;
  Arena* arena = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), cap(o.cap), arena(o.arena) { o.p = nullptr, o.n = o.cap = 0, o.arena = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p, n = o.n, cap = o.cap, arena = o.arena;
      o.p = nullptr, o.n = o.cap = 0, o.arena = nullptr;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void swap(DevBuf& o) noexcept {
    std::swap(p, o.p), std::swap(n, o.n), std::swap(cap, o.cap), std::swap(arena, o.arena);
  }
  void release() {
    if (arena) {
      p = nullptr, n = cap = 0;
      return;
    }
#ifdef CAPSGPU_DEBUG_GUARD
    if (p) guard_free(p);
#else
    // work queued on a non-blocking stream may still use the memory: wait for the device before giving it back
    if (p) (void)hipDeviceSynchronize(), (void)hipFree(p);
#endif
    p = nullptr, n = cap = 0;
  }
  // size n; contents are kept only when no reallocation happens
  void resize(size_t m) {
    if (m > cap && arena) {
      p = static_cast<T*>(arena->take(m * sizeof(T)));
      cap = m;
    } else if (m > cap) {
      release();
#ifdef CAPSGPU_DEBUG_GUARD
      const size_t c = m ? m : 1;
#else
      const size_t c = m + m / 8 + 64;
#endif
#ifdef CAPSGPU_DEBUG_GUARD
      p = static_cast<T*>(guard_alloc(c * sizeof(T)));
#else
      HIPCHECK(hipMalloc(&p, c * sizeof(T)));
#endif
      cap = c;
    }
    n = m;
  }
  // ordered on stream s and complete on return (the host array may be a temporary)
  void upload(const T* h, size_t m, hipStream_t s) {
    resize(m);
    if (m) {
      HIPCHECK(hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, s));
      HIPCHECK(hipStreamSynchronize(s));
    }
  }
  void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
  void download(T* h, size_t m, hipStream_t s) const {
    if (m) HIPCHECK(hipMemcpyAsync(h, p, m * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  std::vector<T> to_host(hipStream_t s) const {
    std::vector<T> h(n);
    download(h.data(), n, s);
    HIPCHECK(hipStreamSynchronize(s));
    return h;
  }
  void zero(hipStream_t s) {
    if (!n) return;
    const size_t b = n * sizeof(T), items = b / 16 + (b & 15);
    hipLaunchKernelGGL(k_zero_bytes, dim3(unsigned((items + 255) / 256)), dim3(256), 0, s, reinterpret_cast<unsigned char*>(p), b);
    HIPCHECK(hipGetLastError());
  }
};

constexpr uint32_t kPad = 0xffffffffu;
constexpr double kCoulomb = 332.06371;    // kcal·Å/(mol·e²) (core/include/caps/kspace.hpp:12)
constexpr double kPi = 3.14159265358979323846;
constexpr double kBoltz = 0.0019872041;   // kcal/(mol·K) (core/src/dynamics.cpp:22)
constexpr double kAcc = 4.184e-4;         // (kcal/mol/Å)/(g/mol) → Å/fs² (core/src/dynamics.cpp:23)
constexpr double kAtm = 68568.415;        // kcal/(mol·Å³) → atm (core/src/dynamics.cpp:24)

inline unsigned blocks_for(size_t n, unsigned bs) { return unsigned((n + bs - 1) / bs); }

// the work item of workgroup b of G: workgroups go round-robin over the 8 XCDs, so XCD x is given the x-th contiguous
// eighth of the items, and neighbouring items (which share data) meet in one XCD's L2 (a bijection on [0, G))
__device__ inline int64_t xcd_block(int64_t b, int64_t G) {
  const int64_t x = b % 8, i = b / 8, q = G / 8, r = G % 8;
  return x * q + (x < r ? x : r) + i;
}

}  // namespace capsgpu
