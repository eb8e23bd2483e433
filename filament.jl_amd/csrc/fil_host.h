// fil_host.h -- host-side helpers shared by the entry points of
// libfilament_hip.so: the error channel behind fil_last_error, device
// buffers, working-precision host arrays and the FFT twiddle tables.
#pragma once

#include <string>
#include <vector>

#include "fil_internal.h"

namespace fil {

// Sets the thread's fil_last_error() message and returns code.
int fail(int code, const std::string& msg);

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fil::fail(FIL_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr double kPi = 3.14159265358979323846;

// Owned device allocation (freed on destruction or re-allocation).
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p; n = o.n;
      o.p = nullptr; o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t bytes) {
    if (p) { (void)hipFree(p); p = nullptr; }
    n = bytes;
    if (!bytes) return hipSuccess;
    return hipMalloc(&p, bytes);
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Host array of the working type (FP32 or FP64, real or interleaved complex).
struct HostArr {
  std::vector<unsigned char> b;
  bool fp64;
  HostArr(size_t n, bool f64, bool complex) : b(n * (f64 ? 8 : 4) * (complex ? 2 : 1), 0), fp64(f64) {}
  void set(size_t i, double v) { if (fp64) reinterpret_cast<double*>(b.data())[i] = v; else reinterpret_cast<float*>(b.data())[i] = (float)v; }
  void setc(size_t i, double re, double im) { set(2 * i, re); set(2 * i + 1, im); }
};

inline hipError_t upload(DevBuf& d, const void* src, size_t bytes) {
  hipError_t e = d.alloc(bytes);
  if (e != hipSuccess) return e;
  return hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice);
}

// Per-pass twiddles of an N-point LDS transform (the layout fft_lds reads)
// appended to tw from entry o; returns the next entry.
size_t append_twiddles(HostArr& tw, size_t o, int N);
// Device twiddle table of an N-point transform; ext: also the half-length
// table and the real-signal factors the row kernel uses.
int build_twiddles(DevBuf& buf, int N, bool f64, bool ext);

}  // namespace fil
