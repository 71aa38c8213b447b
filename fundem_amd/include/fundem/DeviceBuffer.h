// RAII device buffer owned by the host containers (the role of
// HostDeviceArray1D's device half, myHostDeviceArray.h:160-260). Allocation
// happens at set-up only; a step allocates nothing.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/fundem_amd.h"

namespace fundem {

inline void hipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();  // do not leave the error for the next launch to report
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  }
}

// A launcher's return code (FDM_OK or an error code) as an exception.
inline void fdmCheck(int r, const char* what) {
  if (r != FDM_OK) throw std::runtime_error(std::string(what) + " failed with code " + std::to_string(r));
}

template <class T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }

  // Resize without preserving contents; zero-filled when requested.
  void allocate(size_t n, bool zero = true) {
    if (n == n_ && p_) {
      if (zero) hipCheck(hipMemset(p_, 0, n_ * sizeof(T)), "hipMemset");
      return;
    }
    release();
    if (n > 0) {
      hipCheck(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
      n_ = n;
      if (zero) hipCheck(hipMemset(p_, 0, n_ * sizeof(T)), "hipMemset");
    }
  }
  void upload(const T* host, size_t n, hipStream_t s) {
    if (n > n_) allocate(n, false);
    if (n) hipCheck(hipMemcpyAsync(p_, host, n * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void download(T* host, size_t n, hipStream_t s) const {
    if (n) hipCheck(hipMemcpyAsync(host, p_, n * sizeof(T), hipMemcpyDeviceToHost, s), "download");
  }
  std::vector<T> toHost(hipStream_t s, size_t n = (size_t)-1) const {
    if (n == (size_t)-1) n = n_;
    std::vector<T> h(n);
    download(h.data(), n, s);
    hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
    return h;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  double memoryGB() const { return double(n_ * sizeof(T)) / 1073741824.0; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace fundem
