// Host-side launch helpers shared by the launcher translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "fundem_amd.h"

namespace fdm {

inline unsigned blocksFor(long long n, int block) {
  return (unsigned)((n + block - 1) / block);
}

inline int launchStatus() {
  return hipGetLastError() == hipSuccess ? FDM_OK : FDM_ELAUNCH;
}

inline size_t alignUp(size_t x, size_t a = 256) { return (x + a - 1) / a * a; }

// Bump allocator over a caller-provided workspace.
struct Carve {
  char* base;
  size_t cap;
  size_t used = 0;
  bool ok = true;
  Carve(void* b, size_t c) : base((char*)b), cap(c) {}
  template <class T>
  T* take(size_t count) {
    size_t bytes = alignUp(count * sizeof(T));
    if (used + bytes > cap) {
      ok = false;
      return nullptr;
    }
    T* p = (T*)(base + used);
    used += bytes;
    return p;
  }
};

// Number of persistent workgroups for grid-stride kernels.
int persistentBlocks(int wavesPerBlock);
// Launch shapes chosen by fdm_setTuning (contact.hip).
const fdm_tuning& tuning();

// A device flag that switches a kernel off: the kernel returns at once unless
// the flag is null or *flag != 0. Steps captured in a graph launch every
// kernel; gates let the device skip work it decided is not needed.
__device__ __forceinline__ bool gateOff(const int* gate) { return gate && *gate == 0; }

// Device-wide scans over ints (sort.hip). out may alias in.
size_t scanWorkspace(size_t n);
int scanInts(const int* in, int* out, size_t n, bool inclusive, int* total, void* ws, size_t wsBytes,
             hipStream_t stream, const int* gate = nullptr);

// Radix sort and cell ranges (sort.hip). nDev, when not null, holds the
// element count on the device (n bounds it).
size_t radixWorkspace(size_t n);
int radixSort(unsigned* keys, int* vals, size_t n, const int* nDev, int keyBits, void* ws,
              size_t wsBytes, hipStream_t stream, const int* gate = nullptr);
int keyBitsFor(long long numGrid);
int buildStartEnd(fdm_hash_table t, int n, const int* nDev, fdm_grid grid, void* ws, size_t wsBytes,
                  hipStream_t stream, const int* gate = nullptr);
int fillInts(int* p, size_t n, int v, hipStream_t stream, const int* gate = nullptr);

}  // namespace fdm
