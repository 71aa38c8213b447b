// Spatial-hash support: stable LSD radix sort of (key, value) pairs, device
// scans, fills, and half-open cell ranges of a sorted key list.
//
// Radix sort: 8-bit digits. Each pass is an upsweep (per-tile digit
// histogram), an exclusive scan of the digit-major histogram, and a
// downsweep that ranks keys inside the tile with wave64 ballots (the lanes
// holding the same digit form a peer mask; a key's rank is the number of
// lower lanes in its peer mask plus the running count of its digit in the
// wave) and scatters. Tiles, waves inside a tile, iterations inside a wave
// and lanes inside an iteration all follow input order, so equal keys keep
// their input order: the sort is stable, as thrust's radix sort on int keys
// is (buildHashStartEnd.cu:108-111).
#include <algorithm>
#include <mutex>

#include "fdm_device.hpp"
#include "fdm_launch.hpp"

namespace fdm {

// ------------------------------------------------------------------- fill
__global__ void kFillInt(int* p, size_t n, int v, const int* gate) {
  if (gateOff(gate)) return;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------- scan
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int waveInclusiveScan(int v) {
  const int lane = laneId();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Exclusive scan of one value per thread across the block; returns the
// block total through `total`.
__device__ __forceinline__ int blockExclusiveScan(int v, int* lds, int& total) {
  const int lane = laneId();
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int inc = waveInclusiveScan(v);
  if (lane == 63) lds[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < nwaves; ++w) {
      const int t = lds[w];
      lds[w] = run;
      run += t;
    }
    lds[nwaves] = run;
  }
  __syncthreads();
  const int r = lds[wave] + inc - v;
  total = lds[nwaves];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kScanThreads) kScanReduce(const int* in, size_t n, int* sums,
                                                            const int* gate) {
  if (gateOff(gate)) return;
  __shared__ int lds[kScanThreads / 64 + 1];
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += in[base + k];
  int total;
  blockExclusiveScan(s, lds, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Single workgroup: exclusive scan of the tile sums in place, total to *total.
__global__ void __launch_bounds__(1024) kScanSums(int* sums, int nb, int* total, const int* gate) {
  if (gateOff(gate)) return;
  __shared__ int lds[1024 / 64 + 1];
  int carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const int v = i < nb ? sums[i] : 0;
    int t;
    const int ex = blockExclusiveScan(v, lds, t);
    if (i < nb) sums[i] = carry + ex;
    carry += t;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(kScanThreads) kScanApply(const int* in, int* out, size_t n,
                                                           const int* sums, int inclusive,
                                                           const int* gate) {
  if (gateOff(gate)) return;
  __shared__ int lds[kScanThreads / 64 + 1];
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int run = blockExclusiveScan(s, lds, total) + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (inclusive) run += v[k];
    if (base + k < n) out[base + k] = run;
    if (!inclusive) run += v[k];
  }
}

size_t scanWorkspace(size_t n) { return alignUp((blocksFor((long long)n, kScanTile) + 1) * sizeof(int)); }

int scanInts(const int* in, int* out, size_t n, bool inclusive, int* total, void* ws, size_t wsBytes,
             hipStream_t stream, const int* gate) {
  if (n == 0) {
    if (total) hipLaunchKernelGGL(kFillInt, dim3(1), dim3(64), 0, stream, total, (size_t)1, 0, gate);
    return launchStatus();
  }
  if (wsBytes < scanWorkspace(n)) return FDM_EINVAL;
  const unsigned nb = blocksFor((long long)n, kScanTile);
  int* sums = (int*)ws;
  hipLaunchKernelGGL(kScanReduce, dim3(nb), dim3(kScanThreads), 0, stream, in, n, sums, gate);
  hipLaunchKernelGGL(kScanSums, dim3(1), dim3(1024), 0, stream, sums, (int)nb, total, gate);
  hipLaunchKernelGGL(kScanApply, dim3(nb), dim3(kScanThreads), 0, stream, in, out, n, sums,
                     inclusive ? 1 : 0, gate);
  return launchStatus();
}

// ------------------------------------------------------------- radix sort
constexpr int kRadixBits = 8;
constexpr int kRadixBins = 1 << kRadixBits;
constexpr int kSortWaves = 8;
constexpr int kSortThreads = kSortWaves * 64;
constexpr int kSortPerLane = 8;                         // keys per lane per wave sub-tile
constexpr int kSortWaveTile = 64 * kSortPerLane;        // 512
constexpr int kSortTile = kSortWaves * kSortWaveTile;   // 4096

// Number of elements: the host-side bound, or the device count when given.
__device__ __forceinline__ size_t effectiveCount(size_t n, const int* nDev) {
  if (!nDev) return n;
  const int d = *nDev;
  return d < 0 ? 0 : ((size_t)d < n ? (size_t)d : n);
}

__global__ void __launch_bounds__(kSortThreads) kRadixUpsweep(const unsigned* keys, size_t nCap,
                                                              const int* nDev, int shift, int* hist,
                                                              int nblocks, const int* gate) {
  if (gateOff(gate)) return;
  __shared__ int bins[kRadixBins];
  const size_t n = effectiveCount(nCap, nDev);
  for (int d = threadIdx.x; d < kRadixBins; d += kSortThreads) bins[d] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kSortTile;
  for (int k = threadIdx.x; k < kSortTile; k += kSortThreads) {
    const size_t i = base + k;
    if (i < n) atomicAdd(&bins[(keys[i] >> shift) & (kRadixBins - 1)], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadixBins; d += kSortThreads)
    hist[(size_t)d * nblocks + blockIdx.x] = bins[d];
}

// Lanes of the wave whose digit equals this lane's digit (among `valid`).
__device__ __forceinline__ unsigned long long peerMask(unsigned digit, unsigned long long valid) {
  unsigned long long peers = valid;
#pragma unroll
  for (int b = 0; b < kRadixBits; ++b) {
    const bool bit = (digit >> b) & 1u;
    const unsigned long long m = ballot64(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

__global__ void __launch_bounds__(kSortThreads) kRadixDownsweep(const unsigned* keysIn,
                                                                const int* valsIn, unsigned* keysOut,
                                                                int* valsOut, size_t nCap,
                                                                const int* nDev, int shift,
                                                                const int* offsets, int nblocks,
                                                                const int* gate) {
  if (gateOff(gate)) return;
  __shared__ int cnt[kSortWaves][kRadixBins];
  const size_t n = effectiveCount(nCap, nDev);
  const int lane = laneId();
  const int wave = threadIdx.x >> 6;
  for (int d = threadIdx.x; d < kSortWaves * kRadixBins; d += kSortThreads) (&cnt[0][0])[d] = 0;
  __syncthreads();
  const size_t wbase = (size_t)blockIdx.x * kSortTile + (size_t)wave * kSortWaveTile;
  unsigned key[kSortPerLane];
  int val[kSortPerLane];
#pragma unroll
  for (int it = 0; it < kSortPerLane; ++it) {
    const size_t i = wbase + (size_t)it * 64 + lane;
    key[it] = i < n ? keysIn[i] : 0u;
    val[it] = i < n ? valsIn[i] : 0;
  }
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // counts of each digit in this wave's sub-tile
#pragma unroll
  for (int it = 0; it < kSortPerLane; ++it) {
    const size_t i0 = wbase + (size_t)it * 64;
    const unsigned long long valid = ballot64(i0 + lane < n);
    const unsigned d = (key[it] >> shift) & (kRadixBins - 1);
    const unsigned long long peers = peerMask(d, valid);
    if ((valid >> lane) & 1ull) {
      if ((peers & lt) == 0ull) cnt[wave][d] += __popcll(peers);
    }
  }
  __syncthreads();
  // per-wave starting offsets of each digit: the tile's global offset plus
  // the counts of earlier waves
  for (int d = threadIdx.x; d < kRadixBins; d += kSortThreads) {
    int run = offsets[(size_t)d * nblocks + blockIdx.x];
    for (int w = 0; w < kSortWaves; ++w) {
      const int c = cnt[w][d];
      cnt[w][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kSortPerLane; ++it) {
    const size_t i0 = wbase + (size_t)it * 64;
    const unsigned long long valid = ballot64(i0 + lane < n);
    const unsigned d = (key[it] >> shift) & (kRadixBins - 1);
    const unsigned long long peers = peerMask(d, valid);
    int pos = 0;
    if ((valid >> lane) & 1ull) pos = cnt[wave][d] + __popcll(peers & lt);
    __builtin_amdgcn_wave_barrier();
    if (((valid >> lane) & 1ull) && (peers & lt) == 0ull) cnt[wave][d] += __popcll(peers);
    __builtin_amdgcn_wave_barrier();
    if ((valid >> lane) & 1ull) {
      keysOut[pos] = key[it];
      valsOut[pos] = val[it];
    }
  }
}

__global__ void kCopyPairs(const unsigned* ks, const int* vs, unsigned* kd, int* vd, size_t nCap,
                           const int* nDev, const int* gate) {
  if (gateOff(gate)) return;
  const size_t n = effectiveCount(nCap, nDev);
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

size_t radixWorkspace(size_t n) {
  const size_t nb = blocksFor((long long)n, kSortTile);
  return alignUp(n * sizeof(unsigned)) + alignUp(n * sizeof(int)) +
         alignUp(nb * kRadixBins * sizeof(int)) + scanWorkspace(nb * kRadixBins);
}

// Sorts n pairs, or the first *nDev of them when nDev is given (n bounds it).
int radixSort(unsigned* keys, int* vals, size_t n, const int* nDev, int keyBits, void* ws,
              size_t wsBytes, hipStream_t stream, const int* gate) {
  if (n <= 1) return FDM_OK;
  if (keyBits < 1) keyBits = 1;
  if (keyBits > 32) keyBits = 32;
  if (wsBytes < radixWorkspace(n)) return FDM_EINVAL;
  const unsigned nb = blocksFor((long long)n, kSortTile);
  Carve c(ws, wsBytes);
  unsigned* kt = c.take<unsigned>(n);
  int* vt = c.take<int>(n);
  int* hist = c.take<int>((size_t)nb * kRadixBins);
  void* sws = c.base + c.used;
  const size_t swsBytes = c.cap - c.used;
  const int passes = (keyBits + kRadixBits - 1) / kRadixBits;
  unsigned* ki = keys;
  int* vi = vals;
  unsigned* ko = kt;
  int* vo = vt;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * kRadixBits;
    hipLaunchKernelGGL(kRadixUpsweep, dim3(nb), dim3(kSortThreads), 0, stream, ki, n, nDev, shift,
                       hist, (int)nb, gate);
    int r = scanInts(hist, hist, (size_t)nb * kRadixBins, false, nullptr, sws, swsBytes, stream,
                     gate);
    if (r != FDM_OK) return r;
    hipLaunchKernelGGL(kRadixDownsweep, dim3(nb), dim3(kSortThreads), 0, stream, ki, vi, ko, vo, n,
                       nDev, shift, hist, (int)nb, gate);
    unsigned* tk = ki;
    ki = ko;
    ko = tk;
    int* tv = vi;
    vi = vo;
    vo = tv;
  }
  if (ki != keys)
    hipLaunchKernelGGL(kCopyPairs, dim3(blocksFor((long long)n, 256)), dim3(256), 0, stream, ki, vi,
                       keys, vals, n, nDev, gate);
  return launchStatus();
}

// ------------------------------------------------------------ cell ranges
// Reset the cells marked by the previous build (count < 0: all cells).
__global__ void kClearCells(const int* prevKey, const int* prevCount, int* start, int* end,
                            long long num, const int* gate) {
  if (gateOff(gate)) return;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int pc = *prevCount;
  if (pc < 0) {
    for (long long c = i0; c < num; c += stride) {
      start[c] = -1;
      end[c] = -1;
    }
    return;
  }
  for (long long i = i0; i < pc; i += stride) {
    const long long h = (long long)prevKey[i] - 1;
    if (h >= 0 && h < num) {
      start[h] = -1;
      end[h] = -1;
    }
  }
}

// findStartAndEnd (buildHashStartEnd.cu:47-67) on keys hash+1, then remember
// the keys for the next reset.
__global__ void kMarkCells(const unsigned* key, int nCap, const int* nDev, int* start, int* end,
                           long long num, int* prevKey, int* prevCount, const int* gate) {
  if (gateOff(gate)) return;
  const int n = (int)effectiveCount((size_t)nCap, nDev);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *prevCount = n;
  if (i >= n) return;
  const unsigned k = key[i];
  prevKey[i] = (int)k;
  const long long h = (long long)k - 1;
  if (h < 0 || h >= num) return;
  if (i == 0 || key[i - 1] != k) start[h] = i;
  if (i == n - 1 || key[i + 1] != k) end[h] = i + 1;
}

int keyBitsFor(long long numGrid) {
  // keys are hash + 1 with hash in [-1, numGrid); out-of-range hashes (an
  // index rounded up to the grid size) stay below 2 * numGrid + 1
  unsigned long long maxKey = (unsigned long long)numGrid * 2ull + 1ull;
  int bits = 1;
  while (bits < 32 && (maxKey >> bits) != 0ull) ++bits;
  return bits;
}

// Sort the table's (key, index) lists and mark cell ranges; with nDev the
// list length is read on the device (n is its bound).
int buildStartEnd(fdm_hash_table t, int n, const int* nDev, fdm_grid grid, void* ws,
                  size_t wsBytes, hipStream_t stream, const int* gate) {
  int r = radixSort(t.sortKey, t.sortIndex, (size_t)n, nDev, keyBitsFor(grid.num), ws, wsBytes,
                    stream, gate);
  if (r != FDM_OK) return r;
  hipLaunchKernelGGL(kClearCells, dim3(2048), dim3(256), 0, stream, t.prevKey, t.prevCount,
                     t.cellStart, t.cellEnd, grid.num, gate);
  if (n > 0)
    hipLaunchKernelGGL(kMarkCells, dim3(blocksFor(n, 256)), dim3(256), 0, stream, t.sortKey, n,
                       nDev, t.cellStart, t.cellEnd, grid.num, t.prevKey, t.prevCount, gate);
  else
    hipLaunchKernelGGL(kFillInt, dim3(1), dim3(64), 0, stream, t.prevCount, (size_t)1, 0, gate);
  return launchStatus();
}

int fillInts(int* p, size_t n, int v, hipStream_t stream, const int* gate) {
  if (n == 0) return FDM_OK;
  hipLaunchKernelGGL(kFillInt, dim3(blocksFor((long long)n, 256)), dim3(256), 0, stream, p, n, v,
                     gate);
  return launchStatus();
}

int persistentBlocks(int wavesPerBlock) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) cus = 256;
  }
  // enough resident waves to cover memory latency on every CU
  const int wavesPerCU = 32;
  return std::max(1, cus * wavesPerCU / wavesPerBlock);
}

}  // namespace fdm

using namespace fdm;

extern "C" int fdm_fillInt(int* ptr, size_t n, int value, hipStream_t stream) {
  if (n == 0) return FDM_OK;
  if (!ptr) return FDM_EINVAL;
  return fillInts(ptr, n, value, stream);
}

extern "C" size_t fdm_radixSortWorkspaceBytes(size_t n) { return radixWorkspace(n); }

extern "C" int fdm_radixSortPairs(unsigned* keys, int* values, size_t n, int keyBits,
                                  void* workspace, size_t workspaceBytes, hipStream_t stream) {
  if (n > 0 && (!keys || !values)) return FDM_EINVAL;
  return radixSort(keys, values, n, nullptr, keyBits, workspace, workspaceBytes, stream);
}

extern "C" size_t fdm_scanWorkspaceBytes(size_t n) { return scanWorkspace(n); }

extern "C" int fdm_buildPrefixSum(int* prefixSum, const int* count, size_t n, void* workspace,
                                  size_t workspaceBytes, hipStream_t stream) {
  return scanInts(count, prefixSum, n, true, nullptr, workspace, workspaceBytes, stream);
}

extern "C" int fdm_buildHashStartEnd(fdm_hash_table table, int n, fdm_grid grid, void* workspace,
                                     size_t workspaceBytes, hipStream_t stream) {
  if (n < 0 || grid.num <= 0) return FDM_EINVAL;
  return buildStartEnd(table, n, nullptr, grid, workspace, workspaceBytes, stream);
}

extern "C" int fdm_buildHashStartEndGated(fdm_hash_table table, int n, fdm_grid grid,
                                          const int* gate, void* workspace, size_t workspaceBytes,
                                          hipStream_t stream) {
  if (n < 0 || grid.num <= 0) return FDM_EINVAL;
  return buildStartEnd(table, n, nullptr, grid, workspace, workspaceBytes, stream, gate);
}

// ------------------------------------------------ fixed device scratch
// Workspace of fdm_buildHashStartEndNoAlloc: one file-scope device array,
// sized for FDM_NOALLOC_MAX_BODIES bodies (about 16.5 MiB), which every GPU
// the library runs on holds until the process ends. The sort writes every
// byte of it that it later reads (tile histograms, their scan, the
// ping-pong key and index lists), so no call depends on what an earlier
// call left there. The array's name is only a host-side stand-in: its
// address on each device comes from hipGetSymbolAddress, taken once per
// device with that device current, and only that address reaches kernels.
namespace {
constexpr size_t kNoAllocTileBlocks =
    (size_t(FDM_NOALLOC_MAX_BODIES) + size_t(kSortTile) - 1) / size_t(kSortTile);
constexpr size_t kNoAllocHist = kNoAllocTileBlocks * kRadixBins;
constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t kNoAllocBytes =
    2 * align256(size_t(FDM_NOALLOC_MAX_BODIES) * 4) + align256(kNoAllocHist * 4) +
    align256(((kNoAllocHist + kScanTile - 1) / kScanTile + 1) * 4);
constexpr int kNoAllocDevices = 64;

std::mutex g_scratchMutex;
void* g_scratchPtr[kNoAllocDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of this device's scratch array (looked up on first use).
static int scratchOfDevice(int dev, void** out) {
  if (dev < 0 || dev >= kNoAllocDevices) return FDM_ELAUNCH;
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchPtr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FDM_ELAUNCH;
    g_scratchPtr[dev] = p;
  }
  *out = g_scratchPtr[dev];
  return FDM_OK;
}

extern "C" int fdm_buildHashStartEndNoAlloc(fdm_hash_table table, int n, fdm_grid grid,
                                            hipStream_t stream) {
  if (n < 0 || grid.num <= 0) return FDM_EINVAL;
  if ((size_t)n > (size_t)FDM_NOALLOC_MAX_BODIES) return FDM_ECAPACITY;
  if (radixWorkspace((size_t)n) > kNoAllocBytes) return FDM_ECAPACITY;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return FDM_ELAUNCH;
  void* ws = nullptr;
  const int r = scratchOfDevice(dev, &ws);
  if (r != FDM_OK) return r;
  return buildStartEnd(table, n, nullptr, grid, ws, kNoAllocBytes, stream);
}

extern "C" int fdm_deviceInfo(int* numCU, int* warpSize) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return FDM_ELAUNCH;
  if (numCU) hipDeviceGetAttribute(numCU, hipDeviceAttributeMultiprocessorCount, dev);
  if (warpSize) hipDeviceGetAttribute(warpSize, hipDeviceAttributeWarpSize, dev);
  return FDM_OK;
}
