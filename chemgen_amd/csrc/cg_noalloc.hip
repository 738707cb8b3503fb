// Ragged-batch tile lists without any allocation: the tile list lives in a file-scope device
// array (`scratch`) of fixed size, filled on the device by cg_tiles_expand_k from the patch
// offsets, which travel by value in the kernel arguments.  Every call on one GPU shares that
// GPU's array: calls that overlap on one GPU (concurrent streams, or a replayed graph running
// beside another call) race on it.  The array is never cleared: each call writes every tile
// entry its launch reads.
#include <hip/hip_runtime.h>

#include <mutex>

#include "chemgen_amd.h"

extern "C" int cg_internal_fail(int code, const char* msg);

// tile starts (int64) then tile lengths (int32): 12 bytes per tile, 6 MiB in all
static __device__ __align__(256) unsigned char scratch[(int64_t)CG_NOALLOC_MAX_TILES * 12];

namespace {

constexpr int64_t kMaxTiles = CG_NOALLOC_MAX_TILES;
constexpr int kMaxPatches = CG_NOALLOC_MAX_PATCHES;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};  // device address of `scratch` on each device

struct Offsets {
  int64_t v[kMaxPatches + 1];
};

// Tile t of the batch: the patch it falls in (tiles aligned to multiples of `tile` in the
// packed arrays, a short head tile per patch, as cg_batch_tiles builds them on the host).
__global__ void __launch_bounds__(256)
cg_tiles_expand_k(Offsets off, int32_t n_patches, int32_t tile, int64_t nt, int64_t* __restrict__ start,
                  int32_t* __restrict__ len) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int64_t base = 0;  // tiles before patch b
  for (int32_t b = 0; b < n_patches; ++b) {
    const int64_t a = off.v[b], e = off.v[b + 1];
    const int64_t n = e > a ? (e - 1) / tile - a / tile + 1 : 0;
    if (t < base + n) {
      const int64_t k = t - base;
      const int64_t s = k == 0 ? a : (a / tile + k) * tile;
      int64_t end = (a / tile + k + 1) * tile;
      if (end > e) end = e;
      start[t] = s;
      len[t] = (int32_t)(end - s);
      return;
    }
    base += n;
  }
}

}  // namespace

// Device address of this device's `scratch` (looked up once per device, with it current).
static int scratch_ptr(unsigned char** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return cg_internal_fail(CG_EHIP, "cg_getrates_batched_noalloc: no current device");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return cg_internal_fail(CG_EHIP, "cg_getrates_batched_noalloc: hipGetSymbolAddress(scratch) failed");
    g_ptr[dev] = (unsigned char*)p;
  }
  *out = g_ptr[dev];
  return CG_OK;
}

// Writes the tile list of `offsets` into this device's scratch array on `stream` and returns
// the device pointers of its starts and lengths.  The caller has checked the offsets and
// n_tiles <= CG_NOALLOC_MAX_TILES, n_patches <= CG_NOALLOC_MAX_PATCHES.
extern "C" int cg_internal_noalloc_tiles(const int64_t* offsets, int32_t n_patches, int32_t tile, int64_t nt,
                                         void* stream, int64_t** start, int32_t** len) {
  if (n_patches < 0 || n_patches > kMaxPatches || nt < 0 || nt > kMaxTiles || tile <= 0)
    return cg_internal_fail(CG_ERANGE, "cg_getrates_batched_noalloc: batch beyond the scratch limits");
  unsigned char* base = nullptr;
  int rc = scratch_ptr(&base);
  if (rc != CG_OK) return rc;
  *start = (int64_t*)base;
  *len = (int32_t*)(base + kMaxTiles * sizeof(int64_t));
  if (nt == 0) return CG_OK;
  Offsets o;
  for (int32_t b = 0; b <= n_patches; ++b) o.v[b] = offsets[b];
  const unsigned blocks = (unsigned)((nt + 255) / 256);
  hipLaunchKernelGGL(cg_tiles_expand_k, dim3(blocks), dim3(256), 0, (hipStream_t)stream, o, n_patches, tile, nt,
                     *start, *len);
  if (hipGetLastError() != hipSuccess)
    return cg_internal_fail(CG_EHIP, "cg_getrates_batched_noalloc: tile expansion launch failed");
  return CG_OK;
}
