// Handle-based C ABI over generated getrates code objects (see include/chemgen_amd.h).
//
// A mechanism handle owns: the hipModule loaded from the generated code object, the
// kernels it exports (cg_getrates_kernel; cg_getrates_tiles_kernel, a persistent walk of a
// tile list; cg_getrates_tile_kernel, one workgroup per tile of that list) and the device
// copy of the gpu_coef blob.  A ragged batch (cg_batch) owns its tile list: a device copy and
// the pinned host copy it is uploaded from, both alive until cg_batch_free, so a graph
// captured over a launch of it stays valid for as long as the batch does.  The batches the
// handle makes itself (cg_batched_prepare, cg_getrates_batched) are kept by the handle (see
// there); none is freed while a launch may still read it.
#include "chemgen_amd.h"

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define CG_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(CG_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));            \
  } while (0)

struct Meta {
  int32_t abi, ns, nr, ncoef, block, ppb;  // block = threads per workgroup, ppb = points per workgroup
  uint64_t structure_hash;
  uint64_t coef_hash;  // non-zero: the code object has the blob compiled in; it must match
};

// FNV-1a over the bytes of the blob (the generator computes the same for compiled-in blobs)
uint64_t coef_fnv(const double* c, int64_t n) {
  const unsigned char* b = reinterpret_cast<const unsigned char*>(c);
  uint64_t h = 0xcbf29ce484222325ull;
  for (int64_t i = 0; i < n * (int64_t)sizeof(double); ++i) h = (h ^ b[i]) * 0x100000001b3ull;
  return h;
}

}  // namespace

struct cg_batch {
  cg_mech* mech = nullptr;
  void* dev = nullptr;       // device: tile starts (int64) then tile lengths (int32)
  void* host = nullptr;      // pinned host copy the upload reads (kept until the batch is freed)
  int64_t* tile_start_d = nullptr;
  int32_t* tile_len_d = nullptr;
  int64_t n_tiles = 0;
  int64_t n_points = 0;      // offsets[n_patches]: a launch needs ld >= n_points
  std::vector<int64_t> offsets;  // host copy (lets one-shot calls reuse an identical batch)
  hipEvent_t done = nullptr;     // one-shot batches: recorded after their launch
};

struct cg_mech {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  hipFunction_t fn_tiles = nullptr;
  hipFunction_t fn_tile1 = nullptr;  // one workgroup per tile (absent in older code objects)
  double* coef_d = nullptr;
  Meta meta{};
  int device = 0;
  int n_cu = 0;
  // cg_batched_prepare: the current prepared batch, and every earlier one (a graph captured
  // over a launch of it may still be replayed), freed by cg_mech_free
  cg_batch* prepared = nullptr;
  std::vector<cg_batch*> kept;
  // cg_getrates_batched: the batch of the last call (reused while the offsets repeat) and
  // earlier ones whose launch may not have finished (freed once their event has completed)
  cg_batch* oneshot = nullptr;
  std::vector<cg_batch*> retired;
};

extern "C" const char* cg_last_error(void) { return g_err.c_str(); }

// error reporting for the other translation units of the library
extern "C" int cg_internal_fail(int code, const char* msg) { return fail(code, "%s", msg); }

extern "C" int cg_mech_load(const char* path, const double* coef, int64_t n_coef,
                            uint64_t expect_hash, cg_mech_t* out) {
  if (!path || !coef || !out || n_coef <= 0) return fail(CG_EINVAL, "cg_mech_load: bad argument");
  *out = nullptr;
  int dev = 0, n_cu = 0;
  CG_HIP(hipGetDevice(&dev));
  CG_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  cg_mech* h = new cg_mech();
  h->device = dev;
  h->n_cu = n_cu;
  hipError_t e = hipModuleLoad(&h->mod, path);
  if (e != hipSuccess) {
    delete h;
    return fail(CG_ELOAD, "hipModuleLoad(%s): %s", path, hipGetErrorString(e));
  }
  hipDeviceptr_t meta_d = nullptr;
  size_t meta_bytes = 0;
  e = hipModuleGetGlobal(&meta_d, &meta_bytes, h->mod, "cg_meta");
  if (e != hipSuccess || meta_bytes != sizeof(Meta)) {
    (void)hipModuleUnload(h->mod);
    delete h;
    return fail(CG_ELOAD, "code object %s has no cg_meta (%s)", path, hipGetErrorString(e));
  }
  e = hipMemcpyDtoH(&h->meta, meta_d, sizeof(Meta));
  if (e == hipSuccess) e = hipModuleGetFunction(&h->fn, h->mod, "cg_getrates_kernel");
  if (e == hipSuccess) e = hipModuleGetFunction(&h->fn_tiles, h->mod, "cg_getrates_tiles_kernel");
  if (e == hipSuccess && hipModuleGetFunction(&h->fn_tile1, h->mod, "cg_getrates_tile_kernel") != hipSuccess) {
    h->fn_tile1 = nullptr;
    (void)hipGetLastError();
  }
  if (e != hipSuccess) {
    (void)hipModuleUnload(h->mod);
    delete h;
    return fail(CG_ELOAD, "code object %s: %s", path, hipGetErrorString(e));
  }
  if (h->meta.abi != 3 || h->meta.ppb <= 0 || h->meta.ncoef != n_coef ||
      (h->meta.coef_hash != 0 && h->meta.coef_hash != coef_fnv(coef, n_coef)) ||
      (expect_hash != 0 && h->meta.structure_hash != expect_hash)) {
    int code = fail(CG_ELOAD,
                    "code object %s does not match the coefficients (abi %d, ncoef %d vs %lld, "
                    "hash %016llx vs %016llx, or a compiled-in blob that differs)",
                    path, h->meta.abi, h->meta.ncoef, (long long)n_coef,
                    (unsigned long long)h->meta.structure_hash, (unsigned long long)expect_hash);
    (void)hipModuleUnload(h->mod);
    delete h;
    return code;
  }
  e = hipMalloc(&h->coef_d, sizeof(double) * n_coef);
  if (e != hipSuccess) {
    (void)hipModuleUnload(h->mod);
    delete h;
    return fail(CG_ENOMEM, "hipMalloc(coef): %s", hipGetErrorString(e));
  }
  e = hipMemcpy(h->coef_d, coef, sizeof(double) * n_coef, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(h->coef_d);
    (void)hipModuleUnload(h->mod);
    delete h;
    return fail(CG_EHIP, "hipMemcpy(coef): %s", hipGetErrorString(e));
  }
  *out = h;
  return CG_OK;
}

extern "C" int cg_mech_set_coef(cg_mech_t h, const double* coef, int64_t n_coef) {
  if (!h || !coef) return fail(CG_EINVAL, "cg_mech_set_coef: bad argument");
  if (n_coef != h->meta.ncoef) return fail(CG_EINVAL, "cg_mech_set_coef: expected %d coefficients", h->meta.ncoef);
  if (h->meta.coef_hash != 0 && h->meta.coef_hash != coef_fnv(coef, n_coef))
    return fail(CG_EINVAL, "cg_mech_set_coef: this code object has its coefficients compiled in");
  CG_HIP(hipMemcpy(h->coef_d, coef, sizeof(double) * n_coef, hipMemcpyHostToDevice));
  return CG_OK;
}

extern "C" int cg_mech_info(cg_mech_t h, int32_t* ns, int32_t* nr, int32_t* block, uint64_t* hash) {
  if (!h) return fail(CG_EINVAL, "cg_mech_info: null handle");
  if (ns) *ns = h->meta.ns;
  if (nr) *nr = h->meta.nr;
  if (block) *block = h->meta.ppb;
  if (hash) *hash = h->meta.structure_hash;
  return CG_OK;
}

static void batch_destroy(cg_batch* b);

extern "C" int cg_mech_free(cg_mech_t h) {
  if (!h) return CG_OK;
  // the caller guarantees that no launch on this handle is pending: hipFree waits for the
  // device anyway
  for (cg_batch* b : h->kept) batch_destroy(b);
  for (cg_batch* b : h->retired) batch_destroy(b);
  batch_destroy(h->prepared);
  batch_destroy(h->oneshot);
  if (h->coef_d) (void)hipFree(h->coef_d);
  if (h->mod) (void)hipModuleUnload(h->mod);
  delete h;
  return CG_OK;
}

extern "C" int cg_getrates(cg_mech_t h, const double* P, const double* T, int64_t ng, int64_t ld,
                           const double* Y, double* W, void* stream) {
  if (!h) return fail(CG_EINVAL, "cg_getrates: null handle");
  if (ng < 0 || ld < ng) return fail(CG_EINVAL, "cg_getrates: need 0 <= ng <= ld (ng=%lld ld=%lld)",
                                     (long long)ng, (long long)ld);
  if (ld >= (1LL << 32)) return fail(CG_ERANGE, "cg_getrates: ld must be below 2^32");
  if (ng == 0) return CG_OK;
  if (!P || !T || !Y || !W) return fail(CG_EINVAL, "cg_getrates: null array");
  const int64_t nb = (ng + h->meta.ppb - 1) / h->meta.ppb;
  if (nb > 0x7fffffffLL) return fail(CG_ERANGE, "cg_getrates: ng too large");
  void* args[] = {(void*)&P, (void*)&T, (void*)&Y, (void*)&W, (void*)&h->coef_d, (void*)&ng, (void*)&ld};
  CG_HIP(hipModuleLaunchKernel(h->fn, (unsigned)nb, 1, 1, (unsigned)h->meta.block, 1, 1, 0,
                               (hipStream_t)stream, args, nullptr));
  return CG_OK;
}

// ---- ragged batches ----

extern "C" int64_t cg_batch_tiles(const int64_t* offsets, int32_t n_patches, int32_t tile,
                                  int64_t* tile_start, int32_t* tile_len, int64_t cap) {
  if (!offsets || n_patches < 0 || tile <= 0 || cap < 0 || (cap > 0 && (!tile_start || !tile_len)))
    return fail(CG_EINVAL, "cg_batch_tiles: bad argument");
  if (offsets[0] != 0) return fail(CG_EINVAL, "cg_batch_tiles: offsets[0] must be 0");
  int64_t nt = 0;
  for (int32_t b = 0; b < n_patches; ++b) {
    const int64_t a = offsets[b], e = offsets[b + 1];
    if (e < a) return fail(CG_EINVAL, "cg_batch_tiles: offsets must be non-decreasing");
    // tiles aligned to multiples of `tile` in the packed arrays (a short head tile first), so
    // that every full tile reads and writes whole cache lines as the plain launch does
    for (int64_t s = a; s < e;) {
      int64_t t = (s / tile + 1) * tile;
      if (t > e) t = e;
      if (nt < cap) {
        tile_start[nt] = s;
        tile_len[nt] = (int32_t)(t - s);
      }
      ++nt;
      s = t;
    }
  }
  return nt;
}

static void batch_destroy(cg_batch* b) {
  if (!b) return;
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->dev) (void)hipFree(b->dev);
  if (b->host) (void)hipHostFree(b->host);
  delete b;
}

// Tile list of a batch in pinned host memory and on the device; the upload is enqueued on
// `stream` (sync = false) or completed before returning (sync = true).
static int batch_make(cg_mech* h, const int64_t* offsets, int32_t n_patches, hipStream_t stream,
                      bool sync, cg_batch** out) {
  *out = nullptr;
  const int64_t nt = cg_batch_tiles(offsets, n_patches, h->meta.ppb, nullptr, nullptr, 0);
  if (nt < 0) return (int)nt;
  cg_batch* b = new cg_batch();
  b->mech = h;
  b->n_tiles = nt;
  b->n_points = offsets[n_patches];
  b->offsets.assign(offsets, offsets + n_patches + 1);
  if (nt > 0) {
    const size_t bytes = (size_t)nt * (sizeof(int64_t) + sizeof(int32_t));
    if (hipHostMalloc(&b->host, bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      batch_destroy(b);
      return fail(CG_ENOMEM, "ragged batch: pinned host allocation of %zu bytes failed", bytes);
    }
    if (hipMalloc(&b->dev, bytes) != hipSuccess) {
      (void)hipGetLastError();
      batch_destroy(b);
      return fail(CG_ENOMEM, "ragged batch: device allocation of %zu bytes failed", bytes);
    }
    int64_t* hs = (int64_t*)b->host;
    int32_t* hl = (int32_t*)(hs + nt);
    (void)cg_batch_tiles(offsets, n_patches, h->meta.ppb, hs, hl, nt);
    b->tile_start_d = (int64_t*)b->dev;
    b->tile_len_d = (int32_t*)(b->tile_start_d + nt);
    hipError_t e = sync ? hipMemcpy(b->dev, b->host, bytes, hipMemcpyHostToDevice)
                        : hipMemcpyAsync(b->dev, b->host, bytes, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) {
      batch_destroy(b);
      return fail(CG_EHIP, "ragged batch: tile list upload failed: %s", hipGetErrorString(e));
    }
  }
  *out = b;
  return CG_OK;
}

// One launch of the mechanism's tile kernel over nt tiles of a device tile list.
static int tiles_launch(cg_mech* h, const int64_t* tile_start, const int32_t* tile_len, int64_t nt,
                        int64_t n_points, const double* P, const double* T, const double* Y, double* W,
                        int64_t ld, int32_t grid, hipStream_t stream, const char* who) {
  if (nt == 0) return CG_OK;
  if (!P || !T || !Y || !W) return fail(CG_EINVAL, "%s: null array", who);
  if (ld < n_points)
    return fail(CG_EINVAL, "%s: ld = %lld < %lld points of the batch", who, (long long)ld,
                (long long)n_points);
  if (ld >= (1LL << 32)) return fail(CG_ERANGE, "%s: ld must be below 2^32", who);
  // grid > 0: that many persistent workgroups; < 0: one workgroup per tile; 0: one per tile
  // for tiles of at most 16 points (the table kernels of large mechanisms: one tile is ~1e5
  // cycles of work, and a persistent walk over ragged tiles measured 1.2-1.3x the time), else
  // 8 per CU
  int64_t g = grid > 0 ? grid : ((grid < 0 || h->meta.ppb <= 16) ? nt : (int64_t)h->n_cu * 8);
  if (g > nt) g = nt;
  if (g > 0x7fffffffLL) g = 0x7fffffffLL;
  void* args[] = {(void*)&P, (void*)&T, (void*)&Y, (void*)&W, (void*)&h->coef_d,
                  (void*)&tile_start, (void*)&tile_len, (void*)&nt, (void*)&ld};
  // a grid that covers every tile runs the loop-free kernel (same tiles, one each)
  hipFunction_t fn = (g == nt && h->fn_tile1) ? h->fn_tile1 : h->fn_tiles;
  CG_HIP(hipModuleLaunchKernel(fn, (unsigned)g, 1, 1, (unsigned)h->meta.block, 1, 1, 0, stream, args,
                               nullptr));
  return CG_OK;
}

static int batch_launch(cg_batch* b, const double* P, const double* T, const double* Y, double* W,
                        int64_t ld, int32_t grid, hipStream_t stream, const char* who) {
  return tiles_launch(b->mech, b->tile_start_d, b->tile_len_d, b->n_tiles, b->n_points, P, T, Y, W, ld, grid,
                      stream, who);
}

extern "C" int cg_batch_create(cg_mech_t h, const int64_t* offsets, int32_t n_patches, void* stream,
                               cg_batch_t* out) {
  if (!out) return fail(CG_EINVAL, "cg_batch_create: null output");
  *out = nullptr;
  if (!h || !offsets || n_patches < 0) return fail(CG_EINVAL, "cg_batch_create: bad argument");
  return batch_make(h, offsets, n_patches, (hipStream_t)stream, false, out);
}

extern "C" int cg_batch_info(cg_batch_t b, int64_t* n_tiles, int64_t* n_points) {
  if (!b) return fail(CG_EINVAL, "cg_batch_info: null batch");
  if (n_tiles) *n_tiles = b->n_tiles;
  if (n_points) *n_points = b->n_points;
  return CG_OK;
}

extern "C" int cg_getrates_batch(cg_batch_t b, const double* P, const double* T, const double* Y,
                                 double* W, int64_t ld, int32_t grid, void* stream) {
  if (!b) return fail(CG_EINVAL, "cg_getrates_batch: null batch");
  return batch_launch(b, P, T, Y, W, ld, grid, (hipStream_t)stream, "cg_getrates_batch");
}

extern "C" int cg_batch_free(cg_batch_t b) {
  batch_destroy(b);
  return CG_OK;
}

extern "C" int cg_batched_prepare(cg_mech_t h, const int64_t* offsets, int32_t n_patches) {
  if (!h || !offsets || n_patches < 0) return fail(CG_EINVAL, "cg_batched_prepare: bad argument");
  // a new batch every time; the previous one stays alive (a graph captured over a launch of
  // it may still be replayed) until cg_mech_free.  The upload completes before returning, so
  // a following launch on any stream sees it.
  cg_batch* b = nullptr;
  int rc = batch_make(h, offsets, n_patches, nullptr, true, &b);
  if (rc != CG_OK) return rc;
  if (h->prepared) h->kept.push_back(h->prepared);
  h->prepared = b;
  return CG_OK;
}

extern "C" int cg_getrates_batched_prepared(cg_mech_t h, const double* P, const double* T,
                                            const double* Y, double* W, int64_t ld, int32_t grid,
                                            void* stream) {
  if (!h) return fail(CG_EINVAL, "cg_getrates_batched_prepared: null handle");
  if (!h->prepared) return fail(CG_EINVAL, "cg_getrates_batched_prepared: no batch prepared");
  return batch_launch(h->prepared, P, T, Y, W, ld, grid, (hipStream_t)stream,
                      "cg_getrates_batched_prepared");
}

extern "C" int cg_getrates_batched(cg_mech_t h, const double* P, const double* T, const double* Y,
                                   double* W, int64_t ld, const int64_t* offsets, int32_t n_patches,
                                   int32_t grid, void* stream) {
  if (!h || !offsets || n_patches < 0) return fail(CG_EINVAL, "cg_getrates_batched: bad argument");
  if (offsets[n_patches] > ld) return fail(CG_EINVAL, "cg_getrates_batched: offsets exceed ld");
  hipStream_t st = (hipStream_t)stream;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  CG_HIP(hipStreamIsCapturing(st, &cs));
  if (cs != hipStreamCaptureStatusNone)
    return fail(CG_EINVAL, "cg_getrates_batched: stream is capturing; capture cg_getrates_batch over a "
                           "cg_batch_create'd batch instead");
  // free the retired one-shot batches whose launches have finished
  for (size_t i = 0; i < h->retired.size();) {
    cg_batch* r = h->retired[i];
    hipError_t q = hipEventQuery(r->done);
    if (q == hipSuccess) {
      batch_destroy(r);
      h->retired[i] = h->retired.back();
      h->retired.pop_back();
    } else {
      if (q != hipErrorNotReady) (void)hipGetLastError();
      ++i;
    }
  }
  cg_batch* b = h->oneshot;
  const bool same = b && (int64_t)b->offsets.size() == (int64_t)n_patches + 1 &&
                    std::memcmp(b->offsets.data(), offsets, sizeof(int64_t) * (n_patches + 1)) == 0;
  if (!same) {
    cg_batch* nb = nullptr;
    int rc = batch_make(h, offsets, n_patches, st, false, &nb);
    if (rc != CG_OK) return rc;
    if (hipEventCreateWithFlags(&nb->done, hipEventDisableTiming) != hipSuccess) {
      batch_destroy(nb);
      return fail(CG_EHIP, "cg_getrates_batched: hipEventCreate failed");
    }
    if (b) {
      if (b->done && hipEventQuery(b->done) == hipSuccess) {
        batch_destroy(b);
      } else {
        (void)hipGetLastError();
        h->retired.push_back(b);
      }
    }
    h->oneshot = b = nb;
  }
  int rc = batch_launch(b, P, T, Y, W, ld, grid, st, "cg_getrates_batched");
  if (rc != CG_OK) return rc;
  CG_HIP(hipEventRecord(b->done, st));
  return CG_OK;
}

extern "C" int cg_getrates_tiles(cg_mech_t h, const double* P, const double* T, const double* Y, double* W,
                                 int64_t ld, const int64_t* tile_start, const int32_t* tile_len, int64_t n_tiles,
                                 int64_t n_points, int32_t grid, void* stream) {
  if (!h || n_tiles < 0 || n_points < 0 || (n_tiles > 0 && (!tile_start || !tile_len)))
    return fail(CG_EINVAL, "cg_getrates_tiles: bad argument");
  return tiles_launch(h, tile_start, tile_len, n_tiles, n_points, P, T, Y, W, ld, grid, (hipStream_t)stream,
                      "cg_getrates_tiles");
}

extern "C" int cg_internal_noalloc_tiles(const int64_t* offsets, int32_t n_patches, int32_t tile, int64_t nt,
                                         void* stream, int64_t** start, int32_t** len);

extern "C" int cg_getrates_batched_noalloc(cg_mech_t h, const double* P, const double* T, const double* Y,
                                           double* W, int64_t ld, const int64_t* offsets, int32_t n_patches,
                                           int32_t grid, void* stream) {
  if (!h || !offsets || n_patches < 0) return fail(CG_EINVAL, "cg_getrates_batched_noalloc: bad argument");
  if (n_patches > CG_NOALLOC_MAX_PATCHES)
    return fail(CG_ERANGE, "cg_getrates_batched_noalloc: %d patches, at most %d", n_patches,
                CG_NOALLOC_MAX_PATCHES);
  // offsets checked and tiles counted on the host (count only: nothing is written or allocated)
  const int64_t nt = cg_batch_tiles(offsets, n_patches, h->meta.ppb, nullptr, nullptr, 0);
  if (nt < 0) return (int)nt;
  if (nt > CG_NOALLOC_MAX_TILES)
    return fail(CG_ERANGE, "cg_getrates_batched_noalloc: %lld tiles, at most %d (%d points per tile)",
                (long long)nt, CG_NOALLOC_MAX_TILES, h->meta.ppb);
  if (offsets[n_patches] > ld) return fail(CG_EINVAL, "cg_getrates_batched_noalloc: offsets exceed ld");
  if (nt == 0) return CG_OK;
  if (!P || !T || !Y || !W) return fail(CG_EINVAL, "cg_getrates_batched_noalloc: null array");
  int dev = -1;
  CG_HIP(hipGetDevice(&dev));
  if (dev != h->device)
    return fail(CG_EINVAL, "cg_getrates_batched_noalloc: current device %d, mechanism loaded on %d", dev,
                h->device);
  int64_t* ts = nullptr;
  int32_t* tl = nullptr;
  int rc = cg_internal_noalloc_tiles(offsets, n_patches, h->meta.ppb, nt, stream, &ts, &tl);
  if (rc != CG_OK) return rc;
  return tiles_launch(h, ts, tl, nt, offsets[n_patches], P, T, Y, W, ld, grid, (hipStream_t)stream,
                      "cg_getrates_batched_noalloc");
}
