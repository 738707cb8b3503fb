// Handle-based C ABI over generated getrates code objects (see include/chemgen_amd.h).
//
// A mechanism handle owns: the hipModule loaded from the generated code object, the
// kernels it exports (cg_getrates_kernel; cg_getrates_tiles_kernel, a persistent walk of a
// tile list; cg_getrates_tile_kernel, one workgroup per tile of that list), the device
// copy of the gpu_coef blob, and the tile list of the last prepared ragged batch.
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

struct cg_mech {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  hipFunction_t fn_tiles = nullptr;
  hipFunction_t fn_tile1 = nullptr;  // one workgroup per tile (absent in older code objects)
  double* coef_d = nullptr;
  Meta meta{};
  int device = 0;
  int n_cu = 0;
  // ragged-batch tile list (device), capacity in tiles
  int64_t* tile_start_d = nullptr;
  int32_t* tile_len_d = nullptr;
  int64_t tile_cap = 0;
  int64_t n_tiles = 0;
  int64_t n_points = 0;  // offsets[n_patches] of the prepared batch: the launch needs ld >= it
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

extern "C" int cg_mech_free(cg_mech_t h) {
  if (!h) return CG_OK;
  if (h->tile_start_d) (void)hipFree(h->tile_start_d);
  if (h->tile_len_d) (void)hipFree(h->tile_len_d);
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
  if (ng == 0) return CG_OK;
  if (!P || !T || !Y || !W) return fail(CG_EINVAL, "cg_getrates: null array");
  const int64_t nb = (ng + h->meta.ppb - 1) / h->meta.ppb;
  if (nb > 0x7fffffffLL) return fail(CG_ERANGE, "cg_getrates: ng too large");
  void* args[] = {(void*)&P, (void*)&T, (void*)&Y, (void*)&W, (void*)&h->coef_d, (void*)&ng, (void*)&ld};
  CG_HIP(hipModuleLaunchKernel(h->fn, (unsigned)nb, 1, 1, (unsigned)h->meta.block, 1, 1, 0,
                               (hipStream_t)stream, args, nullptr));
  return CG_OK;
}

extern "C" int cg_batched_prepare(cg_mech_t h, const int64_t* offsets, int32_t n_patches) {
  if (!h || !offsets || n_patches < 0) return fail(CG_EINVAL, "cg_batched_prepare: bad argument");
  if (offsets[0] != 0) return fail(CG_EINVAL, "cg_batched_prepare: offsets[0] must be 0");
  // the device tile list may still be read by an earlier batched launch
  if (h->n_tiles > 0) CG_HIP(hipDeviceSynchronize());
  const int64_t B = h->meta.ppb;
  std::vector<int64_t> start;
  std::vector<int32_t> len;
  for (int32_t b = 0; b < n_patches; ++b) {
    const int64_t a = offsets[b], e = offsets[b + 1];
    if (e < a) return fail(CG_EINVAL, "cg_batched_prepare: offsets must be non-decreasing");
    // tiles aligned to multiples of B in the packed arrays (a short head tile first), so that
    // every full tile reads and writes whole cache lines as the plain launch does
    int64_t s = a;
    while (s < e) {
      int64_t t = (s / B + 1) * B;
      if (t > e) t = e;
      start.push_back(s);
      len.push_back((int32_t)(t - s));
      s = t;
    }
  }
  const int64_t nt = (int64_t)start.size();
  if (nt > h->tile_cap) {
    if (h->tile_start_d) (void)hipFree(h->tile_start_d);
    if (h->tile_len_d) (void)hipFree(h->tile_len_d);
    h->tile_start_d = nullptr;
    h->tile_len_d = nullptr;
    h->tile_cap = 0;
    int64_t cap = nt + nt / 4 + 64;
    if (hipMalloc(&h->tile_start_d, sizeof(int64_t) * cap) != hipSuccess ||
        hipMalloc(&h->tile_len_d, sizeof(int32_t) * cap) != hipSuccess)
      return fail(CG_ENOMEM, "cg_batched_prepare: tile list allocation failed");
    h->tile_cap = cap;
  }
  if (nt > 0) {
    CG_HIP(hipMemcpy(h->tile_start_d, start.data(), sizeof(int64_t) * nt, hipMemcpyHostToDevice));
    CG_HIP(hipMemcpy(h->tile_len_d, len.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice));
  }
  h->n_tiles = nt;
  h->n_points = offsets[n_patches];
  return CG_OK;
}

extern "C" int cg_getrates_batched_prepared(cg_mech_t h, const double* P, const double* T,
                                            const double* Y, double* W, int64_t ld, int32_t grid,
                                            void* stream) {
  if (!h) return fail(CG_EINVAL, "cg_getrates_batched_prepared: null handle");
  if (h->n_tiles == 0) return CG_OK;
  if (!P || !T || !Y || !W || ld < h->n_points)
    return fail(CG_EINVAL, "cg_getrates_batched_prepared: bad array or ld < %lld points",
                (long long)h->n_points);
  // grid > 0: that many persistent workgroups; 0: 8 per CU; < 0: one workgroup per tile
  int64_t g = grid > 0 ? grid : (grid < 0 ? h->n_tiles : (int64_t)h->n_cu * 8);
  if (g > h->n_tiles) g = h->n_tiles;
  if (g > 0x7fffffffLL) g = 0x7fffffffLL;
  int64_t nt = h->n_tiles;
  void* args[] = {(void*)&P, (void*)&T, (void*)&Y, (void*)&W, (void*)&h->coef_d,
                  (void*)&h->tile_start_d, (void*)&h->tile_len_d, (void*)&nt, (void*)&ld};
  // a grid that covers every tile runs the loop-free kernel (same tiles, one each)
  hipFunction_t fn = (g == nt && h->fn_tile1) ? h->fn_tile1 : h->fn_tiles;
  CG_HIP(hipModuleLaunchKernel(fn, (unsigned)g, 1, 1, (unsigned)h->meta.block, 1, 1, 0,
                               (hipStream_t)stream, args, nullptr));
  return CG_OK;
}

extern "C" int cg_getrates_batched(cg_mech_t h, const double* P, const double* T, const double* Y,
                                   double* W, int64_t ld, const int64_t* offsets, int32_t n_patches,
                                   int32_t grid, void* stream) {
  if (!h || !offsets || n_patches < 0) return fail(CG_EINVAL, "cg_getrates_batched: bad argument");
  if (offsets[n_patches] > ld) return fail(CG_EINVAL, "cg_getrates_batched: offsets exceed ld");
  int rc = cg_batched_prepare(h, offsets, n_patches);
  if (rc != CG_OK) return rc;
  return cg_getrates_batched_prepared(h, P, T, Y, W, ld, grid, stream);
}
