/* chemgen_amd: C ABI of the MI355X (gfx950) getrates library.
 *
 * The reference exposes its HIP getrates through
 *   extern "C" void getrates_hip_wrapper_vN(double* P, double* Temp, int ng, double* Ysp,
 *                                           double* wdot_rr, double* log_eg, double* log_c);
 *   (chemgen/templates/reactions_hip/hip_v1_h.j2:7-13, hip_v2_h.j2:7-13, hip_v3_h.j2:7-13,
 *    hip_v4_h.j2:26-32)
 * plus a per-mechanism constant upload
 *   extern "C" void copyConstantsToDevice(...)      (chemgen/constants.py:119-240)
 * and calls them from Fortran through chemgen_m::init_chemgen / getrates_gpu
 * (chemgen/templates/chemgen_mod.j2:26-166).
 *
 * This header replaces that pair with a handle-based API: one mechanism per handle
 * (code object + gpu_coef blob) instead of one per process, explicit HIP streams,
 * error codes instead of void, tails of any length, and a leading dimension for the
 * species-major (SoA) arrays.  Every pointer passed to a compute call is a DEVICE
 * pointer; Y and wdot are laid out as Y[k*ld + i] (species k, grid point i), which is
 * the reference's Fortran Y(ng,Ns) column-major layout (device_functions_h.j2:712)
 * with ld = ng.  Calls enqueue work on `stream` (a hipStream_t; NULL = default stream)
 * and do not synchronise; the caller synchronises the stream before reading wdot.
 * Calls on one handle from several host threads must be serialised by the caller.
 *
 * Return value of every function: 0 on success, a negative CG_E* code otherwise;
 * cg_last_error() gives the message of the calling thread's last failure.
 */
#ifndef CHEMGEN_AMD_H
#define CHEMGEN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CG_OK 0
#define CG_EINVAL (-1)   /* bad argument (null pointer, negative size, ld < ng, ...) */
#define CG_EHIP (-2)     /* a HIP runtime call failed */
#define CG_ELOAD (-3)    /* code object missing / malformed / wrong mechanism */
#define CG_ENOMEM (-4)   /* device allocation failed */
#define CG_ERANGE (-5)   /* size beyond what the entry point supports */

typedef struct cg_mech* cg_mech_t;

/* Message of this thread's last failure ("" if none). */
const char* cg_last_error(void);

/* Load a generated getrates code object and its gpu_coef blob on the current device.
 * Replaces copyConstantsToDevice (chemgen/constants.py:119-240): the coefficients are
 * copied into device memory owned by the handle.  `expect_hash` must equal the
 * structure hash compiled into the code object (0 skips the check). */
int cg_mech_load(const char* code_object_path, const double* coef_host, int64_t n_coef,
                 uint64_t expect_hash, cg_mech_t* out);

/* Replace the coefficient values of a loaded mechanism (same structure, e.g. perturbed
 * Arrhenius or thermo data).  n_coef must match. */
int cg_mech_set_coef(cg_mech_t h, const double* coef_host, int64_t n_coef);

/* Sizes compiled into the code object. */
int cg_mech_info(cg_mech_t h, int32_t* n_species, int32_t* n_reactions, int32_t* block,
                 uint64_t* structure_hash);

int cg_mech_free(cg_mech_t h);

/* wdot = getrates(P, T, Y) for ng grid points: P [dyn/cm^2], T [K], Y mass fractions,
 * wdot [mol/cm^3/s].  Replaces getrates_hip_wrapper_vN (hip_v1_h.j2:7-13): no scratch
 * arguments (log_eg/log_c are not needed), no dropped tail when ng is not a multiple
 * of the block, asynchronous on `stream`.  ld >= ng. */
int cg_getrates(cg_mech_t h, const double* P, const double* T, int64_t ng, int64_t ld,
                const double* Y, double* wdot, void* stream);

/* ---- ragged batches (no counterpart in the reference, which launches one kernel per array,
 * hip_v1_cpp.j2:339-372) ----
 * Patch b covers points [offsets[b], offsets[b+1]) of the packed arrays (offsets is a HOST
 * array of n_patches+1 entries, offsets[0] = 0, non-decreasing).  The host splits the patches
 * into tiles of at most the kernel's points per workgroup (never straddling a patch, aligned
 * to tile boundaries of the packed arrays) and one launch walks the tile list.  `grid` =
 * number of persistent workgroups walking the list (0 = the default: 8 per CU, or one
 * workgroup per tile for kernels whose tiles hold at most 16 points -- the table kernels of
 * large mechanisms, where one tile is tens of microseconds of work and the hardware's
 * dispatch balances ragged tiles better than a persistent walk); negative, or at least the
 * number of tiles, = one workgroup per tile (a loop-free kernel over the same list). */

typedef struct cg_batch* cg_batch_t;

/* A batch object: builds the tile list of `offsets` for mechanism h and enqueues its upload
 * on `stream` (launches on other streams must be ordered after it by the caller).  The batch
 * owns its tile list (device memory and the pinned host copy the upload reads) until
 * cg_batch_free, so a graph captured over cg_getrates_batch launches of it may be replayed
 * for as long as the batch lives.  Create it outside stream capture. */
int cg_batch_create(cg_mech_t h, const int64_t* offsets, int32_t n_patches, void* stream,
                    cg_batch_t* out);
int cg_batch_info(cg_batch_t b, int64_t* n_tiles, int64_t* n_points);
/* One launch over the batch's tiles (capturable into a hipGraph); ld >= its point count. */
int cg_getrates_batch(cg_batch_t b, const double* P, const double* T, const double* Y,
                      double* wdot, int64_t ld, int32_t grid, void* stream);
/* Frees the tile list; no launch or graph that uses the batch may still be pending. */
int cg_batch_free(cg_batch_t b);

/* The tile list itself, on the host (no GPU call): writes up to `cap` tiles of at most `tile`
 * points and returns the number of tiles (a negative CG_E* code on bad input). */
int64_t cg_batch_tiles(const int64_t* offsets, int32_t n_patches, int32_t tile,
                       int64_t* tile_start, int32_t* tile_len, int64_t cap);

/* One-shot form: builds (or reuses, when the offsets equal the previous call's) the handle's
 * batch and launches it.  Earlier one-shot tile lists are freed only once the launch that read
 * them has completed (an event per list), never by waiting.  Refused on a capturing stream:
 * capture cg_getrates_batch over a batch object instead. */
int cg_getrates_batched(cg_mech_t h, const double* P, const double* T, const double* Y,
                        double* wdot, int64_t ld, const int64_t* offsets, int32_t n_patches,
                        int32_t grid, void* stream);

/* Handle-held form: cg_batched_prepare makes a new batch (upload complete on return) that the
 * following cg_getrates_batched_prepared calls launch (capturable).  Every earlier prepared
 * batch stays allocated until cg_mech_free, so graphs captured over earlier prepares remain
 * valid after a re-prepare (cost: 12 bytes per tile per prepare). */
int cg_batched_prepare(cg_mech_t h, const int64_t* offsets, int32_t n_patches);
int cg_getrates_batched_prepared(cg_mech_t h, const double* P, const double* T,
                                 const double* Y, double* wdot, int64_t ld, int32_t grid,
                                 void* stream);

/* One launch of the mechanism's tile kernel over a DEVICE tile list the caller owns:
 * tile t covers points [tile_start[t], tile_start[t] + tile_len[t]) (length at most the
 * kernel's points per workgroup, inside [0, n_points) <= ld); tiles of length 0 are skipped, so
 * a list whose lengths are set on the device (cg_count_tiles) makes a launch whose point count
 * is a device value.  No counterpart in the reference. */
int cg_getrates_tiles(cg_mech_t h, const double* P, const double* T, const double* Y, double* wdot,
                      int64_t ld, const int64_t* tile_start, const int32_t* tile_len, int64_t n_tiles,
                      int64_t n_points, int32_t grid, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * cg_getrates_batched without any allocation.  The tile list is built on the device into a
 * fixed array of the library (CG_NOALLOC_MAX_TILES tiles, 6 MiB per GPU, held until the
 * process ends) from the offsets, which travel by value in a kernel's arguments; at most
 * CG_NOALLOC_MAX_PATCHES patches and CG_NOALLOC_MAX_TILES tiles (tiles of the kernel's
 * points per workgroup), larger batches are refused with CG_ERANGE.  Results are identical
 * to cg_getrates_batched's.  Capturable.  Calls that overlap on one GPU (other streams, a
 * graph replayed concurrently) share the array and race: serialise them. */
#define CG_NOALLOC_MAX_PATCHES 256
#define CG_NOALLOC_MAX_TILES (1 << 19)
int cg_getrates_batched_noalloc(cg_mech_t h, const double* P, const double* T, const double* Y,
                                double* wdot, int64_t ld, const int64_t* offsets, int32_t n_patches,
                                int32_t grid, void* stream);

/* ---- learned surrogate (no counterpart in the reference: chemgen/chemistry_nn.py:15-85 is a
 * non-functional prototype).  BF16 values are passed as void* (16-bit, device memory). ---- */

/* C[M,N] = A B (+ bias) on the matrix cores, FP32 accumulation.  A(i,k) = A[i*sAi + k*sAk],
 * B(k,j) = B[k*sBk + j*sBj] (BF16).  Outputs: C32 (FP32, optional) gets acc + bias; Cb (BF16,
 * optional) gets act ? SiLU(acc + bias) : acc + bias.  ksplit > 1 splits K over workgroups
 * and needs accumulate = 1 (C32 += partial, atomically). */
int cg_mlp_gemm(const void* A, int64_t sAi, int64_t sAk, const void* B, int64_t sBk, int64_t sBj,
                int32_t M, int32_t N, int32_t K, int32_t ksplit, const float* bias, int32_t act,
                float* C32, int64_t ldc, void* Cb, int64_t ldcb, int32_t accumulate, void* stream);
/* X[m, :] = ((T-f0)*f1, (ln P - f2)*f3, (ln(Y_k + 1e-12) - f4)*f5 ..., (1000/T - f6)*f7, 0...)
 * in BF16 rows of width kpad (fscale holds f0..f7) */
int cg_mlp_features(const double* P, const double* T, const double* Y, int64_t ld, int64_t ng,
                    int32_t ns, int32_t kpad, const float* fscale, void* X, void* stream);
/* Z[m, k] = asinh(wdot[k, m] / s[k]) */
int cg_mlp_targets(const double* W, int64_t ld, int64_t ng, int32_t ns, const double* s, float* Z,
                   void* stream);
/* G = 2 scale (Zh - Z) (BF16), *loss += scale sum (Zh - Z)^2 */
int cg_mlp_mse_grad(const float* Zh, const float* Z, int64_t n, float scale, void* G, float* loss,
                    void* stream);
/* dZ = dH * SiLU'(Zp) (BF16) */
int cg_mlp_silu_back(const float* dH, const float* Zp, int64_t n, void* dZ, void* stream);
/* db[n] += sum_m G[m, n] (G BF16 [M, N]) */
int cg_mlp_colsum(const void* G, int64_t M, int32_t N, float* db, void* stream);
/* Adam step on FP32 master weights w (bias-corrected, step >= 1); w16 (BF16) refreshed */
int cg_mlp_adam(float* w, const float* g, float* m1, float* m2, int64_t n, float lr, float b1,
                float b2, float eps, int32_t step, void* w16, void* stream);
int cg_mlp_to_bf16(const float* x, int64_t n, void* y, void* stream);
/* wdot[k, m] = s[k] sinh(Zh[m, k]) (FP64) and the element-conservation residual
 * resid[m] = max_e |sum_k a[e,k] wdot_k| / sum_k |a[e,k] wdot_k|; flag[m] = resid > tol */
int cg_mlp_decode(const float* Zh, int64_t ng, int32_t ns, const double* s, const float* aek,
                  int32_t ne, double tol, double* W, int64_t ld, int32_t* flag, double* resid,
                  void* stream);
/* idx[(*count)++] = m for every flagged m that is not audited (m % audit_every != 0; every
 * flagged m when audit_every <= 0), in unspecified order; nothing while *skip != 0 (skip may be
 * NULL).  One atomic add on *count per 4096 points. */
int cg_mlp_compact(const int32_t* flag, int64_t ng, int32_t audit_every, int32_t* count, int64_t* idx,
                   const int32_t* skip, void* stream);
/* The surrogate audit's verdict on the device: *n_bad = number of the n audited points with
 * max_k |Ws - Wr| / (|Wr| + s_k) > err_tol or not finite (an integer count, so deterministic),
 * *reject = (*n_bad > max_bad * n), and *count_clear = 0 when count_clear is not NULL (the
 * compaction's counter).  acc: two uint32 of device workspace, zero before the first call; each
 * call leaves them zero. */
int cg_mlp_audit(const double* Ws, int64_t lds, const double* Wr, int64_t ldr, int64_t n, int32_t ns,
                 const double* s, double err_tol, double max_bad, int32_t* reject, int32_t* n_bad,
                 uint32_t* acc, int32_t* count_clear, void* stream);
/* The fallback plan after the compaction, and the tile lengths of its two FP64 launches over
 * [0, ng) in tiles of `tile` points (tile t starts at t * tile): if *reject or
 * *count > dense_frac * ng, *full = ng and *pcount = 0 (one launch over the whole batch), else
 * *full = 0 and *pcount = *count (the flagged points recomputed packed);
 * len_packed[t] = clamp(*pcount - t tile, 0, tile), len_full[t] = clamp(*full - t tile, 0, tile). */
int cg_mlp_fallback_plan(const int32_t* count, const int32_t* reject, int64_t ng, double dense_frac,
                         int32_t tile, int32_t* full, int32_t* pcount, int32_t* len_packed,
                         int32_t* len_full, void* stream);
/* Device tile list for the first *count points of [0, cap): tile_start[t] = t * tile,
 * tile_len[t] = clamp(*count - t * tile, 0, tile), t < ceil(cap / tile). */
int cg_count_tiles(const int32_t* count, int64_t cap, int32_t tile, int64_t* tile_start,
                   int32_t* tile_len, void* stream);
/* pack / unpack the listed points (first *count of idx; cap = launch bound) */
int cg_mlp_gather(const double* P, const double* T, const double* Y, int64_t ld, const int64_t* idx,
                  const int32_t* count, int64_t cap, int32_t ns, double* Po, double* To, double* Yo,
                  int64_t ldo, void* stream);
int cg_mlp_scatter(const double* Wi, int64_t ldi, const int64_t* idx, const int32_t* count, int64_t cap,
                   int32_t ns, double* W, int64_t ld, void* stream);
/* Weights W [N, K] (BF16, row-major) into the operand order of cg_mlp_infer: NT x KC fragments
 * of 64 lanes x 8 values (NT >= N/16, KC >= K/32; zero padding). */
int cg_mlp_pack(const void* W, int32_t N, int32_t K, int32_t NT, int32_t KC, void* Wf, void* stream);
/* Whole-network inference in one launch: features of (P, T, Y[k*ld + m]) -> n_hidden SiLU layers
 * of width hidden (256 or 512) -> ns outputs z -> wdot W[k*ldw + m] = s[k] sinh(z_k) (FP64) ->
 * residual/flag as cg_mlp_decode.  Wf: the layers packed by cg_mlp_pack back to back
 * (NT = hidden/16 for hidden layers, ceil(ns/16) for the output; KC = ceil(kin/32) for the
 * first layer, hidden/32 after); bias: FP32, hidden values per hidden layer then ns.
 * Requires ns + 2 <= kin <= hidden, ns <= hidden, ne <= 8.  skip (device, may be NULL): when *skip != 0
 * at launch, every workgroup returns at once (a device-side verdict cancels the launch). */
int cg_mlp_infer(const double* P, const double* T, const double* Y, int64_t ld, int64_t ng, int32_t ns,
                 const float* fscale, const void* Wf, const float* bias, int32_t hidden, int32_t kin,
                 int32_t n_hidden, const double* s, const float* aek, int32_t ne, double tol, double* W,
                 int64_t ldw, int32_t* flag, double* resid, const int32_t* skip, void* stream);

/* ---- measurement (no counterpart in the reference) ---- */

/* FP64 FMA throughput probe for the roofline's measured ceiling: `blocks` workgroups of 256
 * lanes, each lane eight independent FMA chains of `iters` steps (2 * 8 * iters * 256 * blocks
 * flops), one result per lane into `out` (device, blocks * 256 doubles). */
int cg_probe_fp64_fma(double* out, int32_t blocks, int32_t iters, void* stream);

/* ---- getrates in the GEMM formulation (the reference's parallel level 4) ----
 * Replaces getrates_hip_wrapper_v4 + allocate_intermediate_DeviceMemory
 * (chemgen/templates/reactions_hip/hip_v4_h.j2:26-32, hip_v4_cpp.j2:422-597,
 * chemgen/constants.py:435-470): the stoichiometric sums as FP64 DGEMMs (rocBLAS) with dense
 * coefficient matrices, element-wise kernels around them.  A plan owns device copies of the
 * matrices, the blob, a rocBLAS handle and a workspace for `slab` grid points
 * ((3 Ns + 4 Nr + Ntb + 4) slab doubles); cg_getrates_gemm works through the points slab by
 * slab on `stream`.  Calls on one plan must not overlap (they share its workspace).
 *   coef/n_coef/off_smh: the gpu_coef blob and the offset of its NASA-7 block; tmid[Ns]: switch
 *   temperatures; nu_r, nu_p [Nr][Ns]: reactant and product coefficients; eff [Ntb][Ns]:
 *   collision efficiencies (1 for unlisted species) of the reactions with a third body;
 *   rec [Nr][8]: kind (0 elementary, 1 third-body, 2 Lindemann, 3 Troe, 4 Troe with T2, 5 SRI,
 *   6 SRI with T^e, 7 PLOG), blob offset of the rate, row of eff (-1: none), reversible,
 *   negative prefactor, first and count of the reaction's pressure points in plog, 0;
 *   dnu[Nr]: sum nu_p - sum nu_r; plog[n_plog][2]: blob offset of a pressure point's record
 *   and its number of Arrhenius terms. */
typedef struct cg_gemm* cg_gemm_t;
int cg_gemm_create(const double* coef, int64_t n_coef, int32_t off_smh, const double* tmid, int32_t ns,
                   int32_t nr, int32_t ntb, const double* nu_r, const double* nu_p, const double* eff,
                   const int32_t* rec, const double* dnu, const int32_t* plog, int32_t n_plog,
                   int64_t slab, cg_gemm_t* out);
int cg_getrates_gemm(cg_gemm_t g, const double* P, const double* T, int64_t ng, int64_t ld,
                     const double* Y, double* wdot, void* stream);
int cg_gemm_free(cg_gemm_t g);

#ifdef __cplusplus
}
#endif

#endif /* CHEMGEN_AMD_H */
