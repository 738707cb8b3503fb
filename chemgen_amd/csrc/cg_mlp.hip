// Kernels of the learned getrates surrogate (an MLP (T, log P, Y) -> asinh-scaled wdot):
// BF16 matrix products on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulation) for training and inference, and the FP32/FP64 element-wise work around
// them (features, loss, SiLU backward, Adam, decode, element-conservation residual,
// fallback compaction).  The reference has no counterpart (chemgen/chemistry_nn.py is a
// non-functional prototype); the surrogate is defined by this build.
#include "chemgen_amd.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

extern "C" int cg_internal_fail(int code, const char* msg);

namespace {

constexpr int TM = 64;   // rows of C per workgroup
constexpr int TN = 64;   // columns of C per workgroup
constexpr int TK = 32;   // K per step (one 16x16x32 MFMA)
constexpr int PADK = 40; // LDS row pitch in bf16 (80 B: rows start on different banks)

// SiLU with the hardware exp2 and reciprocal, z * rcp(1 + 2^(-z log2 e)) (1 ulp-class
// approximations; the activation is rounded to BF16 right after).  Large |z| stays exact in the
// limit: 2^(+big) = inf gives z * 0, 2^(-big) = 0 gives z.
__device__ __forceinline__ float silu(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.44269504088896341f));
}

// C[M,N] = A[M,K] B[K,N] (+ bias[N]) with A(i,k) = A[i*sAi + k*sAk], B(k,j) = B[k*sBk + j*sBj].
// 256 threads = 4 wavefronts in a 2x2 arrangement, each owning a 32x32 block of C as 2x2
// MFMA 16x16 tiles.  MFMA operand layout (16x16x32 bf16): lane l holds row/column l % 16 and
// k = 8 (l / 16) .. 8 (l / 16) + 7; the result tile gives lane l column l % 16 and rows
// 4 (l / 16) + v, v = 0..3.
// Epilogue: v = acc + bias; C32 (FP32) gets v (ACCUM: atomically added -- split-K), Cb (BF16)
// gets act ? SiLU(v) : v.  blockIdx.z selects a K range [z*kchunk, (z+1)*kchunk).
template <bool A_CK, bool B_CK, bool ACCUM>
__global__ void __launch_bounds__(256)
gemm_bf16(const bf16* __restrict__ A, int64_t sAi, int64_t sAk, const bf16* __restrict__ B,
          int64_t sBk, int64_t sBj, int M, int N, int K, int kchunk,
          const float* __restrict__ bias, int act, float* __restrict__ C32, int64_t ldc,
          bf16* __restrict__ Cb, int64_t ldcb) {
  __shared__ __attribute__((aligned(16))) bf16 As[TM][PADK];
  __shared__ __attribute__((aligned(16))) bf16 Bs[TN][PADK];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int i0 = blockIdx.x * TM, j0 = blockIdx.y * TN;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int lr = tid >> 2;        // tile row (A) / tile column (B) this thread loads
  const int lk = (tid & 3) * 8;   // first of its 8 k
  const bf16 zero = (bf16)0.0f;
  for (int k0 = kbeg; k0 < kend; k0 += TK) {
    {
      const int i = i0 + lr;
      const int k = k0 + lk;
      if (A_CK && i < M && k + 8 <= kend) {
        *(bf16x8*)&As[lr][lk] = *(const bf16x8*)(A + (int64_t)i * sAi + k);
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q)
          As[lr][lk + q] = (i < M && k + q < kend) ? A[(int64_t)i * sAi + (int64_t)(k + q) * sAk] : zero;
      }
    }
    {
      const int j = j0 + lr;
      const int k = k0 + lk;
      if (B_CK && j < N && k + 8 <= kend) {
        *(bf16x8*)&Bs[lr][lk] = *(const bf16x8*)(B + (int64_t)j * sBj + k);
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q)
          Bs[lr][lk + q] = (j < N && k + q < kend) ? B[(int64_t)(k + q) * sBk + (int64_t)j * sBj] : zero;
      }
    }
    __syncthreads();
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) af[a] = *(const bf16x8*)&As[wi + a * 16 + (lane & 15)][8 * (lane >> 4)];
#pragma unroll
    for (int b = 0; b < 2; ++b) bfr[b] = *(const bf16x8*)&Bs[wj + b * 16 + (lane & 15)][8 * (lane >> 4)];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = j0 + wj + b * 16 + (lane & 15);
      if (j >= N) continue;
      const float bj = (bias && blockIdx.z == 0) ? bias[j] : 0.0f;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wi + a * 16 + 4 * (lane >> 4) + v;
        if (i >= M) continue;
        const float val = acc[a][b][v] + bj;
        if (C32) {
          if (ACCUM) atomicAdd(&C32[(int64_t)i * ldc + j], val);
          else C32[(int64_t)i * ldc + j] = val;
        }
        if (Cb) Cb[(int64_t)i * ldcb + j] = (bf16)(act ? silu(val) : val);
      }
    }
}

// -------- element-wise kernels --------

// Feature f of point m: 0 -> (T - f0) f1, 1 -> (ln P - f2) f3, 2..ns+1 -> (ln(Y_k + 1e-12) - f4) f5,
// ns+2 -> (1000/T - f6) f7, beyond -> 0 (padding).  Split into the load of the state value
// it depends on and the arithmetic, so that loops can issue many loads before the first use.
__device__ __forceinline__ double feature_load(int f, const double* __restrict__ P, const double* __restrict__ T,
                                               const double* __restrict__ Y, int64_t ld, int64_t m, int ns) {
  if (f == 1) return P[m];
  if (f == 0 || f == ns + 2) return T[m];
  if (f < ns + 2) return Y[(int64_t)(f - 2) * ld + m];
  return 1.0;
}

// single precision throughout: the features are rounded to BF16 (8 bits) right after
__device__ __forceinline__ float feature_value(int f, double v, int ns, const float* __restrict__ fs) {
  const float x = (float)v;
  if (f == 0) return (x - fs[0]) * fs[1];
  if (f == 1) return (__logf(x) - fs[2]) * fs[3];
  if (f < ns + 2) return (__logf(x + 1e-12f) - fs[4]) * fs[5];
  if (f == ns + 2) return (1000.0f * __builtin_amdgcn_rcpf(x) - fs[6]) * fs[7];
  return 0.0f;
}

__device__ __forceinline__ float feature(int f, const double* __restrict__ P, const double* __restrict__ T,
                                         const double* __restrict__ Y, int64_t ld, int64_t m, int ns,
                                         const float* __restrict__ fs) {
  return feature_value(f, feature_load(f, P, T, Y, ld, m, ns), ns, fs);
}

// x[m, :] = features of point m in BF16, row-major with width kpad
__global__ void features_k(const double* __restrict__ P, const double* __restrict__ T,
                           const double* __restrict__ Y, int64_t ld, int64_t ng, int ns, int kpad,
                           const float* __restrict__ fscale, bf16* __restrict__ X) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= ng) return;
  bf16* x = X + m * kpad;
  for (int f = 0; f < kpad; ++f) x[f] = (bf16)feature(f, P, T, Y, ld, m, ns, fscale);
}

// z[m, k] = asinh(wdot[k, m] / s_k)  (training targets, FP32)
__global__ void targets_k(const double* __restrict__ W, int64_t ld, int64_t ng, int ns,
                          const double* __restrict__ s, float* __restrict__ Z) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= ng) return;
  for (int k = 0; k < ns; ++k) Z[m * ns + k] = (float)asinh(W[(int64_t)k * ld + m] / s[k]);
}

// loss = sum (zhat - z)^2 / (M ns); G = d loss / d zhat (BF16 for the backward products)
__global__ void mse_grad_k(const float* __restrict__ Zh, const float* __restrict__ Z, int64_t n,
                           float scale, bf16* __restrict__ G, float* __restrict__ loss) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (e < n) {
    const float d = Zh[e] - Z[e];
    G[e] = (bf16)(2.0f * scale * d);
    l = d * d * scale;
  }
  // wave reduction, one atomic per wavefront
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

// dZ = dH * silu'(z) (z: FP32 pre-activation), in BF16
__global__ void silu_back_k(const float* __restrict__ dH, const float* __restrict__ Zp, int64_t n,
                            bf16* __restrict__ dZ) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float z = Zp[e];
  const float sg = 1.0f / (1.0f + __expf(-z));
  dZ[e] = (bf16)(dH[e] * sg * (1.0f + z * (1.0f - sg)));
}

// db[n] = sum_m G[m, n]
__global__ void colsum_k(const bf16* __restrict__ G, int64_t M, int N, float* __restrict__ db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const int64_t r0 = (int64_t)blockIdx.y * 1024;
  const int64_t r1 = r0 + 1024 < M ? r0 + 1024 : M;
  float s = 0.f;
  for (int64_t r = r0; r < r1; ++r) s += (float)G[r * N + j];
  atomicAdd(&db[j], s);
}

// Adam on FP32 master weights, BF16 copy for the products
__global__ void adam_k(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2, bf16* __restrict__ w16) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gr = g[e];
  const float a = b1 * m1[e] + (1.f - b1) * gr;
  const float b = b2 * m2[e] + (1.f - b2) * gr * gr;
  m1[e] = a;
  m2[e] = b;
  const float v = w[e] - lr * (a * c1) / (sqrtf(b * c2) + eps);
  w[e] = v;
  if (w16) w16[e] = (bf16)v;
}

__global__ void to_bf16_k(const float* __restrict__ x, int64_t n, bf16* __restrict__ y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = (bf16)x[e];
}

// wdot[k, m] = s_k sinh(zhat[m, k]) (FP64 out); residual[m] = max_e |sum_k a_ek wdot_k/W_k...|
// element-conservation residual: r = max_e |sum_k a_ek w_k| / sum_k |a_ek w_k|; flag = r > tol
__global__ void decode_k(const float* __restrict__ Zh, int64_t ng, int ns, const double* __restrict__ s,
                         const float* __restrict__ aek, int ne, double tol, double* __restrict__ W,
                         int64_t ld, int* __restrict__ flag, double* __restrict__ resid) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= ng) return;
  double num[8], den[8];
  for (int e = 0; e < ne; ++e) { num[e] = 0.0; den[e] = 0.0; }
  for (int k = 0; k < ns; ++k) {
    const double w = s[k] * sinh((double)Zh[m * ns + k]);
    W[(int64_t)k * ld + m] = w;
    for (int e = 0; e < ne; ++e) {
      const double t = (double)aek[e * ns + k] * w;
      num[e] += t;
      den[e] += fabs(t);
    }
  }
  // a non-finite wdot (overflowing sinh, NaN) makes r NaN and flags the point
  double r = 0.0;
  for (int e = 0; e < ne; ++e) {
    const double q = den[e] > 0.0 ? fabs(num[e]) / den[e] : (den[e] == 0.0 ? 0.0 : den[e]);
    if (!(q <= r)) r = q;
  }
  resid[m] = r;
  flag[m] = (r <= tol) ? 0 : 1;
}

// Compaction of the flagged points: idx[*count ...] = m for flag[m] != 0 and, with
// audit_every > 0, m % audit_every != 0 (the audited points keep their FP64 values); nothing
// while *skip is set.  A workgroup takes kCompactChunk consecutive points, numbers its flagged
// ones with a prefix sum in LDS and reserves their slots with one atomic add, so the counter
// sees one atomic per 4096 points (same-address atomics serialise in L2).
constexpr int kCompactPer = 16;                       // points per thread
constexpr int kCompactChunk = 256 * kCompactPer;      // points per workgroup

__global__ void __launch_bounds__(256)
compact_k(const int* __restrict__ flag, int64_t ng, int audit_every, int* __restrict__ count,
          int64_t* __restrict__ idx, const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ int scan[256];
  __shared__ int base;
  const int64_t b0 = (int64_t)blockIdx.x * kCompactChunk;
  unsigned mask = 0;
#pragma unroll
  for (int i = 0; i < kCompactPer; ++i) {
    const int64_t m = b0 + (int64_t)i * 256 + threadIdx.x;
    const bool f = m < ng && flag[m] != 0 && (audit_every <= 0 || m % audit_every != 0);
    mask |= f ? (1u << i) : 0u;
  }
  const int mine = __popc(mask);
  scan[threadIdx.x] = mine;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {            // inclusive Hillis-Steele scan
    const int v = (int)threadIdx.x >= o ? scan[threadIdx.x - o] : 0;
    __syncthreads();
    scan[threadIdx.x] += v;
    __syncthreads();
  }
  if (threadIdx.x == 255) base = scan[255] ? atomicAdd(count, scan[255]) : 0;
  __syncthreads();
  int q = base + scan[threadIdx.x] - mine;
#pragma unroll
  for (int i = 0; i < kCompactPer; ++i)
    if (mask & (1u << i)) idx[q++] = b0 + (int64_t)i * 256 + threadIdx.x;
}

// The surrogate audit's verdict: a workgroup takes 64 audited points, four lanes per point
// (species k = g, g + 4, ...: four loads in flight per lane instead of one long serial chain);
// off = some species with |Ws - Wr| / (|Wr| + s_k) > err_tol (or not finite).  The integer count
// is order-independent, so the verdict is deterministic.  The last workgroup to finish (a ticket
// in acc[1]) sets *n_bad and *reject = (*n_bad > max_bad * n), zeroes *count_clear (the
// compaction's counter, when given) and clears acc for the next call.
__global__ void __launch_bounds__(256)
audit_k(const double* __restrict__ Ws, int64_t lds, const double* __restrict__ Wr, int64_t ldr, int64_t n,
        int ns, const double* __restrict__ s, double err_tol, double max_bad, int* __restrict__ reject,
        int* __restrict__ n_bad, unsigned* __restrict__ acc, int* __restrict__ count_clear) {
  __shared__ int off_pt[64];
  __shared__ int red;
  const int p = threadIdx.x & 63, g = threadIdx.x >> 6;
  if (threadIdx.x < 64) off_pt[threadIdx.x] = 0;
  if (threadIdx.x == 0) red = 0;
  __syncthreads();
  const int64_t m = (int64_t)blockIdx.x * 64 + p;
  if (m < n) {
    bool off = false;
#pragma unroll 4
    for (int k = g; k < ns; k += 4) {
      const double wr = Wr[(int64_t)k * ldr + m];
      const double q = fabs(Ws[(int64_t)k * lds + m] - wr) / (fabs(wr) + s[k]);
      off |= !(q <= err_tol);
    }
    if (off) off_pt[p] = 1;
  }
  __syncthreads();
  if (threadIdx.x < 64 && off_pt[threadIdx.x]) atomicAdd(&red, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (red) atomicAdd(&acc[0], (unsigned)red);
    __threadfence();
    if (atomicAdd(&acc[1], 1u) == gridDim.x - 1) {
      __threadfence();
      const unsigned bad = atomicAdd(&acc[0], 0u);
      *n_bad = (int)bad;
      *reject = (double)bad > max_bad * (double)n ? 1 : 0;
      if (count_clear) *count_clear = 0;
      acc[0] = 0u;
      acc[1] = 0u;
    }
  }
}

// The fallback plan and the tile lengths of its two FP64 launches over [0, ng) in tiles of
// `tile` points (tile t starts at t tile): a rejected audit, or more than dense_frac of the
// points flagged, sends the whole batch to one launch (*full = ng, *pcount = 0); otherwise
// *full = 0 and the *pcount = *count flagged points are recomputed packed.  Lengths:
// clamp(*pcount - t tile, 0, tile) and clamp(*full - t tile, 0, tile); zero-length tiles do no
// work.
__global__ void fallback_plan_k(const int* __restrict__ count, const int* __restrict__ reject, int64_t ng,
                                double dense_frac, int64_t nt, int tile, int* __restrict__ full,
                                int* __restrict__ pcount, int* __restrict__ len_packed, int* __restrict__ len_full) {
  const int c = *count;
  const bool all = *reject != 0 || (double)c > dense_frac * (double)ng;
  const int64_t np = all ? 0 : c, nf = all ? ng : 0;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    *full = (int)nf;
    *pcount = (int)np;
  }
  if (t >= nt) return;
  const int64_t a = t * tile;
  int64_t lp = np - a, lf = nf - a;
  lp = lp < 0 ? 0 : (lp > tile ? tile : lp);
  lf = lf < 0 ? 0 : (lf > tile ? tile : lf);
  len_packed[t] = (int)lp;
  len_full[t] = (int)lf;
}

// A tile list for the first *count points of [0, cap): tile t = [t tile, min((t+1) tile, *count)),
// length 0 beyond the count (the getrates tile kernels skip those)
__global__ void count_tiles_k(const int* __restrict__ count, int64_t nt, int tile, int64_t* __restrict__ start,
                              int* __restrict__ len) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int64_t a = t * tile, c = *count;
  int64_t l = c - a;
  l = l < 0 ? 0 : (l > tile ? tile : l);
  start[t] = a;
  len[t] = (int)l;
}

// gather P, T, Y of the listed points into packed arrays (ldo >= cap); a device count above
// cap is clamped to cap.  A fixed grid strides over (row, point) pairs, row ns holding P and T,
// so an empty list costs one small launch and a long one spreads over the whole GPU.
__global__ void gather_k(const double* __restrict__ P, const double* __restrict__ T,
                         const double* __restrict__ Y, int64_t ld, const int64_t* __restrict__ idx,
                         const int* __restrict__ count, int64_t cap, int ns, double* __restrict__ Po,
                         double* __restrict__ To, double* __restrict__ Yo, int64_t ldo) {
  const int64_t c = *count;
  const int64_t n = c < cap ? c : cap;
  const int64_t total = n * (ns + 1);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e / n, q = e - k * n;
    const int64_t m = idx[q];
    if (k < ns) {
      Yo[k * ldo + q] = Y[k * ld + m];
    } else {
      Po[q] = P[m];
      To[q] = T[m];
    }
  }
}

__global__ void scatter_k(const double* __restrict__ Wi, int64_t ldi, const int64_t* __restrict__ idx,
                          const int* __restrict__ count, int64_t cap, int ns, double* __restrict__ W,
                          int64_t ld) {
  const int64_t c = *count;
  const int64_t n = c < cap ? c : cap;
  const int64_t total = n * ns;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e / n, q = e - k * n;
    W[k * ld + idx[q]] = Wi[k * ldi + q];
  }
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
// grid of the grid-stride kernels: one thread per element, at most 1024 workgroups of 256
inline unsigned strided_grid(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}


// ---------------- fused inference ----------------
//
// One launch runs the whole network for 16*PT points per wavefront (NW wavefronts per
// workgroup: 8, two per SIMD, except for batches too small to fill the GPU) and decodes the
// result: features -> n_hidden SiLU layers of width
// H -> output layer -> wdot = s sinh(z) in FP64 -> element-conservation residual -> flag.
// The activations never leave registers: the network is evaluated transposed,
// out[n][m] = sum_k W[n][k] h[m][k], with the weights as the MFMA A operand (read from LDS)
// and the activations as the B operand.  A 16x16 result tile gives lane l the outputs
// n = 4 (l/16) + v of point l % 16, so the k order inside every 32-wide k-chunk is permuted
// to match (kperm): chunk c, lane group g, operand element j holds k = kperm(c, g, j), which
// is exactly what the result tiles 2c and 2c+1 leave in that lane.  The packed weights
// (cg_mlp_pack) use the same order, so one layer's results feed the next after bias + SiLU +
// BF16 rounding, with no data movement.
//
// The weights stream through LDS in 64 KB chunks (NG output tiles x the layer's k-chunks),
// double-buffered: while the wavefronts multiply chunk i out of one buffer, chunk i+1 is
// copied global -> LDS directly (global_load_lds_dwordx4, no registers) into the other.  A
// 1 KB operand fragment (64 lanes x 16 B) is one such copy instruction, so LDS holds the
// fragments in exactly the order the MFMA loads read them.

__device__ __forceinline__ int kperm(int c, int g, int j) { return 32 * c + 4 * g + (j < 4 ? j : 12 + j); }

// Wf[((nt * KC + kc) * 64 + lane) * 8 + j] = W[16 nt + lane % 16][kperm(kc, lane / 16, j)], zero
// outside the N x K matrix W (row-major, BF16).
__global__ void pack_frag_k(const bf16* __restrict__ W, int N, int K, int NT, int KC, bf16* __restrict__ Wf) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)NT * KC * 512) return;
  const int j = (int)(e & 7), lane = (int)((e >> 3) & 63);
  const int64_t f = e >> 9;
  const int kc = (int)(f % KC), nt = (int)(f / KC);
  const int n = 16 * nt + (lane & 15), k = kperm(kc, lane >> 4, j);
  Wf[e] = (n < N && k < K) ? W[(int64_t)n * K + k] : (bf16)0.0f;
}

// sinh in single precision: the odd series below |z| = 1/8 (relative error < 1e-9 there), the
// exponential form above; overflow gives +-inf, which the residual test then flags.
__device__ __forceinline__ float infer_sinh(float z) {
  const float z2 = z * z;
  const float e = __builtin_amdgcn_exp2f(z * 1.44269504088896341f);
  const float big = 0.5f * (e - __builtin_amdgcn_rcpf(e));
  const float small = z * (1.0f + z2 * (1.0f / 6.0f + z2 * (1.0f / 120.0f)));
  return fabsf(z) < 0.125f ? small : big;
}

// PTV: 16-point tiles per wavefront, NWV: wavefronts per workgroup.  Two tiles and eight
// wavefronts (256 points per workgroup, two wavefronts per SIMD) for H = 256 by default; fewer
// cut a workgroup's latency, which pays when the batch does not fill the GPU.
template <int H, int PTV = (H == 256 ? 2 : 1), int NWV = 8>
struct InferShape {
  static constexpr int NW = NWV;                  // wavefronts per workgroup
  static constexpr int KCH = H / 32;              // k-chunks of a hidden layer
  static constexpr int NTH = H / 16;              // output tiles of a hidden layer
  static constexpr int NG = 64 / KCH;             // output tiles per 64 KB weight chunk
  static constexpr int NCH = NTH / NG;            // chunks of a hidden layer
  static constexpr int PT = PTV;                  // 16-point tiles per wavefront
  static constexpr int NP = 16 * PT;              // points per wavefront
  static constexpr int LPP = 64 / NP;             // lanes per point (features, decode)
  static constexpr int KP = 64 / (NW * PT);       // feature k-chunks staged per pass (64 KB)
  static constexpr int TD = 8;                    // output tiles per decode pass
  static constexpr int ZW = TD * 16 + 1;          // decode buffer pitch (floats)
  static constexpr int Z_BYTES = NW * NP * ZW * 4;
  static constexpr int LDS_BYTES = Z_BYTES > 131072 ? Z_BYTES : 131072;
};

typedef __attribute__((address_space(1))) const void* cg_gptr;
typedef __attribute__((address_space(3))) void* cg_lptr;

// Issue the direct global -> LDS copy of weight chunk `ci` of the whole network into buffer
// ci & 1 (fragment f of the chunk -> LDS fragment slot f; wavefront w copies f = w, w+NW, ...).
// Chunks: layer 0 (NCH of NG x kc0 fragments), hidden layers (NCH of NG x KCH), output
// (ceil(nt_out / NG) of up to NG x KCH).
template <int H, int NWV>
__device__ __forceinline__ void infer_prefetch(const bf16x8* __restrict__ Wf, int ci, int kc0, int n_hidden,
                                               int nt_out, bf16x8* wlds) {
  using S = InferShape<H, 1, NWV>;
  int64_t off;
  int nfrag;
  if (ci < S::NCH) {
    off = (int64_t)ci * S::NG * kc0;
    nfrag = S::NG * kc0;
  } else {
    const int64_t base = (int64_t)S::NTH * kc0;
    const int c = ci - S::NCH;
    off = base + (int64_t)c * S::NG * S::KCH;
    const int co = c - (n_hidden - 1) * S::NCH;   // chunk index within the output layer
    if (co < 0) {
      nfrag = S::NG * S::KCH;
    } else {
      const int nt = nt_out - co * S::NG;
      nfrag = (nt < S::NG ? nt : S::NG) * S::KCH;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  bf16x8* dst = wlds + (ci & 1) * 4096;
  const bf16x8* src = Wf + off * 64;
  for (int f = wave; f < nfrag; f += S::NW)
    __builtin_amdgcn_global_load_lds((cg_gptr)(src + f * 64 + lane), (cg_lptr)(dst + f * 64), 16, 0, 0);
}

// One layer: acc[t][pt] = sum over k of weights x act (no bias), chunk by chunk; chunk `ci`
// (running index over the network) is in buffer ci & 1 and the next one is prefetched.  MODE
// fixes which extent is known at compile time, so the MFMA blocks carry no per-fragment
// branches: 0 = first layer (runtime kc0), 1 = hidden layer, 2 = output layer (runtime nt_out).
template <int H, int PTV, int NWV, int MODE>
__device__ __forceinline__ void infer_layer(const bf16x8* __restrict__ Wf, int& ci, int n_chunks, int kc0,
                                            int n_hidden, int nt_out, bf16x8* __restrict__ wlds,
                                            const float* __restrict__ bl,
                                            const bf16x8 (&act)[H / 32][PTV],
                                            f32x4 (&acc)[H / 16][PTV]) {
  using S = InferShape<H, PTV, NWV>;
  const int lane = threadIdx.x & 63;
  const int kc_n = MODE == 0 ? kc0 : S::KCH;
  const int nt_n = MODE == 2 ? nt_out : S::NTH;
  // hidden layers start from their bias (rows 16 t + 4 g + v of this lane); the output layer
  // from zero, its bias is added in the decode
#pragma unroll
  for (int t = 0; t < S::NTH; ++t) {
    const f32x4 b = MODE == 2 ? f32x4{0.f, 0.f, 0.f, 0.f} : *(const f32x4*)(bl + 16 * t + 4 * (lane >> 4));
#pragma unroll
    for (int pt = 0; pt < S::PT; ++pt) acc[t][pt] = b;
  }
#pragma unroll
  for (int c0 = 0; c0 < S::NTH; c0 += S::NG) {
    if (c0 < nt_n) {
      __builtin_amdgcn_s_waitcnt(0);   // this wavefront's copies of chunk ci have landed
      __syncthreads();                 // ... everyone's, and chunk ci-1's buffer is free
      if (ci + 1 < n_chunks) infer_prefetch<H, NWV>(Wf, ci + 1, kc0, n_hidden, nt_out, wlds);
      const bf16x8* __restrict__ buf = wlds + (ci & 1) * 4096;
      // Steps of Q tiles x one k-chunk; the fragments of step st+1 are loaded (into the other
      // half of F) before the MFMAs of step st are issued -- sched_barrier keeps that order --
      // so LDS latency hides behind Q x PT matrix instructions.  The output layer multiplies
      // all NG tiles of a chunk; tiles past nt_out hold stale LDS data and their results are
      // never read.
      constexpr int Q = 4, NQ = S::NG / Q;
      const int n_steps = kc_n * NQ;
      bf16x8 F[2][Q];
#pragma unroll
      for (int j = 0; j < Q; ++j) F[0][j] = buf[j * kc_n * 64 + lane];
#pragma unroll
      for (int st = 0; st < S::KCH * NQ; ++st) {
        const int kc = st / NQ, t0 = (st % NQ) * Q;
        if (MODE != 0 || st < n_steps) {
          if (st + 1 < S::KCH * NQ && (MODE != 0 || st + 1 < n_steps)) {
            const int kn = (st + 1) / NQ, tn = ((st + 1) % NQ) * Q;
#pragma unroll
            for (int j = 0; j < Q; ++j) F[(st + 1) & 1][j] = buf[((tn + j) * kc_n + kn) * 64 + lane];
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int j = 0; j < Q; ++j)
#pragma unroll
            for (int pt = 0; pt < S::PT; ++pt)
              acc[c0 + t0 + j][pt] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[st & 1][j], act[kc][pt], acc[c0 + t0 + j][pt], 0, 0, 0);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      ++ci;
    }
  }
}

// SiLU + BF16 rounding (the bias is already in acc): a hidden layer's results become the next
// layer's operand
template <int H, int PTV>
__device__ __forceinline__ void infer_activate(const f32x4 (&acc)[H / 16][PTV],
                                               bf16x8 (&act)[H / 32][PTV]) {
  using S = InferShape<H, PTV>;
#pragma unroll
  for (int kc = 0; kc < S::KCH; ++kc) {
#pragma unroll
    for (int pt = 0; pt < S::PT; ++pt) {
      bf16x8 x;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        x[v] = (bf16)silu(acc[2 * kc][pt][v]);
        x[4 + v] = (bf16)silu(acc[2 * kc + 1][pt][v]);
      }
      act[kc][pt] = x;
    }
  }
}

// Wf: packed layers back to back (layer 0: NTH x kc0 fragments, hidden layers: NTH x KCH,
// output: nt_out x KCH); bias: hidden layers H each, then the output layer's ns values.
template <int H, int PTV, int NWV>
__global__ void __launch_bounds__(64 * NWV) __attribute__((amdgpu_waves_per_eu(2, 2)))
mlp_infer_k(const double* __restrict__ P, const double* __restrict__ T, const double* __restrict__ Y,
            int64_t ld, int64_t ng, int ns, const float* __restrict__ fs, const bf16x8* __restrict__ Wf,
            const float* __restrict__ bias, int kc0, int n_hidden, int nt_out, const double* __restrict__ s,
            const float* __restrict__ aek, int ne, double tol, double* __restrict__ W, int64_t ldw,
            int* __restrict__ flag, double* __restrict__ resid, const int* __restrict__ skip) {
  using S = InferShape<H, PTV, NWV>;
  constexpr int PT = S::PT, NP = S::NP, LPP = S::LPP;
  // a device-side verdict (the surrogate's audit) can cancel the launch: every workgroup reads
  // it first, before any memory traffic, and leaves
  if (skip && *skip) return;
  extern __shared__ __attribute__((aligned(16))) bf16x8 wlds[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t base = ((int64_t)blockIdx.x * S::NW + wave) * NP;
  const int p = lane % NP, part = lane / NP;  // decode / feature roles: point p, share `part`
  const int64_t m = base + p;
  const int64_t mc = m < ng ? m : ng - 1;
  const int n_chunks = S::NCH * n_hidden + (nt_out + S::NG - 1) / S::NG;
  int ci = 0;
  infer_prefetch<H, NWV>(Wf, 0, kc0, n_hidden, nt_out, wlds);   // lands in buffer 0 meanwhile
  // the decode's per-output constants (s_n, the output bias, the element matrix) in LDS past
  // the weight and decode buffers: read there by every output of every point, they had been
  // ten vector-memory loads per output (10-13% of the kernel's time)
  double* const cs = (double*)((char*)wlds + S::LDS_BYTES);
  float* const cb = (float*)(cs + ns);
  float* const ca = cb + ns;
  {
    const float* bl0 = bias + n_hidden * H;
    for (int q = threadIdx.x; q < ns; q += 64 * S::NW) {
      cs[q] = s[q];
      cb[q] = bl0[q];
    }
    for (int q = threadIdx.x; q < ne * ns; q += 64 * S::NW) ca[q] = aek[q];
  }   // visible to the decode after the layers' barriers

  // 1. features, staged in operand order through this wavefront's part of buffer 1, KP
  //    k-chunks per pass
  bf16x8 act[S::KCH][PT];
  const bf16x8 zero8 = {};
  bf16* xl = (bf16*)(wlds + 4096 + wave * (S::KP * PT * 64));
  for (int k_lo = 0; k_lo < kc0; k_lo += S::KP) {
    const int k_hi = (k_lo + S::KP) < kc0 ? (k_lo + S::KP) : kc0;
    __syncthreads();
    // eight loads in flight per lane before the first logarithm (one wavefront per SIMD:
    // nothing else would hide their latency)
    for (int f0 = 32 * k_lo + part; f0 < 32 * k_hi; f0 += 8 * LPP) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int f = f0 + u * LPP;
        v[u] = f < 32 * k_hi ? feature_load(f, P, T, Y, ld, mc, ns) : 1.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int f = f0 + u * LPP;
        if (f < 32 * k_hi) {
          const int kc = (f >> 5) - k_lo, q = f & 31;
          const int fg = (q >> 2) & 3, j = (q & 3) | ((q >> 4) << 2);
          xl[(((kc * PT + (p >> 4)) * 64 + 16 * fg + (p & 15)) << 3) + j] = (bf16)feature_value(f, v[u], ns, fs);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int kc = 0; kc < S::KCH; ++kc)
      if (kc >= k_lo && kc < k_hi)
#pragma unroll
        for (int pt = 0; pt < PT; ++pt) act[kc][pt] = ((const bf16x8*)xl)[((kc - k_lo) * PT + pt) * 64 + lane];
  }
#pragma unroll
  for (int kc = 0; kc < S::KCH; ++kc)
    if (kc >= kc0)
#pragma unroll
      for (int pt = 0; pt < PT; ++pt) act[kc][pt] = zero8;

  // 2. layers
  f32x4 acc[S::NTH][PT];
  infer_layer<H, PT, NWV, 0>(Wf, ci, n_chunks, kc0, n_hidden, nt_out, wlds, bias, act, acc);
  infer_activate<H, PT>(acc, act);
  for (int l = 1; l < n_hidden; ++l) {
    infer_layer<H, PT, NWV, 1>(Wf, ci, n_chunks, kc0, n_hidden, nt_out, wlds, bias + l * H, act, acc);
    infer_activate<H, PT>(acc, act);
  }
  infer_layer<H, PT, NWV, 2>(Wf, ci, n_chunks, kc0, n_hidden, nt_out, wlds, nullptr, act, acc);

  // 3. decode through LDS, TD output tiles at a time: point p's z values are read by LPP lanes
  float* zl = (float*)wlds + wave * (NP * S::ZW);
  float num[8], den[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) { num[e] = 0.0f; den[e] = 0.0f; }
  const int g = lane >> 4;
#pragma unroll
  for (int t0 = 0; t0 < S::NTH; t0 += S::TD) {
    if (t0 < nt_out) {
      __syncthreads();   // LDS free: last weight chunk (first pass) or the previous pass's z read
#pragma unroll
      for (int tl = 0; tl < S::TD; ++tl) {
        if (t0 + tl < nt_out) {
#pragma unroll
          for (int pt = 0; pt < PT; ++pt)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              zl[(16 * pt + (lane & 15)) * S::ZW + 16 * tl + 4 * g + v] = acc[t0 + tl][pt][v];
        }
      }
      __syncthreads();
      const int nend = (ns - 16 * t0) < 16 * S::TD ? (ns - 16 * t0) : 16 * S::TD;
#pragma unroll 2
      for (int nl = part; nl < nend; nl += LPP) {
        const int n = 16 * t0 + nl;
        const float z = zl[p * S::ZW + nl] + cb[n];
        const float sh = infer_sinh(z);
        const double sn = cs[n];
        if (m < ng) W[(int64_t)n * ldw + m] = sn * (double)sh;
        // the residual needs ~1e-3: single precision, scaled by s_n (its magnitude only)
        const float w = (float)sn * sh;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const float q = (e < ne ? ca[(e < ne ? e : 0) * ns + n] : 0.0f) * w;
          num[e] += q;
          den[e] += fabsf(q);
        }
      }
    }
  }
  float rr = 0.0f;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    float a = num[e], d = den[e];
#pragma unroll
    for (int o = NP; o < 64; o *= 2) {
      a += __shfl_xor(a, o);
      d += __shfl_xor(d, o);
    }
    // rows past ne are zero: d == 0 gives 0; a non-finite wdot makes q NaN and flags the point
    const float q = d > 0.0f ? fabsf(a) / d : (d == 0.0f ? 0.0f : d);
    if (!(q <= rr)) rr = q;
  }
  if (part == 0 && m < ng) {
    resid[m] = rr;
    flag[m] = (rr <= tol) ? 0 : 1;
  }
}

}  // namespace

#define CG_LAUNCH_CHECK(what)                                                  \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) return cg_internal_fail(CG_EHIP, what);              \
  } while (0)

extern "C" int cg_mlp_gemm(const void* A, int64_t sAi, int64_t sAk, const void* B, int64_t sBk,
                           int64_t sBj, int32_t M, int32_t N, int32_t K, int32_t ksplit,
                           const float* bias, int32_t act, float* C32, int64_t ldc, void* Cb,
                           int64_t ldcb, int32_t accumulate, void* stream) {
  if (!A || !B || M < 0 || N < 0 || K < 0 || (!C32 && !Cb) || ksplit < 1)
    return cg_internal_fail(CG_EINVAL, "cg_mlp_gemm: bad argument");
  if (accumulate && !C32) return cg_internal_fail(CG_EINVAL, "cg_mlp_gemm: accumulate needs C32");
  if (M == 0 || N == 0) return CG_OK;
  int kchunk = (K + ksplit - 1) / ksplit;
  kchunk = (kchunk + TK - 1) / TK * TK;
  const int nz = kchunk > 0 ? (K + kchunk - 1) / kchunk : 1;
  if (nz > 1 && !accumulate) return cg_internal_fail(CG_EINVAL, "cg_mlp_gemm: split K needs accumulate");
  dim3 grid(nblk(M, TM), nblk(N, TN), nz > 0 ? nz : 1);
  const bool ack = (sAk == 1) && (sAi % 8 == 0) && (((uintptr_t)A & 15) == 0);
  const bool bck = (sBk == 1) && (sBj % 8 == 0) && (((uintptr_t)B & 15) == 0);
  const bf16* a = (const bf16*)A;
  const bf16* b = (const bf16*)B;
  hipStream_t s = (hipStream_t)stream;
#define CG_GEMM(AC, BC, ACC)                                                                    \
  gemm_bf16<AC, BC, ACC><<<grid, 256, 0, s>>>(a, sAi, sAk, b, sBk, sBj, M, N, K, kchunk, bias, \
                                               act, C32, ldc, (bf16*)Cb, ldcb)
  if (accumulate) {
    if (ack && bck) CG_GEMM(true, true, true);
    else if (ack) CG_GEMM(true, false, true);
    else if (bck) CG_GEMM(false, true, true);
    else CG_GEMM(false, false, true);
  } else {
    if (ack && bck) CG_GEMM(true, true, false);
    else if (ack) CG_GEMM(true, false, false);
    else if (bck) CG_GEMM(false, true, false);
    else CG_GEMM(false, false, false);
  }
#undef CG_GEMM
  CG_LAUNCH_CHECK("cg_mlp_gemm launch");
  return CG_OK;
}

extern "C" int cg_mlp_features(const double* P, const double* T, const double* Y, int64_t ld,
                               int64_t ng, int32_t ns, int32_t kpad, const float* fscale, void* X,
                               void* stream) {
  if (ng < 0 || ns + 2 > kpad || !fscale || !X) return cg_internal_fail(CG_EINVAL, "cg_mlp_features: bad argument");
  if (ng == 0) return CG_OK;
  features_k<<<nblk(ng, 256), 256, 0, (hipStream_t)stream>>>(P, T, Y, ld, ng, ns, kpad, fscale, (bf16*)X);
  CG_LAUNCH_CHECK("cg_mlp_features launch");
  return CG_OK;
}

extern "C" int cg_mlp_targets(const double* W, int64_t ld, int64_t ng, int32_t ns, const double* s,
                              float* Z, void* stream) {
  if (ng < 0 || ns <= 0 || ld < ng || (ng > 0 && (!W || !s || !Z)))
    return cg_internal_fail(CG_EINVAL, "cg_mlp_targets: bad argument");
  if (ng == 0) return CG_OK;
  targets_k<<<nblk(ng, 256), 256, 0, (hipStream_t)stream>>>(W, ld, ng, ns, s, Z);
  CG_LAUNCH_CHECK("cg_mlp_targets launch");
  return CG_OK;
}

extern "C" int cg_mlp_mse_grad(const float* Zh, const float* Z, int64_t n, float scale, void* G,
                               float* loss, void* stream) {
  if (n < 0 || (n > 0 && (!Zh || !Z || !G))) return cg_internal_fail(CG_EINVAL, "cg_mlp_mse_grad: bad argument");
  if (n == 0) return CG_OK;
  mse_grad_k<<<nblk(n, 256), 256, 0, (hipStream_t)stream>>>(Zh, Z, n, scale, (bf16*)G, loss);
  CG_LAUNCH_CHECK("cg_mlp_mse_grad launch");
  return CG_OK;
}

extern "C" int cg_mlp_silu_back(const float* dH, const float* Zp, int64_t n, void* dZ, void* stream) {
  if (n < 0 || (n > 0 && (!dH || !Zp || !dZ))) return cg_internal_fail(CG_EINVAL, "cg_mlp_silu_back: bad argument");
  if (n == 0) return CG_OK;
  silu_back_k<<<nblk(n, 256), 256, 0, (hipStream_t)stream>>>(dH, Zp, n, (bf16*)dZ);
  CG_LAUNCH_CHECK("cg_mlp_silu_back launch");
  return CG_OK;
}

extern "C" int cg_mlp_colsum(const void* G, int64_t M, int32_t N, float* db, void* stream) {
  if (M < 0 || N < 0 || (M > 0 && N > 0 && (!G || !db))) return cg_internal_fail(CG_EINVAL, "cg_mlp_colsum: bad argument");
  if (M == 0 || N == 0) return CG_OK;
  dim3 grid(nblk(N, 256), nblk(M, 1024));
  colsum_k<<<grid, 256, 0, (hipStream_t)stream>>>((const bf16*)G, M, N, db);
  CG_LAUNCH_CHECK("cg_mlp_colsum launch");
  return CG_OK;
}

extern "C" int cg_mlp_adam(float* w, const float* g, float* m1, float* m2, int64_t n, float lr,
                           float b1, float b2, float eps, int32_t step, void* w16, void* stream) {
  if (n < 0 || step < 1 || (n > 0 && (!w || !g || !m1 || !m2)))
    return cg_internal_fail(CG_EINVAL, "cg_mlp_adam: bad argument");
  if (n == 0) return CG_OK;
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
  const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
  adam_k<<<nblk(n, 256), 256, 0, (hipStream_t)stream>>>(w, g, m1, m2, n, lr, b1, b2, eps, c1, c2, (bf16*)w16);
  CG_LAUNCH_CHECK("cg_mlp_adam launch");
  return CG_OK;
}

extern "C" int cg_mlp_to_bf16(const float* x, int64_t n, void* y, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y))) return cg_internal_fail(CG_EINVAL, "cg_mlp_to_bf16: bad argument");
  if (n == 0) return CG_OK;
  to_bf16_k<<<nblk(n, 256), 256, 0, (hipStream_t)stream>>>(x, n, (bf16*)y);
  CG_LAUNCH_CHECK("cg_mlp_to_bf16 launch");
  return CG_OK;
}

extern "C" int cg_mlp_decode(const float* Zh, int64_t ng, int32_t ns, const double* s, const float* aek,
                             int32_t ne, double tol, double* W, int64_t ld, int32_t* flag, double* resid,
                             void* stream) {
  if (ne > 8) return cg_internal_fail(CG_ERANGE, "cg_mlp_decode: at most 8 elements");
  if (ng < 0 || ns <= 0 || ne < 0 || ld < ng || (ng > 0 && (!Zh || !s || !W || !flag || !resid || (ne > 0 && !aek))))
    return cg_internal_fail(CG_EINVAL, "cg_mlp_decode: bad argument");
  if (ng == 0) return CG_OK;
  decode_k<<<nblk(ng, 256), 256, 0, (hipStream_t)stream>>>(Zh, ng, ns, s, aek, ne, tol, W, ld, flag, resid);
  CG_LAUNCH_CHECK("cg_mlp_decode launch");
  return CG_OK;
}

extern "C" int cg_mlp_compact(const int32_t* flag, int64_t ng, int32_t audit_every, int32_t* count, int64_t* idx,
                              const int32_t* skip, void* stream) {
  if (ng < 0 || ng > 0x7fffffffLL || (ng > 0 && (!flag || !count || !idx)))
    return cg_internal_fail(CG_EINVAL, "cg_mlp_compact: bad argument");
  if (ng == 0) return CG_OK;
  compact_k<<<nblk(ng, kCompactChunk), 256, 0, (hipStream_t)stream>>>(flag, ng, audit_every, count, idx, skip);
  CG_LAUNCH_CHECK("cg_mlp_compact launch");
  return CG_OK;
}

extern "C" int cg_mlp_audit(const double* Ws, int64_t lds, const double* Wr, int64_t ldr, int64_t n, int32_t ns,
                            const double* s, double err_tol, double max_bad, int32_t* reject, int32_t* n_bad,
                            uint32_t* acc, int32_t* count_clear, void* stream) {
  if (n <= 0 || ns <= 0 || lds < n || ldr < n || !reject || !n_bad || !acc || !Ws || !Wr || !s)
    return cg_internal_fail(CG_EINVAL, "cg_mlp_audit: bad argument");
  if (n > 0x7fffffffLL) return cg_internal_fail(CG_ERANGE, "cg_mlp_audit: n too large");
  audit_k<<<nblk(n, 64), 256, 0, (hipStream_t)stream>>>(Ws, lds, Wr, ldr, n, ns, s, err_tol, max_bad, reject, n_bad,
                                                         acc, count_clear);
  CG_LAUNCH_CHECK("cg_mlp_audit launch");
  return CG_OK;
}

extern "C" int cg_mlp_fallback_plan(const int32_t* count, const int32_t* reject, int64_t ng, double dense_frac,
                                    int32_t tile, int32_t* full, int32_t* pcount, int32_t* len_packed,
                                    int32_t* len_full, void* stream) {
  if (ng <= 0 || ng > 0x7fffffffLL || tile <= 0 || !count || !reject || !full || !pcount || !len_packed ||
      !len_full)
    return cg_internal_fail(CG_EINVAL, "cg_mlp_fallback_plan: bad argument");
  const int64_t nt = (ng + tile - 1) / tile;
  fallback_plan_k<<<nblk(nt, 256), 256, 0, (hipStream_t)stream>>>(count, reject, ng, dense_frac, nt, tile, full,
                                                                  pcount, len_packed, len_full);
  CG_LAUNCH_CHECK("cg_mlp_fallback_plan launch");
  return CG_OK;
}

extern "C" int cg_count_tiles(const int32_t* count, int64_t cap, int32_t tile, int64_t* tile_start,
                              int32_t* tile_len, void* stream) {
  if (cap < 0 || tile <= 0 || !count || (cap > 0 && (!tile_start || !tile_len)))
    return cg_internal_fail(CG_EINVAL, "cg_count_tiles: bad argument");
  const int64_t nt = (cap + tile - 1) / tile;
  if (nt == 0) return CG_OK;
  count_tiles_k<<<nblk(nt, 256), 256, 0, (hipStream_t)stream>>>(count, nt, tile, tile_start, tile_len);
  CG_LAUNCH_CHECK("cg_count_tiles launch");
  return CG_OK;
}

extern "C" int cg_mlp_gather(const double* P, const double* T, const double* Y, int64_t ld,
                             const int64_t* idx, const int32_t* count, int64_t cap, int32_t ns,
                             double* Po, double* To, double* Yo, int64_t ldo, void* stream) {
  if (cap < 0 || ns <= 0 || ldo < cap || (cap > 0 && (!P || !T || !Y || !idx || !count || !Po || !To || !Yo)))
    return cg_internal_fail(CG_EINVAL, "cg_mlp_gather: bad argument");
  if (cap == 0) return CG_OK;
  gather_k<<<strided_grid(cap * (ns + 1)), 256, 0, (hipStream_t)stream>>>(P, T, Y, ld, idx, count, cap, ns, Po, To,
                                                                          Yo, ldo);
  CG_LAUNCH_CHECK("cg_mlp_gather launch");
  return CG_OK;
}

extern "C" int cg_mlp_scatter(const double* Wi, int64_t ldi, const int64_t* idx, const int32_t* count,
                              int64_t cap, int32_t ns, double* W, int64_t ld, void* stream) {
  if (cap < 0 || ns <= 0 || ldi < cap || (cap > 0 && (!Wi || !idx || !count || !W)))
    return cg_internal_fail(CG_EINVAL, "cg_mlp_scatter: bad argument");
  if (cap == 0) return CG_OK;
  scatter_k<<<strided_grid(cap * ns), 256, 0, (hipStream_t)stream>>>(Wi, ldi, idx, count, cap, ns, W, ld);
  CG_LAUNCH_CHECK("cg_mlp_scatter launch");
  return CG_OK;
}

extern "C" int cg_mlp_pack(const void* W, int32_t N, int32_t K, int32_t NT, int32_t KC, void* Wf, void* stream) {
  if (N <= 0 || K <= 0 || NT * 16 < N || KC * 32 < K) return cg_internal_fail(CG_EINVAL, "cg_mlp_pack: bad shape");
  const int64_t n = (int64_t)NT * KC * 512;
  pack_frag_k<<<nblk(n, 256), 256, 0, (hipStream_t)stream>>>((const bf16*)W, N, K, NT, KC, (bf16*)Wf);
  CG_LAUNCH_CHECK("cg_mlp_pack launch");
  return CG_OK;
}

namespace {
template <int H, int PTV, int NWV>
int launch_infer(const double* P, const double* T, const double* Y, int64_t ld, int64_t ng, int32_t ns,
                 const float* fscale, const void* Wf, const float* bias, int32_t kc0, int32_t n_hidden,
                 int32_t nt_out, const double* s, const float* aek, int32_t ne, double tol, double* W,
                 int64_t ldw, int32_t* flag, double* resid, const int32_t* skip, hipStream_t stream) {
  using S = InferShape<H, PTV, NWV>;
  // LDS: weight / decode buffers, then the decode constants (s_n as doubles, the output bias
  // and ne element rows as floats)
  constexpr int kMaxLds = 160 * 1024;
  const int64_t lds = S::LDS_BYTES + (int64_t)ns * (8 + 4 + 4 * ne);
  if (lds > kMaxLds) return cg_internal_fail(CG_EINVAL, "cg_mlp_infer: too many species for the LDS");
  // the LDS limit is raised once per device and kernel instance
  static bool attr_set[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return cg_internal_fail(CG_EHIP, "cg_mlp_infer: no usable current device");
  if (!attr_set[dev]) {
    hipError_t e = hipFuncSetAttribute((const void*)mlp_infer_k<H, PTV, NWV>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds);
    if (e != hipSuccess) return cg_internal_fail(CG_EHIP, "cg_mlp_infer: cannot reserve LDS");
    attr_set[dev] = true;
  }
  const int64_t nb = (ng + S::NW * S::NP - 1) / (S::NW * S::NP);
  if (nb > 0x7fffffffLL) return cg_internal_fail(CG_ERANGE, "cg_mlp_infer: ng too large");
  mlp_infer_k<H, PTV, NWV><<<(unsigned)nb, 64 * S::NW, (unsigned)lds, stream>>>(
      P, T, Y, ld, ng, ns, fscale, (const bf16x8*)Wf, bias, kc0, n_hidden, nt_out, s, aek, ne, tol, W, ldw,
      flag, resid, skip);
  CG_LAUNCH_CHECK("cg_mlp_infer launch");
  return CG_OK;
}
}  // namespace

extern "C" int cg_mlp_infer(const double* P, const double* T, const double* Y, int64_t ld, int64_t ng,
                            int32_t ns, const float* fscale, const void* Wf, const float* bias, int32_t hidden,
                            int32_t kin, int32_t n_hidden, const double* s, const float* aek, int32_t ne,
                            double tol, double* W, int64_t ldw, int32_t* flag, double* resid,
                            const int32_t* skip, void* stream) {
  if (ng == 0) return CG_OK;
  if (ng < 0 || ld < ng || ldw < ng || ns <= 0 || n_hidden < 1 || ne < 0 || ne > 8)
    return cg_internal_fail(CG_EINVAL, "cg_mlp_infer: bad argument");
  if (hidden != 256 && hidden != 512)
    return cg_internal_fail(CG_EINVAL, "cg_mlp_infer: hidden width must be 256 or 512");
  const int32_t kc0 = (kin + 31) / 32, nt_out = (ns + 15) / 16;
  if (kin < ns + 2 || kc0 > hidden / 32 || nt_out > hidden / 16)
    return cg_internal_fail(CG_EINVAL, "cg_mlp_infer: input/output wider than the hidden layers");
  if (hidden == 256) {
    // a batch of at most one 128-point workgroup per CU (the surrogate's audit, small calls)
    // finishes in one workgroup's latency: half-size tiles halve it.  CU count per device.
    static int n_cu_dev[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
      return cg_internal_fail(CG_EHIP, "cg_mlp_infer: no usable current device");
    int n_cu = n_cu_dev[dev];
    if (n_cu == 0) {
      if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu <= 0)
        return cg_internal_fail(CG_EHIP, "cg_mlp_infer: cannot read the CU count");
      n_cu_dev[dev] = n_cu;
    }
    if (ng <= (int64_t)n_cu * 64)
      return launch_infer<256, 1, 4>(P, T, Y, ld, ng, ns, fscale, Wf, bias, kc0, n_hidden, nt_out, s, aek, ne, tol,
                                     W, ldw, flag, resid, skip, (hipStream_t)stream);
    if (ng <= (int64_t)n_cu * 128)
      return launch_infer<256, 1, 8>(P, T, Y, ld, ng, ns, fscale, Wf, bias, kc0, n_hidden, nt_out, s, aek, ne, tol,
                                  W, ldw, flag, resid, skip, (hipStream_t)stream);
    return launch_infer<256, 2, 8>(P, T, Y, ld, ng, ns, fscale, Wf, bias, kc0, n_hidden, nt_out, s, aek, ne, tol, W,
                                ldw, flag, resid, skip, (hipStream_t)stream);
  }
  return launch_infer<512, 1, 8>(P, T, Y, ld, ng, ns, fscale, Wf, bias, kc0, n_hidden, nt_out, s, aek, ne, tol, W,
                           ldw, flag, resid, skip, (hipStream_t)stream);
}
