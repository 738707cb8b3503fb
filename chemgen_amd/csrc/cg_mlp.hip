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

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

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

// x[m, :] = ((T - f0) f1, (ln P - f2) f3, (ln(Y_k + 1e-12) - f4) f5 ..., 0 ...) in BF16, width kpad
__global__ void features_k(const double* __restrict__ P, const double* __restrict__ T,
                           const double* __restrict__ Y, int64_t ld, int64_t ng, int ns, int kpad,
                           const float* __restrict__ fscale, bf16* __restrict__ X) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= ng) return;
  bf16* x = X + m * kpad;
  x[0] = (bf16)(float)((T[m] - fscale[0]) * fscale[1]);
  x[1] = (bf16)(float)((log(P[m]) - fscale[2]) * fscale[3]);
  for (int k = 0; k < ns; ++k)
    x[2 + k] = (bf16)(float)((log(Y[(int64_t)k * ld + m] + 1e-12) - fscale[4]) * fscale[5]);
  for (int k = ns + 2; k < kpad; ++k) x[k] = (bf16)0.0f;
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

// compaction of flagged points: idx[count++] = m
__global__ void compact_k(const int* __restrict__ flag, int64_t ng, int* __restrict__ count,
                          int64_t* __restrict__ idx) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < ng && flag[m]) idx[atomicAdd(count, 1)] = m;
}

// gather P, T, Y of the listed points into packed arrays (ldo = capacity)
__global__ void gather_k(const double* __restrict__ P, const double* __restrict__ T,
                         const double* __restrict__ Y, int64_t ld, const int64_t* __restrict__ idx,
                         const int* __restrict__ count, int ns, double* __restrict__ Po,
                         double* __restrict__ To, double* __restrict__ Yo, int64_t ldo) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= *count) return;
  const int64_t m = idx[q];
  Po[q] = P[m];
  To[q] = T[m];
  for (int k = 0; k < ns; ++k) Yo[(int64_t)k * ldo + q] = Y[(int64_t)k * ld + m];
}

__global__ void scatter_k(const double* __restrict__ Wi, int64_t ldi, const int64_t* __restrict__ idx,
                          const int* __restrict__ count, int ns, double* __restrict__ W, int64_t ld) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= *count) return;
  const int64_t m = idx[q];
  for (int k = 0; k < ns; ++k) W[(int64_t)k * ld + m] = Wi[(int64_t)k * ldi + q];
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

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
  if (ng == 0) return CG_OK;
  targets_k<<<nblk(ng, 256), 256, 0, (hipStream_t)stream>>>(W, ld, ng, ns, s, Z);
  CG_LAUNCH_CHECK("cg_mlp_targets launch");
  return CG_OK;
}

extern "C" int cg_mlp_mse_grad(const float* Zh, const float* Z, int64_t n, float scale, void* G,
                               float* loss, void* stream) {
  if (n == 0) return CG_OK;
  mse_grad_k<<<nblk(n, 256), 256, 0, (hipStream_t)stream>>>(Zh, Z, n, scale, (bf16*)G, loss);
  CG_LAUNCH_CHECK("cg_mlp_mse_grad launch");
  return CG_OK;
}

extern "C" int cg_mlp_silu_back(const float* dH, const float* Zp, int64_t n, void* dZ, void* stream) {
  if (n == 0) return CG_OK;
  silu_back_k<<<nblk(n, 256), 256, 0, (hipStream_t)stream>>>(dH, Zp, n, (bf16*)dZ);
  CG_LAUNCH_CHECK("cg_mlp_silu_back launch");
  return CG_OK;
}

extern "C" int cg_mlp_colsum(const void* G, int64_t M, int32_t N, float* db, void* stream) {
  if (M == 0) return CG_OK;
  dim3 grid(nblk(N, 256), nblk(M, 1024));
  colsum_k<<<grid, 256, 0, (hipStream_t)stream>>>((const bf16*)G, M, N, db);
  CG_LAUNCH_CHECK("cg_mlp_colsum launch");
  return CG_OK;
}

extern "C" int cg_mlp_adam(float* w, const float* g, float* m1, float* m2, int64_t n, float lr,
                           float b1, float b2, float eps, int32_t step, void* w16, void* stream) {
  if (n == 0) return CG_OK;
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
  const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
  adam_k<<<nblk(n, 256), 256, 0, (hipStream_t)stream>>>(w, g, m1, m2, n, lr, b1, b2, eps, c1, c2, (bf16*)w16);
  CG_LAUNCH_CHECK("cg_mlp_adam launch");
  return CG_OK;
}

extern "C" int cg_mlp_to_bf16(const float* x, int64_t n, void* y, void* stream) {
  if (n == 0) return CG_OK;
  to_bf16_k<<<nblk(n, 256), 256, 0, (hipStream_t)stream>>>(x, n, (bf16*)y);
  CG_LAUNCH_CHECK("cg_mlp_to_bf16 launch");
  return CG_OK;
}

extern "C" int cg_mlp_decode(const float* Zh, int64_t ng, int32_t ns, const double* s, const float* aek,
                             int32_t ne, double tol, double* W, int64_t ld, int32_t* flag, double* resid,
                             void* stream) {
  if (ne > 8) return cg_internal_fail(CG_ERANGE, "cg_mlp_decode: at most 8 elements");
  if (ng == 0) return CG_OK;
  decode_k<<<nblk(ng, 256), 256, 0, (hipStream_t)stream>>>(Zh, ng, ns, s, aek, ne, tol, W, ld, flag, resid);
  CG_LAUNCH_CHECK("cg_mlp_decode launch");
  return CG_OK;
}

extern "C" int cg_mlp_compact(const int32_t* flag, int64_t ng, int32_t* count, int64_t* idx, void* stream) {
  if (ng == 0) return CG_OK;
  compact_k<<<nblk(ng, 256), 256, 0, (hipStream_t)stream>>>(flag, ng, count, idx);
  CG_LAUNCH_CHECK("cg_mlp_compact launch");
  return CG_OK;
}

extern "C" int cg_mlp_gather(const double* P, const double* T, const double* Y, int64_t ld,
                             const int64_t* idx, const int32_t* count, int64_t cap, int32_t ns,
                             double* Po, double* To, double* Yo, int64_t ldo, void* stream) {
  if (cap == 0) return CG_OK;
  gather_k<<<nblk(cap, 256), 256, 0, (hipStream_t)stream>>>(P, T, Y, ld, idx, count, ns, Po, To, Yo, ldo);
  CG_LAUNCH_CHECK("cg_mlp_gather launch");
  return CG_OK;
}

extern "C" int cg_mlp_scatter(const double* Wi, int64_t ldi, const int64_t* idx, const int32_t* count,
                              int64_t cap, int32_t ns, double* W, int64_t ld, void* stream) {
  if (cap == 0) return CG_OK;
  scatter_k<<<nblk(cap, 256), 256, 0, (hipStream_t)stream>>>(Wi, ldi, idx, count, ns, W, ld);
  CG_LAUNCH_CHECK("cg_mlp_scatter launch");
  return CG_OK;
}
