// getrates in the GEMM formulation: the stoichiometric sums of every reaction as dense FP64
// matrix products on the gfx950 matrix cores (rocBLAS DGEMM), the per-point and per-reaction
// element-wise work in two kernels of this file.  It restates the reference's parallel level 4
// (chemgen/templates/reactions_hip/hip_v4_cpp.j2:1-19: rocblas_dgemm of logEG and logC with the
// dense [nr x Nsk] coefficient matrices, an element-wise net-rate kernel, rocblas_dgemm of the
// rates with the [Ns x nr] production matrix) for any mechanism the generator reads: third-body,
// Lindemann, Troe, SRI and PLOG rates, negative prefactors, any stoichiometric coefficient.
//
// Per slab of at most `slab` grid points (column-major in BLAS terms: a row of our species-
// major arrays is a column of length ld):
//   X  [3 Ns][slab]: log2 C_k, s_k = log2 EG_k, C_k              (cg_gemm_species_k)
//   aux[4][slab]:    log2 T, 1/T, log2 pfac, log2 P
//   SG [2 Nr][slab] = [nu_r; nu_p] (2 Nr x Ns) . log2 C           (DGEMM)
//   DS [Nr][slab]   = (nu_p - nu_r) (Nr x Ns) . s                  (DGEMM)
//   M  [Ntb][slab]  = eff (Ntb x Ns) . C    (efficiency 1 for unlisted species)   (DGEMM)
//   Q  [Nr][slab]:   q = kf (2^SG_r - 2^(SG_p - DS - dnu log2 pfac)) [x M]   (cg_gemm_rates_k)
//   wdot [Ns][ld]    = (nu_p - nu_r)^T (Ns x Nr) . Q                (DGEMM, into the caller's array)
// The forward-minus-reverse difference is taken in FP64 from the two exponentials, as the
// reference's level 4 does (hip_v4_cpp.j2:9-16, method 1), so near equilibrium the path is
// less accurate than the generated kernels' factored form; it exists for mechanisms too large
// for those (the reference's motivation for level 4) and is measured against them.
#include "chemgen_amd.h"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <vector>

extern "C" int cg_internal_fail(int code, const char* msg);

namespace {

constexpr double kR0 = 8.314510e7;     // erg / (mol K)   (chemgen/chemistry.py:204-206)
constexpr double kPatm = 1013250.0;    // dyn / cm^2
constexpr double kLog10_2 = 0.30102999566398119521;
constexpr double kLogClamp = -1100.0;  // log2 of a zero concentration: 2^(-1100) = 0

// One grid point per thread: C_k, its log2, s_k = log2 EG_k (NASA-7 in the base-2 form of the
// gpu_coef blob: 7 numbers per range after the inverse molecular weights), and the per-point
// temperature terms.  Species rows of X are `lx` apart.
__global__ void cg_gemm_species_k(const double* __restrict__ P, const double* __restrict__ T,
                                  const double* __restrict__ Y, int64_t ld, int64_t n, int ns,
                                  const double* __restrict__ cf, int off_smh, const double* __restrict__ tmid,
                                  double* __restrict__ X, double* __restrict__ aux, int64_t lx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double t = T[i], p = P[i];
  const double lg2T = log2(t), invT = 1.0 / t;
  const double t2 = t * t, t3 = t2 * t, t4 = t2 * t2;
  double csum = 0.0;
  for (int k = 0; k < ns; ++k) csum += Y[k * ld + i] * cf[k];
  const double scale = p / (kR0 * csum * t);
  for (int k = 0; k < ns; ++k) {
    const double c = Y[k * ld + i] * cf[k] * scale;
    const double* a = cf + off_smh + 14 * k + (t < tmid[k] ? 0 : 7);
    double s = a[6] * invT;
    s = fma(a[1], lg2T, s);
    s = fma(a[2], t, s);
    s = fma(a[3], t2, s);
    s = fma(a[4], t3, s);
    s = fma(a[5], t4, s);
    s = s + a[0];
    X[k * lx + i] = c > 0.0 ? log2(c) : kLogClamp;
    X[(ns + k) * lx + i] = s;
    X[(2 * ns + k) * lx + i] = c;
  }
  aux[i] = lg2T;
  aux[lx + i] = invT;
  aux[2 * lx + i] = log2(kPatm / kR0) - lg2T;
  aux[3 * lx + i] = log2(p);
}

// log2 k at one PLOG pressure point (blob record at q.x: log2 p, p, 1 / (log2 p - previous),
// then q.y rates of (sign, log2 |A|, beta, Ea'), whose rates are summed)
__device__ double plog_point(const double* __restrict__ cf, int2 q, double lg2T, double invT) {
  const double* c = cf + q.x + 3;
  if (q.y == 1) return fma(c[2], lg2T, c[1] - c[3] * invT);
  double s = 0.0;
  for (int r = 0; r < q.y; ++r, c += 4) s += c[0] * exp2(fma(c[2], lg2T, c[1] - c[3] * invT));
  return log2(s);
}

// Record of reaction j (int32 x 8): kind (0 elementary, 1 third-body, 2 Lindemann, 3 Troe,
// 4 Troe with T2, 5 SRI, 6 SRI with T^e, 7 PLOG), blob offset of its rate, row of M (-1: none),
// reversible, negative prefactor, first PLOG point in `plog`, PLOG points, unused.
// One thread per (point, reaction): blockIdx.y = reaction, so a block reads one record.
__global__ void cg_gemm_rates_k(const double* __restrict__ cf, const int32_t* __restrict__ rec,
                                const double* __restrict__ dnu, const int2* __restrict__ plog,
                                const double* __restrict__ aux, const double* __restrict__ SG,
                                const double* __restrict__ DS, const double* __restrict__ M, const double* __restrict__ P,
                                const double* __restrict__ T, int64_t n, int nr, double* __restrict__ Q, int64_t lx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const int32_t* r = rec + 8 * j;
  const int kind = r[0], tb = r[2], rev = r[3], neg = r[4];
  const double* c = cf + r[1];
  const double lg2T = aux[i], invT = aux[lx + i], lg2pfac = aux[2 * lx + i];
  const double t = T[i];
  double kf;
  if (kind == 7) {
    const double p = P[i], lgP = aux[3 * lx + i];
    const int o = r[5], np = r[6];
    int lo = -1;
    for (int q = 0; q < np; ++q) lo = p >= cf[plog[o + q].x + 1] ? q : lo;
    const int2 a = plog[o + (lo < 0 ? 0 : lo)];
    double lk = plog_point(cf, a, lg2T, invT);
    if (lo >= 0 && lo + 1 < np) {
      const int2 b = plog[o + lo + 1];
      lk = fma(plog_point(cf, b, lg2T, invT) - lk, (lgP - cf[a.x]) * cf[b.x + 2], lk);
    }
    kf = exp2(lk);
  } else {
    const double lki = fma(c[1], lg2T, c[0] - c[2] * invT);
    kf = exp2(lki);
    if (kind >= 2) {
      // falloff: Pr = k0 M / k_inf, kf = k_inf Pr / (1 + Pr) F
      const double lk0 = fma(c[4], lg2T, c[3] - c[5] * invT);
      const double pr = M[tb * lx + i] * exp2(lk0 - lki);
      double F = 1.0;
      if (kind == 3 || kind == 4) {
        double fcent = c[6] * exp2(-t * c[7]) + c[8] * exp2(-t * c[9]);
        if (kind == 4) fcent += exp2(-c[10] * invT);
        const double lf = log2(fcent) * kLog10_2;
        const double lp = log2(pr) * kLog10_2;
        const double cc = -0.4 - 0.67 * lf, nn = 0.75 - 1.27 * lf;
        const double f1 = (lp + cc) / (nn - 0.14 * (lp + cc));
        F = exp2(log2(fcent) / (1.0 + f1 * f1));
      } else if (kind >= 5) {
        const double lp = log2(pr) * kLog10_2;
        const double Xs = 1.0 / (1.0 + lp * lp);
        const double base = c[6] * exp2(-c[7] * invT) + exp2(-t * c[8]);
        F = c[9] * exp2(Xs * log2(base));
        if (kind == 6) F *= exp2(c[10] * lg2T);
      }
      kf *= pr / (1.0 + pr) * F;
    }
  }
  if (neg) kf = -kf;
  const double fwd = exp2(SG[j * lx + i]);
  const double bwd = rev ? exp2(SG[(nr + j) * lx + i] - DS[j * lx + i] - dnu[j] * lg2pfac) : 0.0;
  double q = kf * (fwd - bwd);
  if (kind == 1) q *= M[tb * lx + i];
  Q[j * lx + i] = q;
}

// wdot rows [ns][n] (leading dimension ld) = 0: a mechanism without reactions
__global__ void cg_gemm_zero_k(double* __restrict__ W, int64_t ld, int64_t n, int ns) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < ns; ++k) W[k * ld + i] = 0.0;
}

}  // namespace

struct cg_gemm {
  int ns = 0, nr = 0, ntb = 0, off_smh = 0;
  int64_t slab = 0;
  double* coef_d = nullptr;   // gpu_coef blob
  double* tmid_d = nullptr;   // NASA-7 switch temperatures [ns]
  double* nurp_d = nullptr;   // [nu_r; nu_p]: 2 nr x ns, row-major = column-major ns x 2 nr
  double* nud_d = nullptr;    // nu_p - nu_r: nr x ns (row-major)
  double* nut_d = nullptr;    // (nu_p - nu_r)^T: ns x nr (row-major)
  double* eff_d = nullptr;    // ntb x ns
  int32_t* rec_d = nullptr;   // nr x 8
  double* dnu_d = nullptr;    // nr
  int2* plog_d = nullptr;
  double* ws = nullptr;       // X, aux, SG, DS, M, Q for one slab
  rocblas_handle blas = nullptr;
};

namespace {

int gemm_fail(cg_gemm* g, int code, const char* msg) {
  if (g) cg_gemm_free(g);
  return cg_internal_fail(code, msg);
}

template <typename T>
bool upload(T** dst, const T* src, size_t count) {
  if (count == 0) return true;
  if (hipMalloc((void**)dst, count * sizeof(T)) != hipSuccess) return false;
  return hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

// C (m x n, column-major, ldc) = A (m x k, lda) . B (k x n, ldb): the points are the m rows
rocblas_status dgemm(rocblas_handle h, int64_t m, int n, int k, const double* A, int64_t lda, const double* B,
                     int ldb, double* C, int64_t ldc) {
  const double one = 1.0, zero = 0.0;
  return rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)m, n, k, &one, A,
                       (rocblas_int)lda, B, ldb, &zero, C, (rocblas_int)ldc);
}

}  // namespace

extern "C" int cg_gemm_create(const double* coef, int64_t n_coef, int32_t off_smh, const double* tmid, int32_t ns,
                              int32_t nr, int32_t ntb, const double* nu_r, const double* nu_p, const double* eff,
                              const int32_t* rec, const double* dnu, const int32_t* plog, int32_t n_plog,
                              int64_t slab, cg_gemm_t* out) {
  if (!out) return cg_internal_fail(CG_EINVAL, "cg_gemm_create: out is null");
  *out = nullptr;
  if (!coef || n_coef <= 0 || !tmid || ns <= 0 || nr < 0 || ntb < 0 || (nr > 0 && (!nu_r || !nu_p || !rec || !dnu)) ||
      (ntb > 0 && !eff) || n_plog < 0 || (n_plog > 0 && !plog) || slab <= 0 || off_smh < ns ||
      off_smh + 14 * (int64_t)ns > n_coef)
    return cg_internal_fail(CG_EINVAL, "cg_gemm_create: bad argument");
  if (slab > 0x7fffffffLL / 4) return cg_internal_fail(CG_ERANGE, "cg_gemm_create: slab too large");
  cg_gemm* g = new cg_gemm();
  g->ns = ns;
  g->nr = nr;
  g->ntb = ntb;
  g->off_smh = off_smh;
  g->slab = slab;
  for (int j = 0; j < nr; ++j) {
    const int32_t* r = rec + 8 * j;
    const bool bad = r[0] < 0 || r[0] > 7 || r[1] < 0 || r[1] >= n_coef ||
                     ((r[0] >= 1 && r[0] <= 6) ? (r[2] < 0 || r[2] >= ntb) : false) ||
                     (r[0] == 7 && (r[6] < 1 || r[5] < 0 || (int64_t)r[5] + r[6] > n_plog));
    if (bad) return gemm_fail(g, CG_EINVAL, "cg_gemm_create: bad reaction record");
  }
  std::vector<double> nurp((size_t)2 * nr * ns), nud((size_t)nr * ns), nut((size_t)ns * nr);
  for (int j = 0; j < nr; ++j)
    for (int k = 0; k < ns; ++k) {
      const double a = nu_r[(size_t)j * ns + k], b = nu_p[(size_t)j * ns + k];
      nurp[(size_t)j * ns + k] = a;
      nurp[(size_t)(nr + j) * ns + k] = b;
      nud[(size_t)j * ns + k] = b - a;
      nut[(size_t)k * nr + j] = b - a;
    }
  if (!upload(&g->coef_d, coef, (size_t)n_coef) || !upload(&g->tmid_d, tmid, (size_t)ns) ||
      !upload(&g->nurp_d, nurp.data(), nurp.size()) || !upload(&g->nud_d, nud.data(), nud.size()) ||
      !upload(&g->nut_d, nut.data(), nut.size()) || !upload(&g->eff_d, eff, (size_t)ntb * ns) ||
      !upload(&g->rec_d, rec, (size_t)8 * nr) || !upload(&g->dnu_d, dnu, (size_t)nr) ||
      !upload(&g->plog_d, (const int2*)plog, (size_t)n_plog))
    return gemm_fail(g, CG_ENOMEM, "cg_gemm_create: device allocation or copy failed");
  const size_t rows = (size_t)3 * ns + 4 + 2 * (size_t)nr + nr + ntb + nr;
  if (hipMalloc((void**)&g->ws, rows * (size_t)slab * sizeof(double)) != hipSuccess)
    return gemm_fail(g, CG_ENOMEM, "cg_gemm_create: workspace allocation failed");
  if (rocblas_create_handle(&g->blas) != rocblas_status_success)
    return gemm_fail(g, CG_EHIP, "cg_gemm_create: rocblas_create_handle failed");
  rocblas_set_pointer_mode(g->blas, rocblas_pointer_mode_host);
  *out = g;
  return CG_OK;
}

extern "C" int cg_gemm_free(cg_gemm_t g) {
  if (!g) return CG_OK;
  if (g->blas) rocblas_destroy_handle(g->blas);
  for (void* p : {(void*)g->coef_d, (void*)g->tmid_d, (void*)g->nurp_d, (void*)g->nud_d, (void*)g->nut_d,
                  (void*)g->eff_d, (void*)g->rec_d, (void*)g->dnu_d, (void*)g->plog_d, (void*)g->ws})
    if (p) (void)hipFree(p);
  delete g;
  return CG_OK;
}

extern "C" int cg_getrates_gemm(cg_gemm_t g, const double* P, const double* T, int64_t ng, int64_t ld,
                                const double* Y, double* wdot, void* stream) {
  if (!g || ng < 0 || ld < ng || (ng > 0 && (!P || !T || !Y || !wdot)))
    return cg_internal_fail(CG_EINVAL, "cg_getrates_gemm: bad argument");
  if (ld > 0x7fffffffLL) return cg_internal_fail(CG_ERANGE, "cg_getrates_gemm: ld beyond 2^31 - 1");
  if (ng == 0) return CG_OK;
  hipStream_t s = (hipStream_t)stream;
  if (rocblas_set_stream(g->blas, s) != rocblas_status_success)
    return cg_internal_fail(CG_EHIP, "cg_getrates_gemm: rocblas_set_stream failed");
  const int ns = g->ns, nr = g->nr, ntb = g->ntb;
  const int64_t lx = g->slab;
  double* X = g->ws;
  double* aux = X + (size_t)3 * ns * lx;
  double* SG = aux + (size_t)4 * lx;
  double* DS = SG + (size_t)2 * nr * lx;
  double* M = DS + (size_t)nr * lx;
  double* Q = M + (size_t)ntb * lx;
  for (int64_t i0 = 0; i0 < ng; i0 += lx) {
    const int64_t n = ng - i0 < lx ? ng - i0 : lx;
    cg_gemm_species_k<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(P + i0, T + i0, Y + i0, ld, n, ns, g->coef_d,
                                                                  g->off_smh, g->tmid_d, X, aux, lx);
    if (hipGetLastError() != hipSuccess) return cg_internal_fail(CG_EHIP, "cg_getrates_gemm: species launch");
    if (nr == 0) {
      cg_gemm_zero_k<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(wdot + i0, ld, n, ns);
      if (hipGetLastError() != hipSuccess) return cg_internal_fail(CG_EHIP, "cg_getrates_gemm: zero launch");
      continue;
    }
    bool ok = dgemm(g->blas, n, 2 * nr, ns, X, lx, g->nurp_d, ns, SG, lx) == rocblas_status_success &&
              dgemm(g->blas, n, nr, ns, X + (size_t)ns * lx, lx, g->nud_d, ns, DS, lx) == rocblas_status_success &&
              (ntb == 0 ||
               dgemm(g->blas, n, ntb, ns, X + (size_t)2 * ns * lx, lx, g->eff_d, ns, M, lx) == rocblas_status_success);
    if (!ok) return cg_internal_fail(CG_EHIP, "cg_getrates_gemm: rocblas_dgemm (stoichiometric sums) failed");
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)nr);
    cg_gemm_rates_k<<<grid, 256, 0, s>>>(g->coef_d, g->rec_d, g->dnu_d, g->plog_d, aux, SG, DS, M, P + i0, T + i0, n,
                                         nr, Q, lx);
    if (hipGetLastError() != hipSuccess) return cg_internal_fail(CG_EHIP, "cg_getrates_gemm: rates launch");
    if (dgemm(g->blas, n, ns, nr, Q, lx, g->nut_d, nr, wdot + i0, ld) != rocblas_status_success)
      return cg_internal_fail(CG_EHIP, "cg_getrates_gemm: rocblas_dgemm (production rates) failed");
  }
  return CG_OK;
}
