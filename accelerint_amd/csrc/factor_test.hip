// Device LU without integration: the LU routines the integrator runs at each
// order and precision (radau_core.hpp: the row LU lu_real / lu_cplx with
// implicit interchanges; at order 14 in FP64 the column LU clu_real /
// clu_cplx with explicit ones) on caller matrices, so that their pivot
// choice, factors and zero-pivot report can be checked one to one against
// LAPACK dgetrf_/zgetrf_, the routines the reference calls in RK_Decomp
// (radau2a.c:273,277); optionally also the matching triangular solves on one
// right-hand side per matrix, the counterpart of dgetrs_ / zgetrs_
// (radau2a.c:404,417).  Orders 2 (one system per lane), 14 (16-lane
// groups) and 54 (one wave per system).
#include <hip/hip_runtime.h>

#include "radau_core.hpp"

namespace acc {

// One group per matrix.  a_re/a_im: row-major N x N per matrix (a_im unused
// for the real LU).  Row r of the factored matrix is written to the row of
// the LAPACK factors it ends at (its elimination step pos), so lu holds the
// combined unit-L / U factors of P A as dgetrf_ returns them; perm[k] = the
// original row chosen as pivot at step k; info as dgetrf_.  With b_re set,
// x = A^-1 b from the factors (b_im / x_im: imaginary parts, complex LU).
template <int N, int G, int R, class T, bool CPLX>
__global__ __launch_bounds__(64) void k_factor_test(int num, const double* __restrict__ a_re,
                                                    const double* __restrict__ a_im, const double* __restrict__ b_re,
                                                    const double* __restrict__ b_im, double* __restrict__ lu_re,
                                                    double* __restrict__ lu_im, int* __restrict__ perm_out,
                                                    int* __restrict__ info_out, double* __restrict__ x_re,
                                                    double* __restrict__ x_im) {
  using RW = Rows<N, G, R>;
  const int ln = Grp<G>::lane();
  const int c = blockIdx.x * (64 / G) + (int)threadIdx.x / G;
  if (c >= num) return;  // whole groups leave together
  const size_t base = (size_t)c * N * N;
  T ar[R][N], ai[R][N];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const bool v = RW::valid(ln, q);
    const int r = RW::row(ln, q);
#pragma unroll
    for (int j = 0; j < N; ++j) {
      ar[q][j] = v ? (T)a_re[base + (size_t)r * N + j] : T(0);
      ai[q][j] = (CPLX && v) ? (T)a_im[base + (size_t)r * N + j] : T(0);
    }
  }
  int pos[R];
  Perm<N> perm;
#pragma unroll
  for (int i = 0; i < (N + 3) / 4; ++i) perm.w[i] = 0;
  T rdr[R], rdi[R];
  int info = 0;
  if constexpr (CPLX) lu_cplx<N, G, R, T>(ar, ai, pos, perm, rdr, rdi, info);
  else lu_real<N, G, R, T>(ar, pos, perm, rdr, info);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (RW::valid(ln, q) && pos[q] >= 0 && pos[q] < N) {
      const size_t o = base + (size_t)pos[q] * N;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        lu_re[o + j] = (double)ar[q][j];
        if (CPLX) lu_im[o + j] = (double)ai[q][j];
      }
    }
  }
  if (ln == 0) {
    for (int k = 0; k < N; ++k) perm_out[(size_t)c * N + k] = perm.get(k);
    info_out[c] = info;
  }
  if (b_re) {  // uniform over the launch
    T br[R], bi[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const bool v = RW::valid(ln, q);
      const size_t o = (size_t)c * N + RW::row(ln, q);
      br[q] = v ? (T)b_re[o] : T(0);
      bi[q] = (CPLX && v) ? (T)b_im[o] : T(0);
    }
    if constexpr (CPLX) solve_cplx<N, G, R, T>(ar, ai, pos, perm, rdr, rdi, br, bi);
    else solve_real<N, G, R, T>(ar, pos, perm, rdr, br);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (RW::valid(ln, q)) {
        const size_t o = (size_t)c * N + RW::row(ln, q);
        x_re[o] = (double)br[q];
        if (CPLX) x_im[o] = (double)bi[q];
      }
    }
  }
}

// Order 14 in FP64: the column LU (lane j holds column j, rows physically
// interchanged; rmap byte k = the original row now at position k, which is
// also the row chosen at step k), then the factors turned into rows through
// LDS for the row-layout solves, as the integrator does.
template <bool CPLX>
__global__ __launch_bounds__(64) void k_factor_test_col(int num, const double* __restrict__ a_re,
                                                        const double* __restrict__ a_im,
                                                        const double* __restrict__ b_re,
                                                        const double* __restrict__ b_im, double* __restrict__ lu_re,
                                                        double* __restrict__ lu_im, int* __restrict__ perm_out,
                                                        int* __restrict__ info_out, double* __restrict__ x_re,
                                                        double* __restrict__ x_im) {
  constexpr int N = 14, G = 16;
  __shared__ double scratch[64 / G][N * N];
  const int ln = Grp<G>::lane();
  const int g = (int)threadIdx.x / G;
  const int c = blockIdx.x * (64 / G) + g;
  if (c >= num) return;  // whole groups leave together
  const size_t base = (size_t)c * N * N;
  double ar[N], ai[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    ar[i] = ln < N ? a_re[base + (size_t)i * N + ln] : 0.0;
    ai[i] = (CPLX && ln < N) ? a_im[base + (size_t)i * N + ln] : 0.0;
  }
  RowMap<N> rmap;
  double rdr = 0.0, rdi = 0.0;
  int info = 0;
  if constexpr (CPLX) clu_cplx<N, G>(ar, ai, rmap, rdr, rdi, info);
  else clu_real<N, G>(ar, rmap, rdr, info);
  if (ln < N) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      lu_re[base + (size_t)i * N + ln] = ar[i];
      if (CPLX) lu_im[base + (size_t)i * N + ln] = ai[i];
    }
  }
  if (ln == 0) {
    for (int k = 0; k < N; ++k) perm_out[(size_t)c * N + k] = rmap.get(k);
    info_out[c] = info;
  }
  if (b_re) {  // uniform over the launch
    double rr[N], ri[N];
    transpose_cols_to_rows<N, G>(ar, rr, scratch[g]);
    if (CPLX) transpose_cols_to_rows<N, G>(ai, ri, scratch[g]);
    const int src = rmap.get(ln);
    double br = ln < N ? b_re[(size_t)c * N + ln] : 0.0;
    double bi = (CPLX && ln < N) ? b_im[(size_t)c * N + ln] : 0.0;
    if constexpr (CPLX) rsolve_cplx<N, G>(rr, ri, src, rdr, rdi, br, bi);
    else rsolve_real<N, G>(rr, src, rdr, br);
    if (ln < N) {
      x_re[(size_t)c * N + ln] = br;
      if (CPLX) x_im[(size_t)c * N + ln] = bi;
    }
  }
}

template <int N, int G, int R, class T, bool CPLX>
static hipError_t launch_one(int num, const double* are, const double* aim, const double* bre, const double* bim,
                             double* lre, double* lim, int* perm, int* info, double* xre, double* xim,
                             hipStream_t s) {
  constexpr int per_block = 64 / G;
  const int blocks = (num + per_block - 1) / per_block;
  hipLaunchKernelGGL((k_factor_test<N, G, R, T, CPLX>), dim3(blocks), dim3(64), 0, s, num, are, aim, bre, bim, lre,
                     lim, perm, info, xre, xim);
  return hipGetLastError();
}

template <int N, int G, int R>
static hipError_t launch_order(bool fp32, bool cplx, int num, const double* are, const double* aim,
                               const double* bre, const double* bim, double* lre, double* lim, int* perm, int* info,
                               double* xre, double* xim, hipStream_t s) {
  if (fp32) {
    return cplx ? launch_one<N, G, R, float, true>(num, are, aim, bre, bim, lre, lim, perm, info, xre, xim, s)
                : launch_one<N, G, R, float, false>(num, are, aim, bre, bim, lre, lim, perm, info, xre, xim, s);
  }
  return cplx ? launch_one<N, G, R, double, true>(num, are, aim, bre, bim, lre, lim, perm, info, xre, xim, s)
              : launch_one<N, G, R, double, false>(num, are, aim, bre, bim, lre, lim, perm, info, xre, xim, s);
}

// Device buffers of num matrices of order n (2, 14 or 54); bre = nullptr: no
// solve.  Returns hipErrorInvalidValue for any other order.
hipError_t launch_factor_test(int n, bool fp32, bool cplx, int num, const double* are, const double* aim,
                              const double* bre, const double* bim, double* lre, double* lim, int* perm, int* info,
                              double* xre, double* xim, hipStream_t s) {
  switch (n) {
    case 2: return launch_order<2, 1, 2>(fp32, cplx, num, are, aim, bre, bim, lre, lim, perm, info, xre, xim, s);
    case 14:
      if (!fp32) {
        const int blocks = (num + 3) / 4;
        if (cplx)
          hipLaunchKernelGGL(k_factor_test_col<true>, dim3(blocks), dim3(64), 0, s, num, are, aim, bre, bim, lre, lim,
                             perm, info, xre, xim);
        else
          hipLaunchKernelGGL(k_factor_test_col<false>, dim3(blocks), dim3(64), 0, s, num, are, aim, bre, bim, lre,
                             lim, perm, info, xre, xim);
        return hipGetLastError();
      }
      return cplx ? launch_one<14, 16, 1, float, true>(num, are, aim, bre, bim, lre, lim, perm, info, xre, xim, s)
                  : launch_one<14, 16, 1, float, false>(num, are, aim, bre, bim, lre, lim, perm, info, xre, xim, s);
    case 54: return launch_order<54, 64, 1>(fp32, cplx, num, are, aim, bre, bim, lre, lim, perm, info, xre, xim, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace acc
