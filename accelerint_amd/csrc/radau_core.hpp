// Batched Radau IIA (order 5) for gfx950: one condition per lane group.
//
// Layout: a condition is integrated by a group of G lanes inside one
// wavefront (G divides 64).  Lane l of the group owns rows r = l + G*q,
// q < R, of every row-indexed quantity: the state y, the stage vectors, and
// full rows of the Jacobian and of the two iteration matrices.  Matrix rows
// live in VGPRs with compile-time column indices, so the dense LU and the
// triangular solves are straight-line FMA code; the only cross-lane traffic is
// the pivot search (group argmax), the pivot-row broadcast and the scalar
// reductions of the error norms.  Every scalar of the step-size controller is
// group-uniform, so each group runs its own control flow.
//
// The algorithm follows the reference's CPU integrator radau2a/radau2a.c
// (integrate :520-736) step for step; the LU follows LAPACK dgetf2/zgetf2
// semantics (partial pivoting, first maximum, |Re|+|Im| for complex), done with
// implicit row interchanges: rows never move between lanes, each row records
// the LAPACK position it would occupy so ties break exactly as LAPACK would.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace acc {

// ---- control constants: radau2a.c:37-66 ------------------------------------
constexpr int kMaxNoSteps = 200000;
constexpr int kNewtonMaxit = 8;
constexpr double kRoundoff = 2.220446049250313080847e-16;  // DBL_EPSILON
constexpr double kFacMin = 0.2;
constexpr double kFacMax = 8.0;
constexpr double kFacSafe = 0.9;
constexpr double kFacRej = 0.1;
constexpr double kThetaMin = 0.001;
constexpr double kNewtonTol = 0.03;
constexpr double kQmin = 1.0;
constexpr double kQmax = 1.2;
constexpr int kMaxConsecutiveErrs = 5;
constexpr double kDblMin = 2.2250738585072014e-308;

// ---- Radau IIA(5) tableau, as published by Hairer & Wanner (RADAU5) --------
// Same digits as radau2a.c:101-231; every value is re-derived from closed
// forms by the test suite.
constexpr double rkA00 = 1.968154772236604258683861429918299e-1;
constexpr double rkA01 = -6.55354258501983881085227825696087e-2;
constexpr double rkA02 = 2.377097434822015242040823210718965e-2;
constexpr double rkA10 = 3.944243147390872769974116714584975e-1;
constexpr double rkA11 = 2.920734116652284630205027458970589e-1;
constexpr double rkA12 = -4.154875212599793019818600988496743e-2;
constexpr double rkA20 = 3.764030627004672750500754423692808e-1;
constexpr double rkA21 = 5.124858261884216138388134465196080e-1;
constexpr double rkA22 = 1.111111111111111111111111111111111e-1;
constexpr double rkC0 = 1.550510257216821901802715925294109e-1;
constexpr double rkC1 = 6.449489742783178098197284074705891e-1;
constexpr double rkC2 = 1.0;
constexpr double rkE0 = 0.05;
constexpr double rkE1 = -10.04880939982741556246032950764708 * 0.05;
constexpr double rkE2 = 1.382142733160748895793662840980412 * 0.05;
constexpr double rkE3 = -0.3333333333333333333333333333333333 * 0.05;
constexpr double rkGamma = 3.637834252744495732208418513577775;
constexpr double rkAlpha = 2.681082873627752133895790743211112;
constexpr double rkBeta = 3.050430199247410569426377624787569;
constexpr double rkT00 = 9.443876248897524148749007950641664e-2;
constexpr double rkT01 = -1.412552950209542084279903838077973e-1;
constexpr double rkT02 = -3.00291941051474244918611170890539e-2;
constexpr double rkT10 = 2.502131229653333113765090675125018e-1;
constexpr double rkT11 = 2.041293522937999319959908102983381e-1;
constexpr double rkT12 = 3.829421127572619377954382335998733e-1;
constexpr double rkT20 = 1.0;
constexpr double rkT21 = 1.0;
constexpr double rkT22 = 0.0;
constexpr double rkTiAi00 = 1.520148562492775501049204957366528e+1;
constexpr double rkTiAi01 = 1.192055789400527921212348994770778;
constexpr double rkTiAi02 = 1.903956760517560343018332287285119;
constexpr double rkTiAi10 = -9.669512977505946748632625374449567;
constexpr double rkTiAi11 = -8.724028436822336183071773193986487;
constexpr double rkTiAi12 = 3.096043239482439656981667712714881;
constexpr double rkTiAi20 = -1.409513259499574544876303981551774e+1;
constexpr double rkTiAi21 = 5.895975725255405108079130152868952;
constexpr double rkTiAi22 = -1.441236197545344702389881889085515e-1;
constexpr double rkELO = 4.0;

// Return codes: radau2a/radau2a_props.h:30-39.
enum : int {
  EC_success = 0,
  EC_consecutive_steps = 1,
  EC_max_steps_exceeded = 2,
  EC_h_plus_t_equals_h = 3,
  EC_newton_max_iterations_exceeded = 4,
};

// Per-condition event counters (algorithmic-work accounting); the order
// matches the oracle's.
enum : int {
  CNT_STEP = 0, CNT_ACCEPT, CNT_REJECT, CNT_NEWTON_FAIL, CNT_LU, CNT_LU_FAIL,
  CNT_NEWTON, CNT_F, CNT_JAC, CNT_ES, NCNT
};

struct SolverParams {
  double atol;
  double rtol;
  int const_time_step;
};

// ---- lane-group primitives ---------------------------------------------------
template <int G>
struct Grp {
  static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "group width must be a power of two");
  __device__ __forceinline__ static int lane() { return (int)(__lane_id() & (G - 1)); }

  __device__ __forceinline__ static double bcast(double v, int src) {
    if constexpr (G == 1) {
      return v;
    } else if constexpr (G == 64) {
      // src is wave-uniform: read it through SGPRs.
      int s = __builtin_amdgcn_readfirstlane(src);
      int lo = __builtin_amdgcn_readlane(__double2loint(v), s);
      int hi = __builtin_amdgcn_readlane(__double2hiint(v), s);
      return __hiloint2double(hi, lo);
    } else {
      return __shfl(v, src, G);
    }
  }
  __device__ __forceinline__ static int bcast_i(int v, int src) {
    if constexpr (G == 1) {
      return v;
    } else if constexpr (G == 64) {
      int s = __builtin_amdgcn_readfirstlane(src);
      return __builtin_amdgcn_readlane(v, s);
    } else {
      return __shfl(v, src, G);
    }
  }
  __device__ __forceinline__ static double xor_(double v, int m) { return __shfl_xor(v, m, G); }
  __device__ __forceinline__ static int xor_i(int v, int m) { return __shfl_xor(v, m, G); }

  // Butterfly all-reduce: partners add the same two operands, so every lane
  // of the group ends with a bit-identical result.
  __device__ __forceinline__ static double sum(double v) {
#pragma unroll
    for (int m = 1; m < G; m <<= 1) v += xor_(v, m);
    return v;
  }
};

// Row ownership helpers.
template <int N, int G, int R>
struct Rows {
  static_assert(G * R >= N, "group too small for system size");
  __device__ __forceinline__ static int row(int ln, int q) { return ln + G * q; }
  __device__ __forceinline__ static bool valid(int ln, int q) { return ln + G * q < N; }
};

// Select x[q] for a runtime q without dynamic register indexing.
template <int R>
__device__ __forceinline__ double sel(const double (&x)[R], int q) {
  double v = x[0];
#pragma unroll
  for (int i = 1; i < R; ++i) v = (q == i) ? x[i] : v;
  return v;
}

// Packed pivot-row table: entry k = row that was the pivot at elimination step k.
template <int N>
struct Perm {
  uint32_t w[(N + 3) / 4];
  __device__ __forceinline__ void set(int k, int row) {
    const int s = 8 * (k & 3);
    w[k >> 2] = (w[k >> 2] & ~(0xffu << s)) | ((uint32_t)row << s);
  }
  __device__ __forceinline__ int get(int k) const { return (int)((w[k >> 2] >> (8 * (k & 3))) & 0xffu); }
};

// Smith's complex reciprocal 1/(c + i d).
__device__ __forceinline__ void crecip(double c, double d, double& re, double& im) {
  if (fabs(c) >= fabs(d)) {
    double r = d / c, den = c + d * r;
    re = 1.0 / den;
    im = -r / den;
  } else {
    double r = c / d, den = c * r + d;
    re = r / den;
    im = -1.0 / den;
  }
}

// Argmax over the group of (value, key) with key = (LAPACK position << 8) | row;
// larger value wins, equal values go to the smaller position (first maximum).
template <int G>
__device__ __forceinline__ void group_argmax(double& bv, int& bkey) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) {
    double ov = Grp<G>::xor_(bv, m);
    int ok = Grp<G>::xor_i(bkey, m);
    bool better = (ov > bv) || (ov == bv && ok < bkey);
    bv = better ? ov : bv;
    bkey = better ? ok : bkey;
  }
}

// ---- real LU (dgetf2 semantics, implicit interchanges) ---------------------
// On exit a[q][*] holds row r's L multipliers (columns < pos[q]) and U entries
// (columns >= pos[q]); pos[q] is the elimination step at which row r was the
// pivot; perm.get(k) is the pivot row of step k; rd[q] = 1 / U(r, pos[q]).
// info = first step (1-based) with an exactly zero pivot, 0 otherwise.
template <int N, int G, int R>
__device__ __forceinline__ void lu_real(double (&a)[R][N], int (&pos)[R], Perm<N>& perm,
                                        double (&rd)[R], int& info) {
  const int ln = Grp<G>::lane();
#pragma unroll
  for (int q = 0; q < R; ++q) {
    pos[q] = Rows<N, G, R>::valid(ln, q) ? Rows<N, G, R>::row(ln, q) : -1;
    rd[q] = 0.0;
  }
  info = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double bv = -1.0;
    int bkey = 0x7fffffff;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const bool act = pos[q] >= k;
      const double v = fabs(a[q][k]);
      const int key = (pos[q] << 8) | Rows<N, G, R>::row(ln, q);
      const bool better = act && (v > bv || (v == bv && key < bkey));
      bv = better ? v : bv;
      bkey = better ? key : bkey;
    }
    group_argmax<G>(bv, bkey);
    const int prow = bkey & 0xff;
    const int ppos = bkey >> 8;
    const int plane = prow % G;
    const int pq = prow / G;
    perm.set(k, prow);
    double mine[R];
#pragma unroll
    for (int q = 0; q < R; ++q) mine[q] = a[q][k];
    const double piv = Grp<G>::bcast(sel<R>(mine, pq), plane);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (pos[q] == k) pos[q] = ppos;  // the row LAPACK swaps out of position k
      if (Rows<N, G, R>::row(ln, q) == prow) pos[q] = k;
    }
    if (piv != 0.0) {
      const bool tiny = fabs(piv) < kDblMin;
      const double rp = 1.0 / piv;
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (Rows<N, G, R>::row(ln, q) == prow) rd[q] = rp;
      double u[N];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        if (j > k) {
          double m2[R];
#pragma unroll
          for (int q = 0; q < R; ++q) m2[q] = a[q][j];
          u[j] = Grp<G>::bcast(sel<R>(m2, pq), plane);
        }
      }
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (pos[q] > k) {
          const double l = tiny ? a[q][k] / piv : a[q][k] * rp;
          a[q][k] = l;
#pragma unroll
          for (int j = 0; j < N; ++j)
            if (j > k) a[q][j] = fma(-l, u[j], a[q][j]);
        }
      }
    } else if (info == 0) {
      info = k + 1;
    }
  }
}

// ---- complex LU (zgetf2 semantics) -------------------------------------------
template <int N, int G, int R>
__device__ __forceinline__ void lu_cplx(double (&ar)[R][N], double (&ai)[R][N], int (&pos)[R],
                                        Perm<N>& perm, double (&rdr)[R], double (&rdi)[R],
                                        int& info) {
  const int ln = Grp<G>::lane();
#pragma unroll
  for (int q = 0; q < R; ++q) {
    pos[q] = Rows<N, G, R>::valid(ln, q) ? Rows<N, G, R>::row(ln, q) : -1;
    rdr[q] = 0.0;
    rdi[q] = 0.0;
  }
  info = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double bv = -1.0;
    int bkey = 0x7fffffff;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const bool act = pos[q] >= k;
      const double v = fabs(ar[q][k]) + fabs(ai[q][k]);
      const int key = (pos[q] << 8) | Rows<N, G, R>::row(ln, q);
      const bool better = act && (v > bv || (v == bv && key < bkey));
      bv = better ? v : bv;
      bkey = better ? key : bkey;
    }
    group_argmax<G>(bv, bkey);
    const int prow = bkey & 0xff;
    const int ppos = bkey >> 8;
    const int plane = prow % G;
    const int pq = prow / G;
    perm.set(k, prow);
    double mr[R], mi[R];
#pragma unroll
    for (int q = 0; q < R; ++q) { mr[q] = ar[q][k]; mi[q] = ai[q][k]; }
    const double pr_ = Grp<G>::bcast(sel<R>(mr, pq), plane);
    const double pi_ = Grp<G>::bcast(sel<R>(mi, pq), plane);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (pos[q] == k) pos[q] = ppos;
      if (Rows<N, G, R>::row(ln, q) == prow) pos[q] = k;
    }
    if (pr_ != 0.0 || pi_ != 0.0) {
      double rr, ri;
      crecip(pr_, pi_, rr, ri);
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (Rows<N, G, R>::row(ln, q) == prow) { rdr[q] = rr; rdi[q] = ri; }
      double ur[N], ui[N];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        if (j > k) {
          double m2r[R], m2i[R];
#pragma unroll
          for (int q = 0; q < R; ++q) { m2r[q] = ar[q][j]; m2i[q] = ai[q][j]; }
          ur[j] = Grp<G>::bcast(sel<R>(m2r, pq), plane);
          ui[j] = Grp<G>::bcast(sel<R>(m2i, pq), plane);
        }
      }
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (pos[q] > k) {
          const double xr = ar[q][k], xi = ai[q][k];
          const double lr = xr * rr - xi * ri;
          const double li = xr * ri + xi * rr;
          ar[q][k] = lr;
          ai[q][k] = li;
#pragma unroll
          for (int j = 0; j < N; ++j) {
            if (j > k) {
              ar[q][j] = fma(-lr, ur[j], ar[q][j]);
              ar[q][j] = fma(li, ui[j], ar[q][j]);
              ai[q][j] = fma(-lr, ui[j], ai[q][j]);
              ai[q][j] = fma(-li, ur[j], ai[q][j]);
            }
          }
        }
      }
    } else if (info == 0) {
      info = k + 1;
    }
  }
}

// ---- triangular solves (dgetrs / zgetrs, nrhs = 1) ---------------------------
// b[q] is the right-hand side of row r = row(ln, q); on return x[q] holds
// solution component r.  Forward then backward column-oriented substitution in
// pivot order, as reference BLAS dtrsm does it.
template <int N, int G, int R>
__device__ __forceinline__ void solve_real(const double (&a)[R][N], const int (&pos)[R],
                                           const Perm<N>& perm, const double (&rd)[R],
                                           double (&b)[R]) {
  const int ln = Grp<G>::lane();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int prow = perm.get(k);
    const double yk = Grp<G>::bcast(sel<R>(b, prow / G), prow % G);
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (pos[q] > k) b[q] = fma(-a[q][k], yk, b[q]);
  }
  double x[R];
#pragma unroll
  for (int q = 0; q < R; ++q) x[q] = 0.0;
#pragma unroll
  for (int k = N - 1; k >= 0; --k) {
    const int prow = perm.get(k);
    const int pq = prow / G;
    const double xk = Grp<G>::bcast(sel<R>(b, pq) * sel<R>(rd, pq), prow % G);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (pos[q] >= 0 && pos[q] < k) b[q] = fma(-a[q][k], xk, b[q]);
      if (Rows<N, G, R>::row(ln, q) == k) x[q] = xk;
    }
  }
#pragma unroll
  for (int q = 0; q < R; ++q) b[q] = x[q];
}

template <int N, int G, int R>
__device__ __forceinline__ void solve_cplx(const double (&ar)[R][N], const double (&ai)[R][N],
                                           const int (&pos)[R], const Perm<N>& perm,
                                           const double (&rdr)[R], const double (&rdi)[R],
                                           double (&br)[R], double (&bi)[R]) {
  const int ln = Grp<G>::lane();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int prow = perm.get(k);
    const int pl = prow % G, pq = prow / G;
    const double yr = Grp<G>::bcast(sel<R>(br, pq), pl);
    const double yi = Grp<G>::bcast(sel<R>(bi, pq), pl);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (pos[q] > k) {
        const double lr = ar[q][k], li = ai[q][k];
        br[q] = fma(-lr, yr, br[q]);
        br[q] = fma(li, yi, br[q]);
        bi[q] = fma(-lr, yi, bi[q]);
        bi[q] = fma(-li, yr, bi[q]);
      }
    }
  }
  double xr[R], xi[R];
#pragma unroll
  for (int q = 0; q < R; ++q) { xr[q] = 0.0; xi[q] = 0.0; }
#pragma unroll
  for (int k = N - 1; k >= 0; --k) {
    const int prow = perm.get(k);
    const int pl = prow % G, pq = prow / G;
    const double vr = sel<R>(br, pq), vi = sel<R>(bi, pq);
    const double dr = sel<R>(rdr, pq), di = sel<R>(rdi, pq);
    const double kr = Grp<G>::bcast(vr * dr - vi * di, pl);
    const double ki = Grp<G>::bcast(vr * di + vi * dr, pl);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (pos[q] >= 0 && pos[q] < k) {
        const double ur = ar[q][k], ui = ai[q][k];
        br[q] = fma(-ur, kr, br[q]);
        br[q] = fma(ui, ki, br[q]);
        bi[q] = fma(-ur, ki, bi[q]);
        bi[q] = fma(-ui, kr, bi[q]);
      }
      if (Rows<N, G, R>::row(ln, q) == k) { xr[q] = kr; xi[q] = ki; }
    }
  }
#pragma unroll
  for (int q = 0; q < R; ++q) { br[q] = xr[q]; bi[q] = xi[q]; }
}

// ---- the integrator ----------------------------------------------------------
// Mech provides:
//   static constexpr int N;
//   struct Ctx;  (per-group context: LDS pointers, tables)
//   template <int G, int R> dydt(Ctx&, t, pr, const double (&y)[R], double (&f)[R])
//   template <int G, int R> jacob(Ctx&, t, pr, const double (&y)[R], double (&J)[R][N])
//   static constexpr bool kJacInLds;  double* Ctx::jac  (N*N doubles when true)
// When kJacInLds is set the Jacobian is parked in the group's LDS between
// evaluations (it is only read when the iteration matrices are rebuilt),
// which keeps N*R doubles per lane out of the register file.
template <class Mech, int G, int R>
struct Radau {
  static constexpr int N = Mech::N;
  using RW = Rows<N, G, R>;
  using Ctx = typename Mech::Ctx;

  // RK_ErrorNorm, radau2a.c:446-453: max(sqrt(sum (sc d)^2 / N), 1e-10)
  __device__ __forceinline__ static double err_norm(const double (&sc)[R], const double (&d)[R]) {
    const int ln = Grp<G>::lane();
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (RW::valid(ln, q)) s += sc[q] * sc[q] * d[q] * d[q];
    s = Grp<G>::sum(s);
    return fmax(sqrt(s / (double)N), 1e-10);
  }

  // integrate(), radau2a.c:520-736.  y[q] is the owned part of the state.
  __device__ static int integrate(Ctx& ctx, const SolverParams& P, double t_start, double t_end,
                                  double pr, double (&y)[R], uint32_t (&cnt)[NCNT]) {
    const int ln = Grp<G>::lane();
    const double atol = P.atol, rtol = P.rtol;
    double Hmin = 0.0, Hold = 0.0, Hacc = 0.0, ErrOld = 0.0;
    double H = P.const_time_step ? (t_end - t_start) : fmin(5e-7, t_end - t_start);
    double Hnew;
    double t = t_start;
    bool Reject = false, FirstStep = true, SkipJac = false, SkipLU = false;

    double sc[R], y0[R], F0[R], Z1[R], Z2[R], Z3[R], DZ1[R], DZ2[R], DZ3[R];
    double CONT0[R], CONT1[R], CONT2[R];
    double A[R][N], E1[R][N], E2r[R][N], E2i[R][N];
    int pos1[R], pos2[R];
    double rd1[R], rd2r[R], rd2i[R];
    Perm<N> perm1, perm2;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      sc[q] = 1.0 / (atol + fabs(y[q]) * rtol);  // scale_init, :88-94
      y0[q] = y[q];
      F0[q] = 0.0; Z1[q] = 0.0; Z2[q] = 0.0; Z3[q] = 0.0;
      DZ1[q] = 0.0; DZ2[q] = 0.0; DZ3[q] = 0.0;
      CONT0[q] = 0.0; CONT1[q] = 0.0; CONT2[q] = 0.0;
      rd1[q] = 0.0; rd2r[q] = 0.0; rd2i[q] = 0.0;
      pos1[q] = -1; pos2[q] = -1;
#pragma unroll
      for (int j = 0; j < N; ++j) { A[q][j] = 0.0; E1[q][j] = 0.0; E2r[q][j] = 0.0; E2i[q][j] = 0.0; }
    }
#pragma unroll
    for (int i = 0; i < (N + 3) / 4; ++i) { perm1.w[i] = 0; perm2.w[i] = 0; }
    int info = 0, Nconsecutive = 0, Nsteps = 0;
    double NewtonRate = pow(2.0, 1.25);

#pragma unroll 1
    while (t + kRoundoff < t_end) {
      if (!Reject) {
        Mech::template dydt<G, R>(ctx, t, pr, y, F0);
        cnt[CNT_F]++;
      }
      if (!SkipLU) {
        if (!SkipJac) {
          Mech::template jacob<G, R>(ctx, t, pr, y, A);
          cnt[CNT_JAC]++;
          if constexpr (Mech::kJacInLds) {
#pragma unroll
            for (int q = 0; q < R; ++q)
              if (RW::valid(ln, q)) {
#pragma unroll
                for (int j = 0; j < N; ++j) ctx.jac[RW::row(ln, q) * N + j] = A[q][j];
              }
          }
        } else if constexpr (Mech::kJacInLds) {
#pragma unroll
          for (int q = 0; q < R; ++q) {
#pragma unroll
            for (int j = 0; j < N; ++j) A[q][j] = RW::valid(ln, q) ? ctx.jac[RW::row(ln, q) * N + j] : 0.0;
          }
        }
        // RK_Decomp, :256-278
        const double g1 = rkGamma / H, g2r = rkAlpha / H, g2i = rkBeta / H;
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int r = RW::row(ln, q);
#pragma unroll
          for (int j = 0; j < N; ++j) {
            const bool d = (j == r);
            E1[q][j] = -A[q][j] + (d ? g1 : 0.0);
            E2r[q][j] = -A[q][j] + (d ? g2r : 0.0);
            E2i[q][j] = d ? g2i : 0.0;
          }
        }
        lu_real<N, G, R>(E1, pos1, perm1, rd1, info);
        if (info == 0) lu_cplx<N, G, R>(E2r, E2i, pos2, perm2, rd2r, rd2i, info);
        cnt[CNT_LU]++;
        if (info != 0) {
          cnt[CNT_LU_FAIL]++;
          Nconsecutive += 1;
          if (Nconsecutive >= kMaxConsecutiveErrs) return EC_consecutive_steps;
          H *= 0.5;
          Reject = true;
          SkipJac = true;
          SkipLU = false;
          continue;
        } else {
          Nconsecutive = 0;
        }
      }
      Nsteps += 1;
      cnt[CNT_STEP]++;
      if (Nsteps >= kMaxNoSteps) return EC_max_steps_exceeded;
      if (0.1 * fabs(H) <= fabs(t) * kRoundoff) return EC_h_plus_t_equals_h;
      if (FirstStep) {
#pragma unroll
        for (int q = 0; q < R; ++q) { Z1[q] = 0.0; Z2[q] = 0.0; Z3[q] = 0.0; }
      } else {
        // RK_Interpolate, :302-314 (third stage keeps the reference's x2 factor)
        const double rr = H / Hold;
        const double x1 = 1.0 + rkC0 * rr, x2 = 1.0 + rkC1 * rr, x3 = 1.0 + rkC2 * rr;
#pragma unroll
        for (int q = 0; q < R; ++q) {
          Z1[q] = CONT0[q] + x1 * (CONT1[q] + x1 * CONT2[q]);
          Z2[q] = CONT0[q] + x2 * (CONT1[q] + x2 * CONT2[q]);
          Z3[q] = CONT0[q] + x2 * (CONT1[q] + x3 * CONT2[q]);
        }
      }
      bool NewtonDone = false;
      double NewtonIncrementOld = 0.0;
      double Fac = 0.5;
      int NewtonIter = 0;
      double Theta = 0.0;
      NewtonRate = pow(fmax(NewtonRate, kRoundoff), 0.8);

#pragma unroll 1
      for (; NewtonIter < kNewtonMaxit; NewtonIter++) {
        // RK_PrepareRHS, :353-383
        {
          double Tm[R], F[R];
#pragma unroll
          for (int q = 0; q < R; ++q) { DZ1[q] = Z1[q]; DZ2[q] = Z2[q]; DZ3[q] = Z3[q]; }
#pragma unroll
          for (int q = 0; q < R; ++q) Tm[q] = y[q] + Z1[q];
          Mech::template dydt<G, R>(ctx, t + rkC0 * H, pr, Tm, F);
          {
            const double a1 = -H * rkA00, a2 = -H * rkA10, a3 = -H * rkA20;
#pragma unroll
            for (int q = 0; q < R; ++q) { DZ1[q] += a1 * F[q]; DZ2[q] += a2 * F[q]; DZ3[q] += a3 * F[q]; }
          }
#pragma unroll
          for (int q = 0; q < R; ++q) Tm[q] = y[q] + Z2[q];
          Mech::template dydt<G, R>(ctx, t + rkC1 * H, pr, Tm, F);
          {
            const double a1 = -H * rkA01, a2 = -H * rkA11, a3 = -H * rkA21;
#pragma unroll
            for (int q = 0; q < R; ++q) { DZ1[q] += a1 * F[q]; DZ2[q] += a2 * F[q]; DZ3[q] += a3 * F[q]; }
          }
#pragma unroll
          for (int q = 0; q < R; ++q) Tm[q] = y[q] + Z3[q];
          Mech::template dydt<G, R>(ctx, t + rkC2 * H, pr, Tm, F);
          {
            const double a1 = -H * rkA02, a2 = -H * rkA12, a3 = -H * rkA22;
#pragma unroll
            for (int q = 0; q < R; ++q) { DZ1[q] += a1 * F[q]; DZ2[q] += a2 * F[q]; DZ3[q] += a3 * F[q]; }
          }
          cnt[CNT_F] += 3;
        }
        // RK_Solve, :388-441
        {
#pragma unroll
          for (int q = 0; q < R; ++q) {
            const double x1 = DZ1[q] / H, x2 = DZ2[q] / H, x3 = DZ3[q] / H;
            DZ1[q] = rkTiAi00 * x1 + rkTiAi01 * x2 + rkTiAi02 * x3;
            DZ2[q] = rkTiAi10 * x1 + rkTiAi11 * x2 + rkTiAi12 * x3;
            DZ3[q] = rkTiAi20 * x1 + rkTiAi21 * x2 + rkTiAi22 * x3;
          }
          solve_real<N, G, R>(E1, pos1, perm1, rd1, DZ1);
          solve_cplx<N, G, R>(E2r, E2i, pos2, perm2, rd2r, rd2i, DZ2, DZ3);
#pragma unroll
          for (int q = 0; q < R; ++q) {
            const double x1 = DZ1[q], x2 = DZ2[q], x3 = DZ3[q];
            DZ1[q] = rkT00 * x1 + rkT01 * x2 + rkT02 * x3;
            DZ2[q] = rkT10 * x1 + rkT11 * x2 + rkT12 * x3;
            DZ3[q] = rkT20 * x1 + rkT21 * x2 + rkT22 * x3;
          }
          cnt[CNT_NEWTON]++;
        }
        const double d1 = err_norm(sc, DZ1);
        const double d2 = err_norm(sc, DZ2);
        const double d3 = err_norm(sc, DZ3);
        const double NewtonIncrement = sqrt((d1 * d1 + d2 * d2 + d3 * d3) / 3.0);
        Theta = kThetaMin;
        if (NewtonIter > 0) {
          Theta = NewtonIncrement / NewtonIncrementOld;
          if (Theta < 0.99) {
            NewtonRate = Theta / (1.0 - Theta);
          } else {
            break;
          }
          const double NewtonPredictedErr =
              (NewtonIncrement * pow(Theta, (double)(kNewtonMaxit - NewtonIter - 1))) / (1.0 - Theta);
          if (NewtonPredictedErr >= kNewtonTol) {
            const double Qnewton = fmin(10.0, NewtonPredictedErr / kNewtonTol);
            Fac = 0.8 * pow(Qnewton, -1.0 / ((double)(kNewtonMaxit - NewtonIter)));
            break;
          }
        }
        NewtonIncrementOld = fmax(NewtonIncrement, kRoundoff);
#pragma unroll
        for (int q = 0; q < R; ++q) { Z1[q] -= DZ1[q]; Z2[q] -= DZ2[q]; Z3[q] -= DZ3[q]; }
        NewtonDone = (NewtonRate * NewtonIncrement <= kNewtonTol);
        if (NewtonDone) break;
        if (NewtonIter == kNewtonMaxit - 1) return EC_newton_max_iterations_exceeded;
      }
      if (P.const_time_step) {  // :725-733
        t += H;
#pragma unroll
        for (int q = 0; q < R; ++q) y[q] += Z3[q];
        cnt[CNT_ACCEPT]++;
        continue;
      }
      if (!NewtonDone) {
        cnt[CNT_NEWTON_FAIL]++;
        H = Fac * H;
        Reject = true;
        SkipJac = true;
        SkipLU = false;
        continue;
      }
      // RK_ErrorEstimate, :458-508
      double Err;
      {
        const double HrkE1 = rkE1 / H, HrkE2 = rkE2 / H, HrkE3 = rkE3 / H;
        double F2[R], Tm[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
          F2[q] = HrkE1 * Z1[q] + HrkE2 * Z2[q] + HrkE3 * Z3[q];
          Tm[q] = rkE0 * F0[q] + F2[q];
        }
        solve_real<N, G, R>(E1, pos1, perm1, rd1, Tm);
        cnt[CNT_ES]++;
        Err = err_norm(sc, Tm);
        if (Err >= 1.0 && (FirstStep || Reject)) {
          double F1[R];
#pragma unroll
          for (int q = 0; q < R; ++q) Tm[q] += y[q];
          Mech::template dydt<G, R>(ctx, t, pr, Tm, F1);
          cnt[CNT_F]++;
#pragma unroll
          for (int q = 0; q < R; ++q) Tm[q] = F1[q] + F2[q];
          solve_real<N, G, R>(E1, pos1, perm1, rd1, Tm);
          cnt[CNT_ES]++;
          Err = err_norm(sc, Tm);
        }
      }
      Fac = pow(Err, (-1.0 / rkELO)) * (1.0 + 2 * kNewtonMaxit) / (NewtonIter + 1.0 + 2 * kNewtonMaxit);
      Fac = fmin(kFacMax, fmax(kFacMin, Fac));
      Hnew = Fac * H;
      if (Err < 1.0) {
        cnt[CNT_ACCEPT]++;
        if (!FirstStep) {  // Gustafsson, :676-685
          double FacGus = kFacSafe * (H / Hacc) * pow(Err * Err / ErrOld, -0.25);
          FacGus = fmin(kFacMax, fmax(kFacMin, FacGus));
          Fac = fmin(Fac, FacGus);
          Hnew = Fac * H;
        }
        Hacc = H;
        ErrOld = fmax(1e-2, Err);
        FirstStep = false;
        Hold = H;
        t += H;
#pragma unroll
        for (int q = 0; q < R; ++q) y[q] += Z3[q];
        // RK_Make_Interpolate, :283-297
        {
          constexpr double den = (rkC2 - rkC1) * (rkC1 - rkC0) * (rkC0 - rkC2);
#pragma unroll
          for (int q = 0; q < R; ++q) {
            const double z1 = Z1[q], z2 = Z2[q], z3 = Z3[q];
            CONT0[q] = ((-rkC2 * rkC2 * rkC1 * z1 + z3 * rkC1 * rkC0 * rkC0 + rkC1 * rkC1 * rkC2 * z1 -
                         rkC1 * rkC1 * rkC0 * z3 + rkC2 * rkC2 * rkC0 * z2 - z2 * rkC2 * rkC0 * rkC0) /
                        den) - z3;
            CONT1[q] = -(rkC0 * rkC0 * (z3 - z2) + rkC1 * rkC1 * (z1 - z3) + rkC2 * rkC2 * (z2 - z1)) / den;
            CONT2[q] = (rkC0 * (z3 - z2) + rkC1 * (z1 - z3) + rkC2 * (z2 - z1)) / den;
          }
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {  // scale(), :74-81
          sc[q] = 1.0 / (atol + fmax(fabs(y[q]), fabs(y0[q])) * rtol);
          y0[q] = y[q];
        }
        Hnew = fmin(fmax(Hnew, Hmin), t_end - t);
        if (Reject) Hnew = fmin(Hnew, H);
        Reject = false;
        if (t + Hnew / kQmin - t_end >= 0.0) {
          H = t_end - t;
        } else {
          const double Hratio = Hnew / H;
          SkipLU = (Theta <= kThetaMin) && (Hratio >= kQmin) && (Hratio <= kQmax);
          if (!SkipLU) H = Hnew;
        }
        SkipJac = NewtonIter == 1 || NewtonRate <= kThetaMin;
      } else {
        cnt[CNT_REJECT]++;
        if (FirstStep || Reject) {
          H = kFacRej * H;
        } else {
          H = Hnew;
        }
        Reject = true;
        SkipJac = true;
        SkipLU = false;
      }
    }
    return EC_success;
  }
};

}  // namespace acc
