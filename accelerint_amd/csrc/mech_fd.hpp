// Finite-difference Jacobian over any mechanism's dydt: the reference's
// generic/fd_jacob.c:24-127 with FD_ORD 1 (its default), for callbacks that
// come without an analytic Jacobian.  Per column j, with ewt_i = atol +
// rtol |y_i| (the tolerances themselves, as the reference uses them):
//   fac = sqrt(sum_i (ewt_i f_i)^2 / N),  r0 = 1000 rtol eps N fac,
//   r   = max(sqrt(eps) |y_j|, r0 / ewt_j),
//   J(:, j) = (f(y + r e_j) - f(y)) / r.
// One dydt per column plus one at y: N + 1 evaluations per Jacobian.  Each
// lane writes the entries of the rows it owns, into the same place the base
// mechanism's analytic Jacobian goes (registers, or the group's LDS rows).
#pragma once

#include <float.h>

#include "radau_core.hpp"

namespace acc {

template <class Base>
struct MechFD {
  static constexpr int N = Base::N;
  static constexpr bool kJacInLds = Base::kJacInLds;
  struct Params {
    typename Base::Params base;
    double atol, rtol;
  };
  static size_t lds_bytes(const Params& p, int groups, bool scratch) {
    return Base::lds_bytes(p.base, groups, scratch);
  }
  __device__ __forceinline__ static void block_init(const Params& p, double* lds) { Base::block_init(p.base, lds); }
  struct Ctx : Base::Ctx {
    double atol, rtol;
    __device__ Ctx(const Params& p, int g, double* lds, bool scratch)
        : Base::Ctx(p.base, g, lds, scratch), atol(p.atol), rtol(p.rtol) {}
    __device__ Ctx(const Params& p, const double* tables, double* group, bool scratch)
        : Base::Ctx(p.base, tables, group, scratch), atol(p.atol), rtol(p.rtol) {}
  };

  template <int G, int R>
  __device__ __forceinline__ static void dydt(Ctx& c, double t, double pr, const double (&y)[R], double (&f)[R]) {
    Base::template dydt<G, R>(c, t, pr, y, f);
  }

  // Column j of the difference quotient, passed to put(j, col) (col[q] = the
  // entry of row ln + G q).
  template <int G, int R, class Put>
  __device__ __forceinline__ static void columns(Ctx& c, double t, double pr, const double (&y)[R], Put&& put) {
    using RW = Rows<N, G, R>;
    const int ln = Grp<G>::lane();
    double f0[R];
    Base::template dydt<G, R>(c, t, pr, y, f0);
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const double ew = c.atol + c.rtol * fabs(y[q]);
      if (RW::valid(ln, q)) s += (ew * f0[q]) * (ew * f0[q]);
    }
    s = Grp<G>::sum(s);
    const double fac = sqrt(s / (double)N);
    const double r0 = 1000.0 * c.rtol * DBL_EPSILON * (double)N * fac;
    const double srur = sqrt(DBL_EPSILON);
#pragma unroll 1
    for (int j = 0; j < N; ++j) {
      double mine[R];
#pragma unroll
      for (int q = 0; q < R; ++q) mine[q] = y[q];
      const double yj = Grp<G>::bcast(sel<R>(mine, j / G), j % G);
      const double ej = c.atol + c.rtol * fabs(yj);
      const double r = fmax(srur * fabs(yj), r0 / ej);
      double yp[R], ft[R], col[R];
#pragma unroll
      for (int q = 0; q < R; ++q) yp[q] = (RW::valid(ln, q) && RW::row(ln, q) == j) ? yj + r : y[q];
      Base::template dydt<G, R>(c, t, pr, yp, ft);
#pragma unroll
      for (int q = 0; q < R; ++q) col[q] = (ft[q] - f0[q]) / r;
      put(j, col);
    }
  }

  // Jacobian parked in the group's LDS rows (jac[r*N + j])
  template <int G, int R>
  __device__ __forceinline__ static void jacob(Ctx& c, double t, double pr, const double (&y)[R]) {
    using RW = Rows<N, G, R>;
    const int ln = Grp<G>::lane();
    columns<G, R>(c, t, pr, y, [&](int j, const double (&col)[R]) {
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (RW::valid(ln, q)) c.jac[RW::row(ln, q) * N + j] = col[q];
    });
    group_lds_sync();
  }

  // Jacobian in registers (J[q][j] = d f_row / d y_j)
  template <int G, int R>
  __device__ __forceinline__ static void jacob(Ctx& c, double t, double pr, const double (&y)[R], double (&J)[R][N]) {
    columns<G, R>(c, t, pr, y, [&](int j, const double (&col)[R]) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
#pragma unroll
        for (int k = 0; k < N; ++k)
          if (k == j) J[q][k] = col[q];
      }
    });
  }
};

}  // namespace acc
