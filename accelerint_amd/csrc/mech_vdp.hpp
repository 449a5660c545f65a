// van der Pol oscillator in the pyJac callback form:
//   examples/van_der_pol/dydt.c:28-35   dy0 = y1, dy1 = mu (1 - y0^2) y1 - y0
//   examples/van_der_pol/jacob.c:29-41  column-major analytic Jacobian
// pr is mu.  In the lane-group layout each lane evaluates only the rows it
// owns, after gathering the (two-component) state.
#pragma once

#include "radau_core.hpp"

namespace acc {

struct MechVdP {
  static constexpr int N = 2;
  static constexpr bool kJacInLds = false;
  struct Params {};
  static size_t lds_bytes(const Params&, int, bool) { return 0; }
  __device__ __forceinline__ static void block_init(const Params&, double*) {}
  struct Ctx {
    double* jac = nullptr;
    double* scratch = nullptr;
    __device__ Ctx(const Params&, int /*group*/, double* /*lds*/, bool /*scratch*/) {}
    __device__ Ctx(const Params&, const double* /*tables*/, double* /*group*/, bool /*scratch*/) {}
  };

  template <int G, int R>
  __device__ __forceinline__ static void gather(const double (&y)[R], double& y0, double& y1) {
    double a[R];
#pragma unroll
    for (int q = 0; q < R; ++q) a[q] = y[q];
    y0 = Grp<G>::bcast(sel<R>(a, 0 / G), 0 % G);
    y1 = Grp<G>::bcast(sel<R>(a, 1 / G), 1 % G);
  }

  template <int G, int R>
  __device__ __forceinline__ static void dydt(Ctx&, double, double mu, const double (&y)[R],
                                              double (&f)[R]) {
    double y0, y1;
    gather<G, R>(y, y0, y1);
    const int ln = Grp<G>::lane();
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = ln + G * q;
      f[q] = (r == 0) ? y1 : mu * (1 - y0 * y0) * y1 - y0;
    }
  }

  template <int G, int R>
  __device__ __forceinline__ static void jacob(Ctx&, double, double mu, const double (&y)[R],
                                               double (&J)[R][N]) {
    double y0, y1;
    gather<G, R>(y, y0, y1);
    const int ln = Grp<G>::lane();
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = ln + G * q;
      J[q][0] = (r == 0) ? 0.0 : -2 * mu * y0 * y1 - 1;
      J[q][1] = (r == 0) ? 1.0 : mu * (1 - y0 * y0);
    }
  }
};

}  // namespace acc
